"""GPU fits: PSO identity (oracle PSO fed the GPU's chi2 == host-driven plain fit ==
device-driven batched fit, bit for bit), the end-to-end fit against the oracle's
own fit, and the server's continuous batching (each result identical alone and
in the mix)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _fitter(small, pasta, wlg, interp, P=40, iters=100):
    d = pasta.FixedResolutionDispersion(small.wl, 85_000.0, wlg)
    c = pasta.ChunkContinuumFitter(small.wl, 5, 4, 0.2)
    return interp.get_fitter(d, c, pasta.PSOSettings(P, iters), (0.0, 300.0), (-150.0, 150.0)), d, c


def test_pso_identity_oracle_host_device(small, gpu_small):
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=24, iters=30)
    print("plain fit P=24 iters=30", flush=True)
    res = fitter.fit(interp, small.flux, small.var, seed=77)
    star = pasta.Star(interp, d, c, small.flux, small.var)
    b = O.bounds(small.og, (0.0, 300.0), (-150.0, 150.0))
    best, bc, trace, ne = O.pso_run(b, O.pso_settings(24, 30), 77, lambda pos: star.chi2(pos)[0])
    assert ne == 24 * 31
    np.testing.assert_array_equal(trace, res._trace)
    np.testing.assert_array_equal(best, res.label.as_list())
    assert bc == res.chi2
    print("batched fit of the same star", flush=True)
    bulk = fitter.fit_bulk(interp, [small.flux], [small.var], seeds=[77], trace=True)[0]
    np.testing.assert_array_equal(bulk._trace, res._trace)
    assert bulk.label.as_list() == res.label.as_list() and bulk.chi2 == res.chi2
    np.testing.assert_array_equal(bulk.continuum_params, res.continuum_params)


def test_fit_recovers_labels_like_oracle(small, gpu_small):
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp)
    print("plain fit P=40 iters=100 (small config)", flush=True)
    res = fitter.fit(interp, small.flux, small.var, seed=3000)
    b = O.bounds(small.og, (0.0, 300.0), (-150.0, 150.0))
    best, bc, params, ne = O.fit(small.og, small.odisp, small.ocont, small.flux, small.var, b,
                                 O.pso_settings(40, 100), 3000, n_threads=8)
    gpu = np.array(res.label.as_list())
    print("gpu", gpu, res.chi2, "oracle", best, bc, "identical:", np.array_equal(gpu, best), flush=True)
    # both land on the same minimum: within 0.5 % of each label's range and chi2 within 1e-4
    span = np.array([2000.0, 1.5, 1.5, 300.0, 300.0])
    assert np.all(np.abs(gpu - best) <= 5e-3 * span)
    assert abs(res.chi2 - bc) / bc < 1e-4
    assert np.all(np.abs(gpu - small.true) <= np.array([60.0, 0.06, 0.1, 2.0, 0.5]))
    # the GPU's chi2 at the oracle's best matches the oracle's
    star = pasta.Star(interp, d, c, small.flux, small.var)
    chi, st, _ = star.chi2(best[None])
    assert abs(chi[0] - bc) / bc < 1e-5


def _observed(small, lab, seed):
    from pasta_amd import synth
    model = O.produce_model(small.og, small.odisp, lab)
    return synth.observe(model, small.wl, 80.0, seed)


def test_bulk_equals_individual_fits(small, gpu_small):
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=16, iters=25)
    labs = [[5600.0, -0.5, 4.4, 35.0, -40.0], [6500.0, 0.2, 3.8, 120.0, 60.0],
            [6123.0, -0.3, 4.21, 10.0, 20.0]]
    obs = [_observed(small, l, 10 + k) for k, l in enumerate(labs)]
    seeds = [5, 6, 7]
    print("fit_bulk of 3 stars", flush=True)
    bulk = fitter.fit_bulk(interp, [o[0] for o in obs], [o[1] for o in obs], seeds=seeds)
    for k in range(3):
        one = fitter.fit(interp, obs[k][0], obs[k][1], seed=seeds[k])
        assert bulk[k].label.as_list() == one.label.as_list()
        assert bulk[k].chi2 == one.chi2


def test_server_alone_vs_mix(small, gpu_small):
    import ctypes as C
    from pasta_amd import _lib
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=16, iters=20)
    labs = [[5600.0, -0.5, 4.4, 35.0, -40.0], [6500.0, 0.2, 3.8, 120.0, 60.0],
            [6123.0, -0.3, 4.21, 10.0, 20.0], [5200.0, 0.4, 4.9, 250.0, 100.0]]
    obs = [_observed(small, l, 20 + k) for k, l in enumerate(labs)]
    st = pasta.api._settings_c(fitter.settings, (0.0, 300.0), (-150.0, 150.0), None)

    def run(order, submit_at, iters_per_step=3, max_stars=3):
        cfg = _lib.ServerConfig(max_stars, 64, iters_per_step, 0)
        sv = C.c_void_p()
        _lib.check(_lib.lib.pasta_server_create(interp._h, C.byref(cfg), C.byref(sv)))
        ids, out, step = {}, {}, 0
        try:
            while len(out) < len(order):
                for k in order:
                    if k not in ids and submit_at[k] <= step:
                        s = pasta.Star(interp, d, c, obs[k][0], obs[k][1])
                        rid = C.c_uint64()
                        _lib.check(_lib.lib.pasta_server_submit(sv, s.release(), C.byref(st), 100 + k,
                                                                C.byref(rid)))
                        ids[k] = rid.value
                a, w = C.c_int(), C.c_int()
                _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)))
                for k, rid in ids.items():
                    done = C.c_int()
                    r = _lib.FitResult()
                    _lib.check(_lib.lib.pasta_server_poll(sv, rid, C.byref(done), C.byref(r)))
                    if done.value and k not in out:
                        out[k] = (list(r.label), r.chi2, list(r.continuum_params[:r.n_params]))
                step += 1
                assert step < 500
        finally:
            _lib.lib.pasta_server_destroy(sv)
        return out

    print("server: each request alone", flush=True)
    alone = {k: run([k], {k: 0})[k] for k in range(4)}
    print("server: mixed arrivals", flush=True)
    mixed = run([0, 1, 2, 3], {0: 0, 1: 2, 2: 3, 3: 5}, iters_per_step=2, max_stars=2)
    for k in range(4):
        assert mixed[k] == alone[k], k
    # and the server agrees with the plain fit
    one = fitter.fit(interp, obs[2][0], obs[2][1], seed=102)
    assert one.label.as_list() == alone[2][0] and one.chi2 == alone[2][1]


def _server_run(pasta, interp, d, c, obs, st, order, submit_at, cancel_at=None, iters_per_step=2,
                max_stars=3, budget=0):
    """Steps a server, submitting request k at step submit_at[k] and cancelling request k at
    step cancel_at[k]; returns {k: (label, chi2, params)} of the finished requests and the
    request ids."""
    import ctypes as C
    from pasta_amd import _lib
    cfg = _lib.ServerConfig(max_stars, 64, iters_per_step, budget)
    sv = C.c_void_p()
    _lib.check(_lib.lib.pasta_server_create(interp._h, C.byref(cfg), C.byref(sv)))
    ids, out, step, cancelled = {}, {}, 0, set()
    cancel_at = cancel_at or {}
    try:
        while len(out) + len(cancelled) < len(order):
            for k in order:
                if k not in ids and submit_at[k] <= step:
                    s = pasta.Star(interp, d, c, obs[k][0], obs[k][1])
                    rid = C.c_uint64()
                    _lib.check(_lib.lib.pasta_server_submit(sv, s.release(), C.byref(st), 100 + k,
                                                            C.byref(rid)))
                    ids[k] = rid.value
                if k in ids and cancel_at.get(k) == step and k not in out and k not in cancelled:
                    _lib.check(_lib.lib.pasta_server_cancel(sv, ids[k]))
                    cancelled.add(k)
            a, w = C.c_int(), C.c_int()
            _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)))
            for k, rid in ids.items():
                if k in cancelled or k in out:
                    continue
                done = C.c_int()
                r = _lib.FitResult()
                _lib.check(_lib.lib.pasta_server_poll(sv, rid, C.byref(done), C.byref(r)))
                if done.value:
                    out[k] = (list(r.label), r.chi2, list(r.continuum_params[:r.n_params]))
            step += 1
            assert step < 500
        for k in cancelled:  # a cancelled request reports so
            done = C.c_int()
            assert _lib.lib.pasta_server_poll(sv, ids[k], C.byref(done), None) == -6
    finally:
        _lib.lib.pasta_server_destroy(sv)
    return out


def test_server_cancel_running_and_queued(small, gpu_small):
    """A request cancelled while running and one cancelled while queued leave the other
    requests' results exactly as when each runs alone (the batch is rebuilt around them)."""
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=16, iters=20)
    labs = [[5600.0, -0.5, 4.4, 35.0, -40.0], [6500.0, 0.2, 3.8, 120.0, 60.0],
            [6123.0, -0.3, 4.21, 10.0, 20.0], [5200.0, 0.4, 4.9, 250.0, 100.0]]
    obs = [_observed(small, l, 30 + k) for k, l in enumerate(labs)]
    st = pasta.api._settings_c(pasta.PSOSettings(16, 20), (0.0, 300.0), (-150.0, 150.0), None)
    print("server: alone", flush=True)
    alone = {k: _server_run(pasta, interp, d, c, obs, st, [k], {k: 0})[k] for k in (0, 2)}
    print("server: request 1 cancelled while running, 3 while queued", flush=True)
    # 2 iterations per step: request 0 retires in step 9; request 1 (18 iterations done) is
    # cancelled before step 10, right after that retirement, and request 3 while it waits
    # for a slot; request 2 resumes in the rebuilt batch
    mixed = _server_run(pasta, interp, d, c, obs, st, [0, 1, 2, 3], {0: 0, 1: 1, 2: 4, 3: 6},
                        cancel_at={1: 10, 3: 9}, max_stars=3)
    assert set(mixed) == {0, 2}
    assert mixed[0] == alone[0] and mixed[2] == alone[2]


def test_server_memory_flat_over_many_requests(small, gpu_small):
    """Finished requests free their stars: free device memory after a second wave of short
    requests is what it was after the first."""
    from pasta_amd import _lib
    import ctypes as C
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=8, iters=3)
    obs = [_observed(small, [6123.0, -0.3, 4.21, 10.0, 20.0], 50 + k) for k in range(3)]
    srv = pasta.FitServer(interp, max_stars=4, max_particles=8, iters_per_step=3)

    def wave(n):
        ids = [srv.submit(pasta.Star(interp, d, c, *obs[k % 3]), pasta.PSOSettings(8, 3),
                          (0.0, 300.0), (-150.0, 150.0), seed=k) for k in range(n)]
        for _ in range(200):
            a, w = srv.step()
            if a == 0 and w == 0:
                break
        assert all(srv.poll(i) is not None for i in ids)

    def free():
        f, t = C.c_uint64(), C.c_uint64()
        _lib.check(_lib.lib.pasta_device_mem_info(C.byref(f), C.byref(t)))
        return f.value

    try:
        wave(12)
        f0 = free()
        wave(40)
        f1 = free()
        stats = srv.stats()
        assert stats[1] == 52 and stats[3] == 0, stats
        print(f"free device memory after 12 / 52 requests: {f0} / {f1} bytes", flush=True)
        # 40 more stars of ~0.3 MB each would hold ~12 MB if they were not freed
        assert f1 >= f0 - (4 << 20), (f0, f1)
    finally:
        srv.close()


def test_fit_observers_write_reference_files(small, gpu_small, tmp_path):
    """trace_directory writes it_{k}.json per PSO iteration (the swarm after that iteration:
    position and cost of every particle, src/fitting.rs:217-256) and best_particle_file the
    best particle after each iteration (:260-296); the fit itself is unchanged."""
    import json
    pasta, wlg, interp = gpu_small
    fitter, d, c = _fitter(small, pasta, wlg, interp, P=12, iters=6)
    plain = fitter.fit(interp, small.flux, small.var, seed=5)
    res = fitter.fit(interp, small.flux, small.var, seed=5, trace_directory=str(tmp_path))
    assert res.label.as_list() == plain.label.as_list() and res.chi2 == plain.chi2
    files = sorted(tmp_path.glob("it_*.json"), key=lambda p: int(p.stem[3:]))
    assert [p.name for p in files] == [f"it_{k}.json" for k in range(6)]
    star = pasta.Star(interp, d, c, small.flux, small.var)
    for k, p in enumerate(files):
        pop = json.load(open(p))
        assert len(pop) == 12
        pos = np.array([q["position"] for q in pop])
        cost = np.array([q["cost"] for q in pop])
        chi, st, _ = star.chi2(pos)
        np.testing.assert_array_equal(chi, cost)  # the costs of those positions
        # every global best is at most every personal best, so the new global best is the
        # lower of the old one and this swarm's lowest cost
        assert plain._trace[k + 1, 0] == min(plain._trace[k, 0], cost.min())
    bp = tmp_path / "best.json"
    fitter.fit(interp, small.flux, small.var, seed=5, best_particle_file=str(bp))
    best = json.load(open(bp))
    assert len(best) == 6
    np.testing.assert_array_equal([b["cost"] for b in best], plain._trace[1:, 0])
    np.testing.assert_array_equal([b["position"] for b in best], plain._trace[1:, 1:])
