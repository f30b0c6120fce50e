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
