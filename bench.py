#!/usr/bin/env python3
"""bench.py -- chi2 evaluations per second of the batched, served PSO path on one MI355X.

Workload (synthetic data, seeded): the full-size grid (Teff 4000-10000/250 x [M/H]
-1..0.5/0.1 x logg 3-5/0.1 = 8,400 models of 200,000 uint16 pixels over 3800-9000 A,
5,000 lines, 3.36 GB resident in HBM) and 512 stars, each with its own log-uniform
sampling of 40,000-200,000 pixels, labels uniform in the grid, vsini U(0,300),
RV U(-150,150), SNR U(30,300), R = 85,000, a 10-chunk degree-8 continuum; 64 PSO
particles per star.  One step = one PSO iteration of every star (32,768 chi2
evaluations) through the server's hipGraph-captured iteration.

Also measured in the same run: the plain one-fit-at-a-time GPU path (one star,
150,000 pixels, 100 particles, 200 iterations, host PSO) and the CPU baseline:
the C oracle (a restatement of the reference algorithm) on the host cores, over a
bounded sample of evaluations of that star.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
R_INSTR = 85_000.0
# PMC figures of k_model (rocprofv3 counter passes of this script's batched leg), written in
# profiles/ as "<key>: <value>" lines under "window: <stars>x<particles> warmup <W> steps <K>"
# (the k_model launches of the timed iterations of that command), so that they are only used
# beside kernel times of the same iterations
PMC_FILES = [os.path.join(ROOT, "profiles", "k_model_batched_512x64_hbm_traffic.md"),
             os.path.join(ROOT, "profiles", "k_model_k_moments_batched_512x64_mfma_valu.md")]


def measured_pmc(stars, particles, warmup, steps):
    want = f"{stars}x{particles} warmup {warmup} steps {steps}"
    out, on = {}, False
    for path in PMC_FILES:
        if not os.path.exists(path):
            continue
        for line in open(path):
            k, _, v = line.strip().partition(":")
            if k == "window":
                on = v.strip() == want
            elif on and k and v.strip():
                try:
                    out[k.strip()] = float(v)
                except ValueError:
                    pass
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_stars(pasta, synth, interp, wlg, spec, n_stars, seed=4000, n_range=(40_000, 200_000)):
    rng = np.random.default_rng(seed)
    stars, meta = [], []
    for s in range(n_stars):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        wl = synth.observed_wavelengths(spec, n, 600)
        lab = [rng.uniform(4000, 10000), rng.uniform(-1.0, 0.5), rng.uniform(3.0, 5.0),
               rng.uniform(0, 300), rng.uniform(-150, 150)]
        snr = rng.uniform(30, 300)
        disp = pasta.FixedResolutionDispersion(wl, R_INSTR, wlg)
        model = interp.produce_model(disp, *lab)
        flux, var = synth.observe(model, wl, snr, seed + 1 + s)
        cont = pasta.ChunkContinuumFitter(wl, 10, 8, 0.2)
        stars.append(pasta.Star(interp, disp, cont, flux, var))
        fidx = (np.log10(wl) - spec.log10_first) / spec.log10_step
        span = int(np.floor(fidx[-1]) - np.floor(fidx[0]) + 2)
        meta.append(dict(n=n, span=span, wl=wl, flux=flux, var=var, label=lab))
    return stars, meta


def settings(pasta, P, iters, vsini_range=(0.0, 300.0), constraints=None):
    st = pasta.api._settings_c(pasta.PSOSettings(P, iters), vsini_range, (-150.0, 150.0),
                               constraints)
    return st


def run_batched(pasta, _lib, interp, stars, meta, P, warmup, steps, timing_every=5,
                min_iters=200, st=None):
    """Every star runs one PSO fit of max(min_iters, warmup + steps + 3) iterations in the
    server.  The timed steps are graph replays, except every timing_every-th, which runs the
    same launches eagerly with HIP events around k_model and the whole iteration."""
    cfg = _lib.ServerConfig(len(stars), P, 1, 0, 0)
    sv = C.c_void_p()
    _lib.check(_lib.lib.pasta_server_create(interp._h, C.byref(cfg), C.byref(sv)), "server")
    iters = max(min_iters, warmup + steps + 3)
    st = st or settings(pasta, P, iters)
    try:
        for k, s in enumerate(stars):
            rid = C.c_uint64()
            _lib.check(_lib.lib.pasta_server_submit(sv, s.release(), C.byref(st), 5000 + k,
                                                    C.byref(rid)), "submit")
        a, w = C.c_int(), C.c_int()
        t_fit = time.perf_counter()
        _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)), "step")
        log(f"[bench] admitted {a.value} stars ({w.value} waiting), first step "
            f"{time.perf_counter() - t_fit:.2f} s")
        for _ in range(warmup):
            _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)), "step")
        tm = (C.c_double * 7)()
        _lib.lib.pasta_server_timing(sv, 1, tm)
        t0 = time.perf_counter()
        for k in range(steps):
            eager = k % timing_every == timing_every - 1 or (k == steps - 1 and tm[4] == 0)
            if eager:
                _lib.check(_lib.lib.pasta_server_set_timing(sv, 1), "timing")
            _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)), "step")
            if eager:
                _lib.check(_lib.lib.pasta_server_set_timing(sv, 0), "timing")
                _lib.lib.pasta_server_timing(sv, 0, tm)
        elapsed = time.perf_counter() - t0
        assert a.value == len(stars), "a star left the batch during the timed steps"
        _lib.lib.pasta_server_timing(sv, 0, tm)
        while a.value > 0:  # the rest of the fits
            _lib.check(_lib.lib.pasta_server_step(sv, C.byref(a), C.byref(w)), "step")
        fit_s = time.perf_counter() - t_fit
        return dict(elapsed=elapsed, model_ms=tm[1] / tm[4], iter_ms=tm[3] / tm[4], fit_s=fit_s,
                    iters=iters, flops=tm[5] / tm[4], bytes=tm[6] / tm[4], timed=int(tm[4]))
    finally:
        _lib.lib.pasta_server_destroy(sv)


def rotation_taps(vsini, log10_step):
    """Taps of the reference's rotation kernel (src/convolve_rv.rs:438-449)."""
    eps, vrot = 0.6, vsini / 299792.0
    a = 4.0 * (eps - 1.0) / (np.pi * eps)
    n = int(np.ceil(2.0 * vrot / (np.log(10.0) * log10_step) * np.sqrt(1.0 + a * a))) + 2
    return n + 1 - n % 2


def run_rotation_sweep(pasta, _lib, synth, interp, wlg, spec, taps, n_stars=128, P=64):
    """k_model time per rotation-kernel size on the direct (matrix-core) and FFT paths:
    n_stars of the batched workload's stars, P particles each, vsini fixed (FixConstraint)
    so that every evaluation's kernel has the bucket's taps (plus the 6 of the composed
    R = 85,000 instrumental kernel); 2 warm-up and 4 timed iterations, each timed with HIP
    events around k_model."""
    rows = []
    for K in taps:
        lo, hi = 0.1, 3000.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            lo, hi = (lo, mid) if rotation_taps(mid, spec.log10_step) >= K else (mid, hi)
        v = hi
        row = {"taps": rotation_taps(v, spec.log10_step), "vsini": round(v, 3)}
        for name, fft_taps in (("direct", 1 << 20), ("fft", 3)):
            interp.set_fft_taps(fft_taps)
            try:
                stars, meta = make_stars(pasta, synth, interp, wlg, spec, n_stars)
                st = settings(pasta, P, 9, (0.0, 3000.0), [pasta.FixConstraint(3, v)])
                rb = run_batched(pasta, _lib, interp, stars, meta, P, 2, 4, timing_every=1,
                                 min_iters=9, st=st)
            finally:
                interp.set_fft_taps(0)
            row[name + "_ms"] = round(rb["model_ms"], 3)
        log(f"[bench] rotation sweep: {row}")
        rows.append(row)
    return rows


def run_plain(pasta, synth, interp, wlg, spec, iters=200, P=100, seed=2000):
    """The plain path on the full grid: one star, 150,000 pixels, 10 chunks of degree 8,
    100 particles."""
    rng = np.random.default_rng(seed)
    wl = synth.observed_wavelengths(spec, 150_000, 600)
    lab = [rng.uniform(4500, 9500), rng.uniform(-0.9, 0.4), rng.uniform(3.2, 4.8),
           rng.uniform(0, 300), rng.uniform(-150, 150)]
    disp = pasta.FixedResolutionDispersion(wl, R_INSTR, wlg)
    model = interp.produce_model(disp, *lab)
    flux, var = synth.observe(model, wl, rng.uniform(50, 200), seed + 1)
    cont = pasta.ChunkContinuumFitter(wl, 10, 8, 0.2)
    warm = interp.get_fitter(disp, cont, pasta.PSOSettings(P, 2), (0.0, 300.0), (-150.0, 150.0))
    warm.fit(interp, flux, var, seed=1)  # warm-up (code objects, allocations)
    fitter = interp.get_fitter(disp, cont, pasta.PSOSettings(P, iters), (0.0, 300.0), (-150.0, 150.0))
    t0 = time.perf_counter()
    res = fitter.fit(interp, flux, var, seed=3000)
    dt = time.perf_counter() - t0
    n_evals = P * (iters + 1) + 1
    return dict(evals_per_s=n_evals / dt, fit_latency_s=dt, label=res.label.as_list(), true=lab,
                chi2=res.chi2, wl=wl, flux=flux, var=var)


def run_binary(pasta, synth, interp, wlg, spec, P=128, reps=6, fit_iters=20, seed=7000):
    """Binary stars on the full grid: a full-size continuum grid (Planck continua, 8400 x
    200000 u16 with factors), one observed binary (N = 150,000, R = 85,000, 10 chunks of
    degree 8); chi2 of P binary label vectors per batch (BinaryCostFunction) timed over
    `reps` batches of new labels, then one PSO fit of P particles x fit_iters iterations."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    cs, cf = synth.make_continuum_grid(spec)
    cinterp = pasta.InMemInterpolator.from_arrays(cs, spec.labels(), wlg, factors=cf)
    del cs
    prep_s = time.perf_counter() - t0
    wl = synth.observed_wavelengths(spec, 150_000, 600)
    disp = pasta.FixedResolutionDispersion(wl, R_INSTR, wlg)
    l1, l2 = [6100.0, -0.2, 4.3, 12.0, 25.0], [5200.0, 0.1, 4.6, 40.0, -45.0]
    m = interp.produce_binary_model_norm(cinterp, disp, l1, l2, 0.6)
    flux, var = synth.observe(m, wl, 120.0, seed + 1)
    cont = pasta.ChunkContinuumFitter(wl, 10, 8, 0.2)
    fitter = interp.get_binary_fitter(disp, cont, pasta.PSOSettings(P, fit_iters), (0.0, 300.0),
                                      (-150.0, 150.0))
    b = fitter._binary(interp, cinterp, flux, var)

    def labels(k):
        a = synth.random_labels(spec, P, seed + 10 * k)
        c = synth.random_labels(spec, P, seed + 10 * k + 1)
        return np.concatenate([a, c, rng.uniform(0.05, 0.95, (P, 1))], axis=1)

    c, st, _ = b.chi2(labels(0))  # warm-up (allocations, code objects)
    bad = int(np.count_nonzero(st))
    t0 = time.perf_counter()
    for k in range(reps):
        c, st, _ = b.chi2(labels(k + 1))
        bad += int(np.count_nonzero(st))
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    res = fitter.fit(interp, cinterp, flux, var, seed=1)
    fit_s = time.perf_counter() - t1
    return dict(evals_per_s=P * reps / dt, ms_per_batch=dt / reps * 1e3, P=P, reps=reps,
                failed=bad, fit_latency_s=fit_s, fit_evals=P * (fit_iters + 1) + 1,
                fit_chi2=res.chi2, continuum_grid_prep_s=prep_s)


def serving_requests(pasta, synth, interp, wlg, spec, n, seed=5000):
    """Serving-config requests (the client's data): lengths lognormal (median 80,000, clipped
    20,000-200,000), labels uniform in the grid, vsini U(0,300), RV U(-150,150), SNR
    U(30,300), continuum modes cycling chunk(10, 8) / linear (30 smooth columns) /
    constant.  Only wavelengths, flux, variance and the mode travel with a request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        npx = int(np.clip(rng.lognormal(np.log(80_000), 0.5), 20_000, 200_000))
        wl = synth.observed_wavelengths(spec, npx, 600)
        lab = [rng.uniform(4000, 10000), rng.uniform(-1.0, 0.5), rng.uniform(3.0, 5.0),
               rng.uniform(0, 300), rng.uniform(-150, 150)]
        disp = pasta.FixedResolutionDispersion(wl, R_INSTR, wlg)
        flux, var = synth.observe(interp.produce_model(disp, *lab), wl, rng.uniform(30, 300),
                                  seed + 1 + k)
        mode = ("chunk", "linear", "constant")[k % 3]
        reqs.append(dict(wl=wl, flux=flux, var=var, mode=mode, n=npx,
                         design=synth.smooth_design(npx, 30, seed + k) if mode == "linear" else None))
    return reqs


def _continuum(pasta, r):
    if r["mode"] == "chunk":
        return pasta.ChunkContinuumFitter(r["wl"], 10, 8, 0.2)
    if r["mode"] == "linear":
        return pasta.LinearModelContinuumFitter(r["design"])
    return pasta.ConstantContinuum()


def prepare_star(pasta, interp, wlg, r):
    """A request's server-side preparation: dispersion (float indices, instrumental taps),
    continuum model (chunk bounds, design rows, Cholesky factors) and the star's upload."""
    disp = pasta.FixedResolutionDispersion(r["wl"], R_INSTR, wlg)
    return pasta.Star(interp, disp, _continuum(pasta, r), r["flux"], r["var"])


def pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def run_serving(pasta, interp, wlg, reqs, rate, n_req, P, iters, iters_per_step, n_plain, seed,
                workers=8):
    """Poisson arrivals at `rate` req/s of the first n_req requests into the server (continuous
    batching), then the first n_plain of the same arrivals through the plain path, one fit at
    a time in arrival order.  Each request is prepared on a pool of host threads from its
    arrival on (the C calls release the GIL, so preparation overlaps the device's work);
    latency = result time - scheduled arrival, preparation included."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    settings = pasta.PSOSettings(P, iters)
    srv = pasta.FitServer(interp, max_stars=1024, max_particles=P, iters_per_step=iters_per_step)
    pool = ThreadPoolExecutor(max_workers=workers)
    try:
        pending, ids, lat, prep, sub, steps, busy = {}, {}, {}, [], 0, 0, 0
        t0 = time.perf_counter()
        while len(lat) < n_req:
            now = time.perf_counter() - t0
            while sub < n_req and arr[sub] <= now:
                pending[sub] = (pool.submit(prepare_star, pasta, interp, wlg, reqs[sub]),
                                time.perf_counter())
                sub += 1
            for k in [k for k, (f, _) in pending.items() if f.done()]:
                f, tq = pending.pop(k)
                prep.append(time.perf_counter() - tq)
                ids[k] = srv.submit(f.result(), settings, (0.0, 300.0), (-150.0, 150.0), seed=k)
            active, waiting = srv.step()
            steps += 1
            busy = max(busy, active)
            now = time.perf_counter() - t0
            for k, rid in ids.items():
                if k not in lat and srv.poll(rid) is not None:
                    lat[k] = now - arr[k]
            if active == 0 and waiting == 0 and not pending and sub < n_req and len(lat) == sub:
                time.sleep(max(0.0, arr[sub] - (time.perf_counter() - t0)))
            elif active == 0 and waiting == 0 and pending:
                time.sleep(0.0005)  # only preparation in flight
        makespan = time.perf_counter() - t0
        stats = srv.stats()
    finally:
        pool.shutdown()
        srv.close()
    evals, busy_us, rebuilds, rebuild_us = stats[0], stats[4], stats[5], stats[6]
    lat_s = [lat[k] for k in range(n_req)]
    # the plain path, same arrivals, one fit at a time: the service time (preparation + fit)
    # of each of the first n_plain requests is measured; FIFO latencies over all n_req
    # arrivals follow from finish_k = max(arrival_k, finish_(k-1)) + service_(k mod n_plain)
    from pasta_amd import _lib
    n_plain = min(n_plain, n_req)
    st = pasta.api._settings_c(settings, (0.0, 300.0), (-150.0, 150.0), None)
    res = _lib.FitResult()
    service = []
    for k in range(n_plain):
        t1 = time.perf_counter()
        star = prepare_star(pasta, interp, wlg, reqs[k])
        _lib.check(_lib.lib.pasta_fit(star._h, C.byref(st), k, None, C.byref(res), None), "fit")
        service.append(time.perf_counter() - t1)
    plain_lat, fin = [], 0.0
    for k in range(n_req):
        fin = max(arr[k], fin) + service[k % n_plain]
        plain_lat.append(fin - arr[k])
    return dict(rate=rate, requests=n_req, p50_s=pct(lat_s, 50), p99_s=pct(lat_s, 99),
                evals_per_s=evals / makespan, makespan_s=makespan, steps=steps, max_active=busy,
                gpu_busy_s=busy_us * 1e-6, gpu_idle_frac=max(0.0, 1.0 - busy_us * 1e-6 / makespan),
                batch_rebuilds=rebuilds, rebuild_host_s=rebuild_us * 1e-6,
                prep_p50_ms=1e3 * pct(prep, 50),
                prep_p99_ms=1e3 * pct(prep, 99), prep_workers=workers,
                plain_service_ms=1e3 * float(np.mean(service)), plain_measured=n_plain,
                plain_p50_s=pct(plain_lat, 50), plain_p99_s=pct(plain_lat, 99))


def cpu_baseline(interp, spec, plain, budget_s):
    """The oracle on the host cores over a bounded sample of the plain star's evaluations."""
    from oracle import oracle as O
    from pasta_amd import synth
    import pasta_amd as pasta

    t0 = time.perf_counter()
    spectra = synth.read_grid(interp)
    og = O.Grid(spec.labels(), spectra, None, spec.log10_first, spec.log10_step)
    od = O.Dispersion(og, O.DISP_FIXED, plain["wl"], resolution=R_INSTR)
    oc = O.Continuum("chunk", len(plain["wl"]), wl=plain["wl"], n_chunks=10, p_order=8, overlap=0.2)
    log(f"[bench] oracle set-up {time.perf_counter() - t0:.1f} s")
    # the host's CPU share for one GPU on the GPU box is 16 threads: the pool is capped there
    affinity = len(os.sched_getaffinity(0))
    threads = min(affinity, 16)
    rng = np.random.default_rng(77)
    pool = 16384
    labels = np.stack([rng.uniform(4200, 9800, pool), rng.uniform(-0.95, 0.45, pool),
                       rng.uniform(3.1, 4.9, pool), rng.uniform(0, 300, pool),
                       rng.uniform(-150, 150, pool)], axis=1)
    t0 = time.perf_counter()
    first = O.cost_batch(og, od, oc, plain["flux"], plain["var"], labels[:threads], threads)
    t_first = time.perf_counter() - t0
    n = int(max(threads, min(pool - threads, threads * round(budget_s / max(t_first, 1e-3)))))
    t0 = time.perf_counter()
    rest = O.cost_batch(og, od, oc, plain["flux"], plain["var"], labels[threads:threads + n], threads)
    dt = time.perf_counter() - t0
    # one core: a sample of a fifth of the budget
    n1 = int(max(1, min(64, round(0.2 * budget_s * threads / max(t_first, 1e-3)))))
    t0 = time.perf_counter()
    O.cost_batch(og, od, oc, plain["flux"], plain["var"], labels[:n1], 1)
    dt1 = time.perf_counter() - t0
    # full-size parity on the same sample: GPU chi2 vs the oracle's
    disp = pasta.FixedResolutionDispersion(plain["wl"], R_INSTR, pasta.WlGrid(
        spec.log10_first, spec.log10_step, spec.n_pix, log=True))
    star = pasta.Star(interp, disp, pasta.ChunkContinuumFitter(plain["wl"], 10, 8, 0.2),
                      plain["flux"], plain["var"])
    lab_all = labels[: threads + n]
    gpu, st, _ = star.chi2(lab_all)
    ref = np.concatenate([first, rest])
    ok = st == 0
    rels = np.abs(gpu[ok] - ref[ok]) / np.abs(ref[ok])
    # both against FP64 over the reference's f32-stored values, on a subsample: which side
    # the differences come from
    nq = min(len(lab_all), 8 * threads)
    t0 = time.perf_counter()
    ref64 = O.cost_batch(og, od, oc, plain["flux"], plain["var"], lab_all[:nq], threads, f64="q")
    okq = ok[:nq]
    e_gpu = np.abs(gpu[:nq][okq] - ref64[okq]) / ref64[okq]
    e_ref = np.abs(ref[:nq][okq] - ref64[okq]) / ref64[okq]
    log(f"[bench] FP64 oracle on {nq} evaluations in {time.perf_counter() - t0:.1f} s")
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=n / dt, cores=threads, n=n, seconds=dt, parity_max_rel=float(rels.max()),
                parity_p50_rel=float(np.percentile(rels, 50)),
                parity_p99_rel=float(np.percentile(rels, 99)),
                parity_over_1e5=int((rels > 1e-5).sum()), n_parity=int(ok.sum()),
                one_core=n1 / dt1, n1=n1, nproc=os.cpu_count(), affinity=affinity, cpu_model=model,
                fp64_n=int(okq.sum()), gpu_vs_fp64_max=float(e_gpu.max()),
                gpu_vs_fp64_p50=float(np.percentile(e_gpu, 50)),
                ref_vs_fp64_max=float(e_ref.max()), ref_vs_fp64_p50=float(np.percentile(e_ref, 50)))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed PSO iterations (the fit has max(200, warmup + steps + 3))")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--stars", type=int, default=512)
    ap.add_argument("--particles", type=int, default=64)
    ap.add_argument("--n-range", type=int, nargs=2, default=(40_000, 200_000),
                    metavar=("LO", "HI"), help="observed pixels per batched star, U{LO..HI}")
    ap.add_argument("--plain-iters", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-binary", action="store_true")
    ap.add_argument("--serve-rates", default="50,200,500")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="arrival window per rate (requests = rate x seconds, at most 1000)")
    ap.add_argument("--serve-particles", type=int, default=32)
    ap.add_argument("--serve-iters", type=int, default=100)
    ap.add_argument("--rotation-sweep", default="",
                    help="comma-separated rotation-kernel taps: time k_model on the direct and "
                         "FFT paths per size, print the table and exit")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="also run one plain copy of this many GiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(stars would shard across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import pasta_amd as pasta
    from pasta_amd import _lib, synth

    t_all = time.perf_counter()
    spec = synth.full_grid_spec()
    wlg = pasta.WlGrid(spec.log10_first, spec.log10_step, spec.n_pix, log=True)
    t0 = time.perf_counter()
    interp = pasta.InMemInterpolator.empty(spec.labels(), wlg)
    synth.fill_grid_gpu(interp, spec)
    log(f"[bench] grid {interp.n_models} x {interp.n_pix} filled in HBM in {time.perf_counter() - t0:.1f} s")
    if args.copy_probe_gib > 0:
        synth.copy_probe(int(args.copy_probe_gib * (1 << 30)))

    if args.rotation_sweep:
        taps = [int(x) for x in args.rotation_sweep.split(",") if x]
        rows = run_rotation_sweep(pasta, _lib, synth, interp, wlg, spec, taps)
        print(json.dumps({"rotation_sweep": rows}), flush=True)
        return

    t0 = time.perf_counter()
    stars, meta = make_stars(pasta, synth, interp, wlg, spec, args.stars,
                             n_range=tuple(args.n_range))
    log(f"[bench] {len(stars)} stars set up in {time.perf_counter() - t0:.1f} s "
        f"(N {min(m['n'] for m in meta)}..{max(m['n'] for m in meta)})")

    rb = run_batched(pasta, _lib, interp, stars, meta, args.particles, args.warmup, args.steps)
    elapsed, model_ms, iter_ms, fit_s, fit_iters = (rb["elapsed"], rb["model_ms"], rb["iter_ms"],
                                                    rb["fit_s"], rb["iters"])
    n_evals_step = args.particles * len(stars)
    value = n_evals_step * args.steps / elapsed
    tflops = rb["flops"] / (model_ms * 1e-3) / 1e12
    pmc = measured_pmc(len(stars), args.particles, args.warmup, args.steps)
    traffic = pmc.get("traffic_bytes_per_launch")
    log(f"[bench] batched: {value:,.0f} evals/s, {elapsed / args.steps * 1e3:.1f} ms/step, "
        f"k_model {model_ms:.2f} ms ({tflops:.1f} TFLOP/s FP32 algorithmic, "
        f"{rb['bytes'] / (model_ms * 1e-3) / 1e9:,.0f} GB/s algorithmic bytes; timed over "
        f"{rb['timed']} eager launches), iteration {iter_ms:.2f} ms; {len(stars)} fits of "
        f"{fit_iters} iterations in {fit_s:.2f} s")

    plain = None
    if not args.no_plain:
        plain = run_plain(pasta, synth, interp, wlg, spec, iters=args.plain_iters)
        log(f"[bench] plain: {plain['evals_per_s']:,.0f} evals/s, fit {plain['fit_latency_s']:.2f} s, "
            f"label {np.round(plain['label'], 3).tolist()} true {np.round(plain['true'], 3).tolist()}")

    binary = None
    if not args.no_binary:
        binary = run_binary(pasta, synth, interp, wlg, spec)
        log(f"[bench] binary: {binary['evals_per_s']:,.0f} binary chi2 evals/s "
            f"({binary['ms_per_batch']:.1f} ms per batch of {binary['P']}), fit of "
            f"{binary['fit_evals']} evaluations {binary['fit_latency_s']:.2f} s")

    serving = []
    if not args.no_serving:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        counts = [int(min(1000, max(20, r * args.serve_seconds))) for r in rates]
        t0 = time.perf_counter()
        reqs = serving_requests(pasta, synth, interp, wlg, spec, max(counts))
        log(f"[bench] {len(reqs)} serving requests generated in {time.perf_counter() - t0:.1f} s")
        for k, (r, n) in enumerate(zip(rates, counts)):
            sv = run_serving(pasta, interp, wlg, reqs, r, n, args.serve_particles,
                             args.serve_iters, 2, 40, 6000 + k)
            serving.append(sv)
            log(f"[bench] serving {r:g} req/s: {n} requests, p50 {sv['p50_s']:.3f} s, p99 "
                f"{sv['p99_s']:.3f} s, {sv['evals_per_s']:,.0f} evals/s, max {sv['max_active']} "
                f"stars in flight, GPU idle {sv['gpu_idle_frac']:.2f}, preparation p50 "
                f"{sv['prep_p50_ms']:.1f} ms; plain path one at a time (service "
                f"{sv['plain_service_ms']:.1f} ms): p50 {sv['plain_p50_s']:.3f} s, p99 "
                f"{sv['plain_p99_s']:.3f} s")

    cpu = None
    if not args.no_cpu_baseline and plain is not None:
        cpu = cpu_baseline(interp, spec, plain, args.cpu_seconds)
        log(f"[bench] cpu oracle: {cpu['value']:.1f} evals/s on {cpu['cores']} threads "
            f"({cpu['cpu_model']}, nproc {cpu['nproc']}), {cpu['one_core']:.1f} evals/s on 1 "
            f"({cpu['n1']} evals) ({cpu['n']} evals, {cpu['seconds']:.1f} s); GPU-vs-oracle chi2 "
            f"max rel {cpu['parity_max_rel']:.2e} (p50 {cpu['parity_p50_rel']:.1e}, p99 "
            f"{cpu['parity_p99_rel']:.1e}) over {cpu['n_parity']} full-size evaluations; against "
            f"FP64 ({cpu['fp64_n']}): GPU max {cpu['gpu_vs_fp64_max']:.2e}, f32 oracle max "
            f"{cpu['ref_vs_fp64_max']:.2e}")

    out = {
        # value is the batched throughput; the per-star fit latencies (p50/p99) of the
        # served path and of the plain path one fit at a time are under "serving"
        "metric": "chi2 evals/s (batched PSO) + p50/p99 per-star fit latency vs plain GPU path",
        "value": round(value, 1),
        "unit": "chi2 evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded grid, spectra and noise)",
        "config": {
            "workload": f"batched PSO: {len(stars)} stars x {args.particles} particles, full grid "
                        f"8400 x 200000 u16, timed iterations {args.warmup + 2}-"
                        f"{args.warmup + 1 + args.steps} of a {fit_iters}-iteration fit, "
                        f"N U({args.n_range[0] // 1000}k,{args.n_range[1] // 1000}k), R=85000, "
                        "10 chunks deg 8",
            "stars": len(stars),
            "particles": args.particles,
            "evals_per_step": n_evals_step,
            "grid": [25, 16, 21, 200_000],
            "parallelism": "single GPU",
        },
        "roofline": {
            # k_model is bound on chip (its HBM traffic is ~0.1 of peak): its roofline is the
            # FP32 compute one, with the algorithmic flops of the reference's direct count
            # (interpolation 128 W, rotation + instrumental 2 (K_rot + K_instr) W with each
            # evaluation's own K_rot, resample 6 N) over the FP32 peak
            "bound": "mfma",
            "kernel": "k_model",
            "achieved": round(tflops, 2),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "flops_per_launch": rb["flops"],
            "kernel_ms": round(model_ms, 3),
            "timed_launches": rb["timed"],
            "hbm_GBs": round(traffic / (model_ms * 1e-3) / 1e9, 1) if traffic else None,
            "hbm_frac": round(traffic / (model_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            if traffic else None,
            # algorithmic bytes (64 u16 neighbour rows per synthetic pixel + 8 B per observed
            # pixel) over the HBM bytes moved: the reuse the L2 / MALL provide
            "algorithmic_bytes_per_launch": rb["bytes"],
            "reuse_factor": round(rb["bytes"] / traffic, 2) if traffic else None,
            "l2_to_cu_bytes_per_launch": pmc.get("l2_to_cu_bytes_per_launch"),
            "iteration_ms": round(iter_ms, 3),
            # PMC of the same iterations: matrix-core busy fraction and rate of the rotation
            # (f16 x 2) in k_model, VALU busy fraction of k_model
            "pmc": {k: v for k, v in pmc.items() if k not in
                    ("traffic_bytes_per_launch", "l2_to_cu_bytes_per_launch")} or None,
        },
    }
    out["batched_fit_latency_s"] = round(fit_s, 3)  # all stars' fits, submit to last result
    if plain is not None:
        out["plain_gpu"] = {
            "workload": "one star, N=150000, 100 particles x 200 iterations, host PSO",
            "evals_per_s": round(plain["evals_per_s"], 1),
            "fit_latency_s": round(plain["fit_latency_s"], 3),
            "batched_over_plain": round(value / plain["evals_per_s"], 2),
        }
    if cpu is not None:
        out["cpu_baseline"] = {
            "value": round(cpu["value"], 2),
            "unit": "chi2 evals/s",
            "cores": cpu["cores"],
            "kind": "port",
            "sample": f"{cpu['n']} chi2 evaluations of the N=150000 star on the full grid "
                      f"(C oracle, {cpu['cores']} threads: the host's CPU share for one GPU)",
            "one_core_value": round(cpu["one_core"], 2),
            "one_core_sample": f"{cpu['n1']} evaluations on 1 thread",
            "nproc": cpu["nproc"],
            "affinity_cpus": cpu["affinity"],
            "cpu_model": cpu["cpu_model"],
        }
        out["parity_full_size"] = {
            "max_rel_chi2": cpu["parity_max_rel"], "p50_rel_chi2": cpu["parity_p50_rel"],
            "p99_rel_chi2": cpu["parity_p99_rel"], "over_1e-5": cpu["parity_over_1e5"],
            "n": cpu["n_parity"],
            # against FP64 over the reference's f32-stored values (weights, taps, resample
            # weight): the GPU's error and the f32 restatement's own, on the first fp64_n
            "fp64_n": cpu["fp64_n"],
            "gpu_vs_fp64_max_rel": cpu["gpu_vs_fp64_max"], "gpu_vs_fp64_p50_rel": cpu["gpu_vs_fp64_p50"],
            "f32_oracle_vs_fp64_max_rel": cpu["ref_vs_fp64_max"],
            "f32_oracle_vs_fp64_p50_rel": cpu["ref_vs_fp64_p50"]}
    if binary is not None:
        out["binary"] = {
            "workload": f"binary stars (BinaryCostFunction): full grid + full continuum grid "
                        f"(8400 x 200000 u16 each), N=150000, R=85000, 10 chunks deg 8, "
                        f"{binary['P']} label vectors per batch x {binary['reps']} batches; fit "
                        f"{binary['P']} particles x 20 iterations",
            **{k: (round(v, 4) if isinstance(v, float) else v) for k, v in binary.items()},
        }
    if serving:
        out["serving"] = {
            "workload": f"Poisson arrivals, lognormal N (median 80k, 20k-200k), continuum modes "
                        f"chunk(10,8)/linear(30 columns)/constant, PSO {args.serve_particles} "
                        f"particles x {args.serve_iters} iterations per request; latency from "
                        f"scheduled arrival to result, including each request's preparation "
                        f"(dispersion, continuum factorisation, upload) on a pool of host "
                        f"threads from its arrival on; gpu_idle_frac = 1 - device-busy time "
                        f"(HIP events around first evaluations and iterations) / makespan",
            "rates": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in sv.items()}
                      for sv in serving],
        }
    log(f"[bench] total {time.perf_counter() - t_all:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
