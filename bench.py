#!/usr/bin/env python3
"""Benchmark of the lens-map hot path on one MI355X.

A step is one cluster-scale map: the Hessian of 503 dPIE halos (2 cluster + 1 BCG + 500 galaxies)
summed over a 4096 x 4096 pixel grid (300" field) with the kappa, gamma1, gamma2, det A and mu maps
written in the same kernel, the lens model and the five output planes resident in HBM. The value is
halo x pixel evaluations per second over the timed steps; per-step latencies come from HIP events on
the stream the kernel runs on.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--precision fp64|fp32c]

The headline value is the FP64 map (--precision fp64, default); the same steps are then timed in
the other precision and reported under "fp32c" (or "fp64"), with the FP32-compensated map's maximum
absolute error against the FP64 map (det A, kappa, gamma1, gamma2) and its relative error on kappa
and gamma where |det A| >= 1e-3, measured after the timed region; then gradient maps of the same
lens and grid in both precisions ("gradient_fp64", "gradient_fp32c" with its error against FP64).
Prints one JSON line. Runs on one GPU only (--gpus > 1 exits with a message).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "lenstool-hpc_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import ltmodel as M  # noqa: E402
import lthip  # noqa: E402

PEAK_FP64_TFLOPS = 78.6   # MI355X vector FP64
PEAK_FP32_TFLOPS = 157.3  # MI355X vector FP32 (published figure, counts packed FP32)
PEAK_HBM_GBS = 8000.0
METRIC = "halo·pixel evals/s, Hessian maps (FP32-comp vs FP64 GPU); p50/p99 map latency"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hessian_fields_fp64_4096_503halos_traffic.md")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "hessian_fields_4096_503halos_valu_counters.md")
SIMDS = 1024          # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9      # MI355X peak engine clock


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q))


def run_maps(model, outs, grid, r, precision, steps, warmup, stream, px=0, kind=lthip.MAP_FIELDS, mode=0):
    ptrs = [o.ptr for o in outs]
    for _ in range(warmup):
        model.map(kind, mode, ptrs, grid, r=r, precision=precision, stream=stream, px=px)
    lthip.sync(stream)
    ev = [lthip.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(steps):
        model.map(kind, mode, ptrs, grid, r=r, precision=precision, stream=stream, px=px)
        ev[k + 1].record(stream)
    ev[-1].synchronize()
    wall = time.perf_counter() - t0
    per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)]
    total_ms = ev[0].elapsed_ms(ev[-1])
    return per, total_ms, wall


def read_counters():
    """Per dPIE halo x pixel figures of the FP64 Hessian + fields kernel from its committed rocprofv3
    counter summary: hardware FP64 flops (64 lanes x (ADD + MUL + 2 FMA) from SQ_INSTS_VALU_*_F64) and
    VALU issue slots (SQ_ACTIVE_INST_VALU); None when absent."""
    if not os.path.exists(COUNTERS_FILE):
        return None
    txt = open(COUNTERS_FILE).read()
    out = {}
    for key in ("fp64_hw_flop_per_halo_px", "valu_issue_slots_per_halo_px"):
        m = re.search(key + r":\s*([0-9.eE+]+)", txt)
        if not m:
            return None
        out[key] = float(m.group(1))
    return out


def read_traffic():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), committed as a
    summary next to the kernel traces; None when absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(TRAFFIC_FILE).read())
    return float(m.group(1)) if m else None


def cpu_model():
    try:
        return next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown"


def cpu_baseline(lens, grid, run_s=2.0, runs=5):
    """The reference's own CPU Hessian grid (gradient2_grid_CPU compiled from the reference sources in
    oracle/_ref, in its AVX2 and, where the host has AVX-512, its AVX-512 build) on whole rows of the
    same 4096-wide grid with all halos, all host cores through OpenMP. The row count is sized so that
    one run takes about run_s; one warm-up run, then the median of `runs` timed runs, for each build;
    the faster build is the baseline. Falls back to the oracle's restatement when no build exists."""
    import ctypes as C
    import statistics
    import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    builds = O.ref_builds()

    def runner(name):
        if name is None:
            O.set_threads(threads)
            return lambda out, sub: O.lib().orc_gradient2_grid(out, C.byref(sub.frame), C.byref(lens.soa()), lens.n,
                                                               sub.dx, sub.dy, sub.nx, sub.ny, 0, 2048)
        R = O.ref_lib(name)
        R.ref_set_threads(threads)
        return lambda out, sub: R.ref_gradient2_grid(out, C.byref(sub.frame), C.byref(lens.soa()), lens.n, sub.dx,
                                                     sub.dy, sub.nx, sub.ny, 0, 2048)

    # rows for ~run_s with the first build
    first = runner(builds[0] if builds else None)
    rows = 4
    while True:
        sub = M.Grid(grid.frame, grid.nx, rows, step=(grid.dx, grid.dy))
        out = (M.LtMatrix * sub.npix)()
        t0 = time.perf_counter()
        first(out, sub)
        dt = time.perf_counter() - t0
        if dt >= 0.25 * run_s or rows >= 1024:
            break
        rows *= 4
    rows = int(min(grid.ny - 2048, max(4, rows * run_s / max(dt, 1e-3))))
    sub = M.Grid(grid.frame, grid.nx, rows, step=(grid.dx, grid.dy))
    out = (M.LtMatrix * sub.npix)()
    per_build = {}
    for name in (builds or [None]):
        f = runner(name)
        f(out, sub)  # warm-up
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            f(out, sub)
            ts.append(time.perf_counter() - t0)
        per_build[name or "oracle port"] = lens.n * sub.npix / statistics.median(ts)
    best = max(per_build, key=per_build.get)
    flags = O.REF_FLAGS.get(best, "gcc -O3 -march=x86-64-v3 -fopenmp -std=c99 (oracle port)")
    return {"value": per_build[best], "unit": "halo·px/s", "cores": threads,
            "kind": "reference" if builds else "port",
            "sample": f"gradient2_grid_CPU (the reference's CPU Hessian grid) of all {lens.n} halos on {rows} full "
                      f"rows (4096 px) of the same grid from row 2048, median of {runs} runs after one warm-up",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads, "compiler_flags": flags,
            "per_build_halo_px_per_s": per_build}


def bench_bayes(a):
    """Batched Bayes maps: nsamples perturbations of the 1 dPIE + 10 SIS cluster lens (cluster x, y
    by N(0, 0.5"), e, theta, rcore, sigma by 5%, galaxy sigma by a common 5% factor), each a 2048^2
    amplification map (mode 1, mu), all samples in one kernel launch inside a captured graph that
    also uploads the staged parameters. A step = stage a new sample set on the host + replay."""
    S, n = a.samples, 2048
    base = M.group_lens(seed=1)
    sets = [M.perturbed_samples(base, S, seed=100 + k) for k in range(2)]
    # the samples' SOA arrays, as a C caller holds them (a step stages one: the library packs it)
    arrs = [lthip.Model.soa_array(x) for x in sets]
    grid = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), n, n)
    r = lthip.cosmo_dlsds(0.4, 2.0, M.flat_lcdm(70.0, 0.3))
    model = lthip.Model(sets[0], lthip.MAP_HESSIAN)
    out = lthip.DeviceArray(S * grid.npix)
    stream = lthip.stream_create()
    res = {}
    keep = {}
    for name, prec in (("fp64", lthip.PREC_FP64), ("fp32c", lthip.PREC_FP32C)):
        g = lthip.Graph(stream)
        model.upload(stream)
        model.map(lthip.MAP_AMPLI, 1, [out.ptr], grid, r=r, precision=prec, stream=stream)
        g.end()
        for k in range(a.warmup):
            model.stage(arrs[k % 2])
            g.launch()
            lthip.sync(stream)
        ev = [lthip.Event() for _ in range(a.steps + 1)]
        per = []
        t0 = time.perf_counter()
        for k in range(a.steps):
            model.stage(arrs[k % 2])
            ev[k].record(stream)
            g.launch()
            ev[k + 1].record(stream)
            lthip.sync(stream)
            per.append(ev[k].elapsed_ms(ev[k + 1]))
        wall = time.perf_counter() - t0
        work = S * grid.npix * base.n
        res[name] = {"value": a.steps * work / wall, "unit": "halo·px/s", "ms_per_step": wall / a.steps * 1e3,
                     "device_ms_p50": percentile(per, 50), "device_ms_p99": percentile(per, 99),
                     "device_halo_px_per_s": work / (np.mean(per) * 1e-3)}
        # both precisions end on the same sample set: keep four samples' mu for the error check
        keep[name] = [out.download_range(s * grid.npix, grid.npix) for s in (0, S // 3, 2 * S // 3, S - 1)]
        g.close()
    det_err = max(float(np.abs(1 / x - 1 / y).max()) for x, y in zip(keep["fp32c"], keep["fp64"]))
    res["fp32c"]["max_abs_err_detA_vs_fp64"] = det_err
    return {"metric": "halo·pixel evals/s, batched Bayes amplification maps (one graph launch per step)",
            "value": res["fp64"]["value"], "unit": "halo·px/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": res["fp64"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{S} Bayes samples x 2048^2 mu maps, 1 dPIE + 10 SIS", "samples": S,
                       "pixels_per_sample": grid.npix, "halos": base.n},
            "fp64": res["fp64"], "fp32c": res["fp32c"]}


def refine_images(lens, pos, sx, sy, dr, iters=8, h=1e-5):
    """Newton iteration on theta - dr alpha(theta) = s for image positions pos [k, 2] of one source
    (s = (sx, sy)), with the library's FP64 point gradients and a central-difference Jacobian; an
    image whose residual does not shrink keeps its starting position."""
    th = np.array(pos, dtype=np.float64)
    def resid(t):
        gx, gy = lthip.point_gradients(lens, t[:, 0], t[:, 1])
        return np.stack([t[:, 0] - dr * gx - sx, t[:, 1] - dr * gy - sy], axis=1)
    r0 = np.hypot(*resid(th).T)
    for _ in range(iters):
        F = resid(th)
        J = np.empty((len(th), 2, 2))
        for j, e in enumerate(((h, 0.0), (0.0, h))):
            d = (resid(th + np.array(e)) - resid(th - np.array(e))) / (2 * h)
            J[:, :, j] = d
        th = th - np.linalg.solve(J, F[:, :, None])[:, :, 0]
    r1 = np.hypot(*resid(th).T)
    keep = ~(r1 < r0) | ~np.isfinite(r1) | ~(np.hypot(*(th - np.asarray(pos)).T) < 1.0)
    th[keep] = np.asarray(pos)[keep]
    return th


def bench_chi(a):
    """Delensing + chi^2 batched over lens samples: the 1 dPIE + 10 SIS lens with 40 families of 3
    images (sources uniform in [-20, 20]^2 arcsec, z_s ~ U[1, 6]; images = this library's own
    delensing of the true model on a 2048^2 grid, refined by Newton iteration on
    theta - dr alpha(theta) = s, + N(0, 0.1") noise), delensing grid 512^2, samples perturbed as for
    the Bayes maps. A step = one ChiBatch run over all samples. The CPU baseline is the oracle's
    chi^2 path (its restatement of the reference's gradient grid, delensing and chi_computation) on a
    few of the same samples."""
    rng = np.random.default_rng(4)
    lens = M.group_lens(seed=1)
    cosmo = M.flat_lcdm(70.0, 0.3)
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    fine = M.Grid(frame, 2048, 2048)
    gx, gy = lthip.gradient_grid(lens, fine)
    fams = []
    while len(fams) < 40:
        src = M.galaxies_array(16)
        src["cx"] = rng.uniform(-20, 20, 16)
        src["cy"] = rng.uniform(-20, 20, 16)
        zs = rng.uniform(1.0, 6.0, 16)
        src["dr"] = [lthip.cosmo_dlsds(0.4, z, cosmo) for z in zs]
        _, pos, loc, _ = lthip.delense_barycenter(16, fine.nx, frame, src, gx.ravel(), gy.ravel())
        for s in range(16):
            if loc[s] >= 3 and len(fams) < 40:
                th = refine_images(lens, pos[s, :3], src["cx"][s], src["cy"][s], src["dr"][s])
                fams.append((th, src["dr"][s], zs[s]))
    nimages = np.full(40, 3, dtype=np.int32)
    images = M.galaxies_array(120)
    for f, (p, dr, z) in enumerate(fams):
        for k in range(3):
            images[3 * f + k]["cx"] = p[k, 0] + rng.normal(0, 0.1)
            images[3 * f + k]["cy"] = p[k, 1] + rng.normal(0, 0.1)
            images[3 * f + k]["dr"] = dr
            images[3 * f + k]["redshift"] = z
    S = a.samples
    samples = M.perturbed_samples(lens, S, seed=5)
    b = lthip.ChiBatch(S, lens.type, 512, frame, nimages, images)
    for _ in range(a.warmup):
        b.run(samples)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        chi, found, status = b.run(samples)
    wall = time.perf_counter() - t0
    b.close()
    cpu = None
    if not a.no_cpu_baseline:
        import pyoracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        O.set_threads(threads)
        n, t1 = 0, time.perf_counter()
        while n < S and (n < 4 or time.perf_counter() - t1 < 10.0):
            O.chi_barycentersource(samples[n], 512, frame, nimages, images)
            n += 1
        dt = time.perf_counter() - t1
        cpu = {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "port",
               "sample": f"the oracle's chi^2 path on the first {n} of the same samples, {dt:.1f} s, "
                         f"OMP threads = {threads}"}
    return {"metric": "chi^2 samples/s (gradient grid + barycentres + delensing + chi^2 per sample)",
            "value": a.steps * S / wall, "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{S} samples, 40 families x 3 images, delensing grid 512^2, 1 dPIE + 10 SIS",
                       "samples": S},
            "gradient_halo_px_per_s": a.steps * S * 512 * 512 * lens.n / wall,
            "chi_p50": float(np.median(chi)), "found_mean": float(found.mean()),
            "overflow_samples": int((status != 0).sum()), "cpu_baseline": cpu}


def serve_pool(n_models=32, seed=9):
    """Resident models of the serving workload: cluster-like lenses with H = floor(10^U[1, log10 2000])
    halos (3 cluster-scale dPIE + luminosity-scaled galaxies)."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_models):
        H = int(10 ** rng.uniform(1.0, np.log10(2000.0)))
        L = M.cluster_lens(seed=1000 + k, ngal=max(H - 3, 0))
        L.dlsds[:] = 0.7
        L._soa = None
        pool.append(L)
    return pool


def serve_sweep(pool, rates, requests, slots=8, budget_gb=64, seed=9, until_saturation=False):
    """Poisson arrivals at each rate (req/s): grid n = 2^U[8,13] rounded to a multiple of 64; map kind
    uniform over {gradient (dpl), ampli (mu), shear, mass (Kmap), pot}; precision uniform over {FP32C,
    FP64}. A producer thread submits at the arrival times, the consumer
    fetches in submission order. Per rate: offered and achieved request rate, throughput (halo·px/s)
    and p50/p99 latency from submission to result ready on the device (and to result on the host).
    until_saturation: keep doubling the last rate until the achieved rate falls below 90 % of the
    offered one."""
    import threading
    rng = np.random.default_rng(seed)
    r = lthip.cosmo_dlsds(0.4, 2.0, M.flat_lcdm(70.0, 0.3))
    srv = lthip.Server(nslots=slots, hbm_budget_bytes=budget_gb << 30)
    mids = [srv.add_model(L) for L in pool]
    kinds = [(lthip.MAP_DPL, 1), (lthip.MAP_AMPLI, 1), (lthip.MAP_SHEAR, 1), (lthip.MAP_MASS, 1), (lthip.MAP_POT, 1)]

    def make_req():
        n = int(round(2 ** rng.uniform(8, 13) / 64.0)) * 64
        n = min(max(n, 256), 8192)
        k, m = kinds[rng.integers(len(kinds))]
        mi = int(rng.integers(len(pool)))
        grid = M.Grid(M.grid_param(-150.0, 150.0, -150.0, 150.0), n, n)
        p = int(rng.integers(2))
        return mi, k, m, grid, p

    rows = []
    todo = list(rates)
    while todo:
        rate = todo.pop(0)
        reqs = [make_req() for _ in range(requests)]
        gaps = rng.exponential(1.0 / rate, len(reqs))
        srv.reset_stats()
        tickets = [None] * len(reqs)
        ready = threading.Semaphore(0)

        def producer():
            t_next = time.perf_counter()
            for i, (mi, k, m, grid, p) in enumerate(reqs):
                t_next += gaps[i]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(mids[mi], k, m, grid, precision=p, r=r, mult=1.0)
                ready.release()

        th = threading.Thread(target=producer)
        t0 = time.perf_counter()
        th.start()
        lat = np.empty(len(reqs))
        for i in range(len(reqs)):
            ready.acquire()
            _, k, _, grid, _ = reqs[i]
            _, tm = srv.fetch(tickets[i], k, grid.npix)
            lat[i] = tm.done_ms - tm.submit_ms
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        # latency by request size: halos x pixels, the quarter of smallest and the tenth of largest requests
        work = np.array([pool[mi].n * grid.npix for mi, _, _, grid, _ in reqs], dtype=float)
        small, large = work <= np.quantile(work, 0.25), work >= np.quantile(work, 0.9)
        row = {"offered_req_per_s": rate, "achieved_req_per_s": len(reqs) / wall, "requests": len(reqs),
               "wall_s": wall, "halo_px_per_s": st["halo_px"] / wall, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
               "p50_fetch_ms": st["p50_fetch_ms"], "p99_fetch_ms": st["p99_fetch_ms"], "bytes_peak": st["bytes_peak"],
               "smallest_quarter": {"halo_px_max": float(work[small].max()), "p50_ms": float(np.median(lat[small])),
                                    "p99_ms": float(np.quantile(lat[small], 0.99))},
               "largest_tenth": {"halo_px_min": float(work[large].min()), "p50_ms": float(np.median(lat[large])),
                                 "p99_ms": float(np.quantile(lat[large], 0.99))}}
        rows.append(row)
        print(f"[serve] {json.dumps(row)}", file=sys.stderr, flush=True)
        if until_saturation and not todo and row["achieved_req_per_s"] >= 0.9 * rate:
            todo.append(rate * 2)
    srv.close()
    return rows


def bench_serve(a):
    """Map-serving loop under Poisson load, rates swept (--rates; --until-saturation doubles the last
    rate until the achieved rate falls below 90 % of the offered one), --requests per rate."""
    rows = serve_sweep(serve_pool(), a.rates, a.requests, a.slots, a.budget_gb, until_saturation=a.until_saturation)
    best = max(rows, key=lambda x: x["halo_px_per_s"])
    return {"metric": "map-serving throughput (halo·px/s) and p50/p99 latency under Poisson load",
            "value": best["halo_px_per_s"], "unit": "halo·px/s", "n_gpus": 1, "steps": a.requests,
            "warmup": 0, "ms_per_step": best["wall_s"] / best["requests"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64|f32+f64 (per request)", "data": "synthetic",
            "config": {"workload": "Poisson map requests, n in 256..8192, H in 10..2000, 5 map kinds, "
                                   "FP32C or FP64", "slots": a.slots, "hbm_budget_gb": a.budget_gb},
            "latency_ms": {"p50": best["p50_ms"], "p99": best["p99_ms"]}, "sweep": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("cluster", "bayes", "chi", "serve"), default="cluster")
    ap.add_argument("--samples", type=int, default=0, help="bayes: 256, chi: 1024 by default")
    ap.add_argument("--rates", type=float, nargs="*", default=[8.0, 16.0, 32.0, 64.0])
    ap.add_argument("--requests", type=int, default=500)
    ap.add_argument("--until-saturation", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="cluster workload only: skip the short Bayes, chi^2 and serving sub-results")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "fp32c"), default="fp64")
    ap.add_argument("--no-secondary", action="store_true", help="do not time the other precision")
    ap.add_argument("--px", type=int, default=0, help="pixels per thread (0: pick with the autotuner)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run one device copy kernel of this many bytes (hardware-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the lens-map path has no "
              "multi-GPU code in this library", file=sys.stderr)
        sys.exit(2)
    lthip.check(lthip.lib().lt_init(0))
    if a.workload != "cluster":
        if not a.samples:
            a.samples = 256 if a.workload == "bayes" else 1024
        fn = {"bayes": bench_bayes, "chi": bench_chi, "serve": bench_serve}[a.workload]
        print(json.dumps(fn(a)), flush=True)
        return

    lens = M.cluster_lens(seed=2)
    grid = M.cluster_grid(4096)
    cosmo = M.flat_lcdm(70.0, 0.3)
    zl, zs = 0.4, 2.0
    r = lthip.cosmo_dlsds(zl, zs, cosmo)
    model = lthip.Model(lens, lthip.MAP_HESSIAN)
    outs = [lthip.DeviceArray(grid.npix) for _ in range(5)]
    stream = lthip.stream_create()

    prec = lthip.PREC_FP64 if a.precision == "fp64" else lthip.PREC_FP32C
    other = lthip.PREC_FP32C if prec == lthip.PREC_FP64 else lthip.PREC_FP64
    # launch shape (pixels per thread) picked by the library's autotuner for both precisions, untimed
    tuned = {}
    for name, pr in (("fp64", lthip.PREC_FP64), ("fp32c", lthip.PREC_FP32C)):
        if a.px:
            tuned[name] = {"px": a.px, "ms_px1_px2_px4": None}
            continue
        best, ms = model.autotune(lthip.MAP_FIELDS, 0, grid, pr, stream)
        tuned[name] = {"px": best, "ms_px1_px2_px4": [round(x, 3) for x in ms]}
    per, total_ms, wall = run_maps(model, outs, grid, r, prec, a.steps, a.warmup, stream, a.px)
    halo_px = lens.n * grid.npix
    value = a.steps * halo_px / (total_ms * 1e-3)
    avg_ms = float(np.mean(per))
    flops_per_map = lens.halo_flops("hessian") * grid.npix
    algorithmic = flops_per_map / (avg_ms * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if prec == lthip.PREC_FP64 else PEAK_FP32_TFLOPS
    traffic = read_traffic() if prec == lthip.PREC_FP64 else None
    counters = read_counters() if prec == lthip.PREC_FP64 else None
    rate_dev = halo_px / (avg_ms * 1e-3)  # halo x px per second of kernel time
    if counters:
        achieved = counters["fp64_hw_flop_per_halo_px"] * rate_dev / 1e12
        roofline = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": traffic,
                    "flops": "hardware FP64 flops, 64 lanes x (ADD + MUL + 2 FMA) per wavefront instruction, "
                             f"{counters['fp64_hw_flop_per_halo_px']:.2f} per dPIE halo x px (rocprofv3 "
                             "SQ_INSTS_VALU_{ADD,MUL,FMA}_F64, profiles/hessian_fields_4096_503halos_valu_counters.md)",
                    "valu_issue_frac": counters["valu_issue_slots_per_halo_px"] * 4 * rate_dev / (SIMDS * CLOCK_HZ),
                    "valu_issue": f"SQ_ACTIVE_INST_VALU slots ({counters['valu_issue_slots_per_halo_px']:.3f} per "
                                  "dPIE halo x px) x 4 cycles / (1024 SIMDs x 2.4 GHz x time)",
                    "algorithmic": {"achieved": algorithmic, "frac": algorithmic / peak,
                                    "convention": "262 flop per dPIE Hessian halo*px (division, sqrt = 8 each); "
                                                  "this kernel has no division or sqrt, so the convention "
                                                  "overstates its work and can exceed the peak"}}
    else:
        roofline = {"bound": "valu", "achieved": algorithmic, "peak": peak, "unit": "TFLOP/s",
                    "frac": algorithmic / peak, "traffic": traffic,
                    "flops": "algorithmic convention: 262 flop per dPIE Hessian halo*px (division, sqrt = 8 each)"}

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "halo·px/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": total_ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == lthip.PREC_FP64 else "f32+f64",
        "data": "synthetic (seeded lens model and grid; no external data)",
        "config": {
            "workload": "cluster Hessian + kappa/gamma1/gamma2/detA/mu maps, 4096x4096 px over 300 arcsec, "
                        "503 dPIE halos (2 cluster + 1 BCG + 500 luminosity-scaled galaxies), z_l=0.4, z_s=2.0",
            "pixels": grid.npix, "halos": lens.n, "halo_px_per_step": halo_px, "precision": a.precision,
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": percentile(per, 50), "p99": percentile(per, 99), "mean": avg_ms},
        "autotune": tuned,
        "wall_s": wall,
        "roofline": roofline,
    }
    if not a.no_secondary:
        # the other precision on the same steps, then the FP32C error against FP64 (untimed)
        first = [o.download() for o in outs[:4]]
        per2, total2, _ = run_maps(model, outs, grid, r, other, a.steps, a.warmup, stream, a.px)
        second = [o.download() for o in outs[:4]]
        f64, f32 = (first, second) if prec == lthip.PREC_FP64 else (second, first)
        errs = {n: float(np.abs(x - y).max()) for n, x, y in zip(("kappa", "gamma1", "gamma2", "detA"), f32, f64)}
        # relative errors of kappa and gamma away from the critical lines (|det A| >= 1e-3)
        away = np.abs(f64[3]) >= 1e-3
        g64, g32 = np.hypot(f64[1], f64[2]), np.hypot(f32[1], f32[2])
        mk, mg = away & (f64[0] != 0), away & (g64 != 0)
        rel = {"kappa": float((np.abs(f32[0] - f64[0])[mk] / np.abs(f64[0][mk])).max()),
               "gamma": float((np.abs(g32 - g64)[mg] / g64[mg]).max())}
        name = "fp32c" if other == lthip.PREC_FP32C else "fp64"
        v2 = a.steps * halo_px / (total2 * 1e-3)
        res[name] = {"value": v2, "unit": "halo·px/s", "ms_per_step": total2 / a.steps,
                     "latency_ms": {"p50": percentile(per2, 50), "p99": percentile(per2, 99)},
                     "fp32c_over_fp64": (v2 / value) if name == "fp32c" else (value / v2),
                     "fp32c_max_abs_err_vs_fp64": errs,
                     "fp32c_max_rel_err_vs_fp64_where_abs_detA_ge_1e-3": rel}
        # the FP32C map with the lanes-over-halos launch shape (not among the autotuner's shapes: its sums
        # run in another order)
        per3, total3, _ = run_maps(model, outs, grid, r, lthip.PREC_FP32C, a.steps, a.warmup, stream,
                                   lthip.PX_HALO_LANES)
        hl = [o.download() for o in outs[:4]]
        res["fp32c_halo_lanes"] = {
            "value": a.steps * halo_px / (total3 * 1e-3), "unit": "halo·px/s", "ms_per_step": total3 / a.steps,
            "latency_ms": {"p50": percentile(per3, 50), "p99": percentile(per3, 99)},
            "over_fp64": (a.steps * halo_px / (total3 * 1e-3)) / (value if prec == lthip.PREC_FP64 else v2),
            "max_abs_err_vs_fp64": {n: float(np.abs(x - y).max())
                                    for n, x, y in zip(("kappa", "gamma1", "gamma2", "detA"), hl, f64)}}
        # gradient maps of the same lens and grid in both precisions, each autotuned first
        gm = lthip.Model(lens, lthip.MAP_GRADIENT)
        gsteps = min(a.steps, 10)
        grads = {}
        for gname, gprec in (("fp64", lthip.PREC_FP64), ("fp32c", lthip.PREC_FP32C)):
            gm.autotune(lthip.MAP_GRADIENT, 0, grid, gprec, stream)
            perg, totalg, _ = run_maps(gm, outs[:2], grid, 1.0, gprec, gsteps, 1, stream, 0, lthip.MAP_GRADIENT)
            grads[gname] = [o.download() for o in outs[:2]]
            res["gradient_" + gname] = {
                "value": gsteps * halo_px / (totalg * 1e-3), "unit": "halo·px/s", "ms_per_step": totalg / gsteps,
                "steps": gsteps, "latency_ms": {"p50": percentile(perg, 50), "p99": percentile(perg, 99)},
                "algorithmic_TFLOP/s": lens.halo_flops("gradient") * grid.npix / (totalg / gsteps * 1e-3) / 1e12}
        gm.close()
        res["gradient_fp32c"]["max_abs_err_vs_fp64_arcsec"] = float(
            np.hypot(grads["fp32c"][0] - grads["fp64"][0], grads["fp32c"][1] - grads["fp64"][1]).max())
        res["gradient_fp32c"]["fp32c_over_fp64"] = res["gradient_fp32c"]["value"] / res["gradient_fp64"]["value"]
        # potential maps (mode 3, no scaling) of the same lens and grid in both precisions
        pm = lthip.Model(lens, lthip.MAP_POT)
        psteps = min(a.steps, 3)
        pots = {}
        for pname, pprec in (("fp64", lthip.PREC_FP64), ("fp32c", lthip.PREC_FP32C)):
            pm.autotune(lthip.MAP_POT, 3, grid, pprec, stream)
            perp, totalp, _ = run_maps(pm, outs[:1], grid, 1.0, pprec, psteps, 1, stream, 0, lthip.MAP_POT, mode=3)
            pots[pname] = outs[0].download()
            res["potential_" + pname] = {
                "value": psteps * halo_px / (totalp * 1e-3), "unit": "halo·px/s", "ms_per_step": totalp / psteps,
                "steps": psteps, "latency_ms": {"p50": percentile(perp, 50), "p99": percentile(perp, 99)}}
        pm.close()
        dphi = np.abs(pots["fp32c"] - pots["fp64"])
        res["potential_fp32c"]["max_abs_err_vs_fp64_arcsec2"] = float(dphi.max())
        res["potential_fp32c"]["max_rel_err_vs_fp64"] = float((dphi / np.maximum(np.abs(pots["fp64"]), 1e-300)).max())
        res["potential_fp32c"]["fp32c_over_fp64"] = res["potential_fp32c"]["value"] / res["potential_fp64"]["value"]
    if a.copy_bytes > 0:
        src = lthip.DeviceArray(a.copy_bytes, np.uint8)
        dst = lthip.DeviceArray(a.copy_bytes, np.uint8)
        e0, e1 = lthip.Event(), lthip.Event()
        e0.record(stream)
        lthip.check(lthip.lib().lt_memcpy_d2d(dst.ptr, src.ptr, a.copy_bytes, stream))
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1)
        res["copy_calibration"] = {"bytes": a.copy_bytes, "ms": ms, "GB/s": 2 * a.copy_bytes / ms / 1e6}
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(lens, grid)
    if not a.no_workloads:
        # short runs of the other workloads, so that every bench line measures them too
        for o in outs:
            o.close()
        model.close()
        sub = argparse.Namespace(samples=256, steps=5, warmup=1, no_cpu_baseline=True)
        b = bench_bayes(sub)
        res["bayes_maps"] = {"workload": b["config"]["workload"], "steps": sub.steps,
                             "fp64": b["fp64"], "fp32c": b["fp32c"],
                             "fp32c_over_fp64": b["fp32c"]["value"] / b["fp64"]["value"]}
        sub = argparse.Namespace(samples=1024, steps=3, warmup=1, no_cpu_baseline=True)
        c = bench_chi(sub)
        res["chi"] = {"workload": c["config"]["workload"], "steps": sub.steps, "value": c["value"],
                      "unit": c["unit"], "ms_per_step": c["ms_per_step"], "chi_p50": c["chi_p50"],
                      "found_mean": c["found_mean"], "overflow_samples": c["overflow_samples"]}
        rows = serve_sweep(serve_pool(), [20.0, 80.0], 120)
        res["serving"] = {"workload": "Poisson map requests, n in 256..8192, H in 10..2000, 5 map kinds, FP32C or "
                                      "FP64, 8 slots, 64 GB arena (short sweep; the full one is bench.py "
                                      "--workload serve)", "sweep": rows}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
