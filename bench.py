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
absolute error against the FP64 map (det A, kappa, gamma1, gamma2) measured after the timed region.
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


def percentile(v, q):
    return float(np.percentile(np.asarray(v), q))


def run_maps(model, outs, grid, r, precision, steps, warmup, stream):
    ptrs = [o.ptr for o in outs]
    for _ in range(warmup):
        model.map(lthip.MAP_FIELDS, 0, ptrs, grid, r=r, precision=precision, stream=stream)
    lthip.sync(stream)
    ev = [lthip.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(steps):
        model.map(lthip.MAP_FIELDS, 0, ptrs, grid, r=r, precision=precision, stream=stream)
        ev[k + 1].record(stream)
    ev[-1].synchronize()
    wall = time.perf_counter() - t0
    per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(steps)]
    total_ms = ev[0].elapsed_ms(ev[-1])
    return per, total_ms, wall


def read_traffic():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), committed as a
    summary next to the kernel traces; None when absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(TRAFFIC_FILE).read())
    return float(m.group(1)) if m else None


def cpu_baseline(lens, grid, budget_s=6.0):
    """The reference's own CPU Hessian grid (oracle/_ref, gradient2_grid_CPU compiled from the
    reference sources) when it was built, else the oracle's restatement; on whole rows of the same
    4096-wide grid with all halos, doubling the rows until one run takes >= budget_s."""
    import ctypes as C
    import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    R = O.ref_lib()
    kind = "reference" if R is not None else "port"
    if R is not None:
        R.ref_set_threads(threads)
    else:
        O.set_threads(threads)
    rows = 2
    while True:
        sub = M.Grid(grid.frame, grid.nx, rows, step=(grid.dx, grid.dy))
        out = (M.LtMatrix * sub.npix)()
        t0 = time.perf_counter()
        if R is not None:
            R.ref_gradient2_grid(out, C.byref(sub.frame), C.byref(lens.soa()), lens.n, sub.dx, sub.dy, sub.nx,
                                 sub.ny, 0, 2048)
        else:
            O.lib().orc_gradient2_grid(out, C.byref(sub.frame), C.byref(lens.soa()), lens.n, sub.dx, sub.dy,
                                       sub.nx, sub.ny, 0, 2048)
        dt = time.perf_counter() - t0
        if dt >= budget_s or rows >= grid.ny:
            break
        rows = min(grid.ny, rows * max(2, int(budget_s / max(dt, 1e-3) * 0.6)))
    return {"value": lens.n * sub.npix / dt, "unit": "halo·px/s", "cores": threads, "kind": kind,
            "sample": f"Hessian grid of all {lens.n} halos on {rows} full rows (4096 px) of the same "
                      f"grid, rows 2048.., {dt:.1f} s, OMP threads = {threads}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=("fp64", "fp32c"), default="fp64")
    ap.add_argument("--no-secondary", action="store_true", help="do not time the other precision")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run one device copy kernel of this many bytes (hardware-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the lens-map path has no "
              "multi-GPU code in this build", file=sys.stderr)
        sys.exit(2)
    lthip.check(lthip.lib().lt_init(0))

    lens = M.config2_lens(seed=2)
    grid = M.config2_grid(4096)
    cosmo = M.flat_lcdm(70.0, 0.3)
    zl, zs = 0.4, 2.0
    r = lthip.cosmo_dlsds(zl, zs, cosmo)
    model = lthip.Model(lens, lthip.MAP_HESSIAN)
    outs = [lthip.DeviceArray(grid.npix) for _ in range(5)]
    stream = lthip.stream_create()

    prec = lthip.PREC_FP64 if a.precision == "fp64" else lthip.PREC_FP32C
    other = lthip.PREC_FP32C if prec == lthip.PREC_FP64 else lthip.PREC_FP64
    per, total_ms, wall = run_maps(model, outs, grid, r, prec, a.steps, a.warmup, stream)
    halo_px = lens.n * grid.npix
    value = a.steps * halo_px / (total_ms * 1e-3)
    avg_ms = float(np.mean(per))
    flops_per_map = lens.halo_flops("hessian") * grid.npix
    achieved = flops_per_map / (avg_ms * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if prec == lthip.PREC_FP64 else PEAK_FP32_TFLOPS
    traffic = read_traffic() if prec == lthip.PREC_FP64 else None

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
        "wall_s": wall,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "flops_convention": "algorithmic: 262 flop per dPIE Hessian halo*px (div, sqrt = 8)",
        },
    }
    if not a.no_secondary:
        # the other precision on the same steps, then the FP32C error against FP64 (untimed)
        first = [o.download() for o in outs[:4]]
        per2, total2, _ = run_maps(model, outs, grid, r, other, a.steps, a.warmup, stream)
        second = [o.download() for o in outs[:4]]
        f64, f32 = (first, second) if prec == lthip.PREC_FP64 else (second, first)
        errs = {n: float(np.abs(x - y).max()) for n, x, y in zip(("kappa", "gamma1", "gamma2", "detA"), f32, f64)}
        name = "fp32c" if other == lthip.PREC_FP32C else "fp64"
        v2 = a.steps * halo_px / (total2 * 1e-3)
        res[name] = {"value": v2, "unit": "halo·px/s", "ms_per_step": total2 / a.steps,
                     "latency_ms": {"p50": percentile(per2, 50), "p99": percentile(per2, 99)},
                     "fp32c_over_fp64": (v2 / value) if name == "fp32c" else (value / v2),
                     "fp32c_max_abs_err_vs_fp64": errs}
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
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
