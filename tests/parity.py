"""Parity criteria shared by the GPU tests.

FP64 parity: a GPU value g of a halo sum is accepted when
    |g - x| <= 1e-12 * S + E
where x is the oracle's extended-precision (80-bit) value of the same formulas, S the sum over
halos of |per-halo term| (the conditioning of the sum) and E the oracle's estimate of the FP64
rounding error of the closed form itself, which only matters next to the removable 0/0 point of
the PIEMD terms (there the reference's own FP64 result is off the exact value by the same amount).
"""
import numpy as np

REL = 1e-12


def assert_close(name, got, ext, scale, cond, rel=REL):
    got = np.asarray(got)
    tol = rel * np.asarray(scale) + np.asarray(cond)
    err = np.abs(got - ext)
    bad = ~(err <= tol)
    if bad.any():
        k = np.unravel_index(np.argmax(np.where(bad, err / np.maximum(tol, 1e-300), 0)), err.shape)
        raise AssertionError(f"{name}: {bad.sum()} of {err.size} values outside tolerance; worst at {k}: "
                             f"got {got[k]!r} want {ext[k]!r} err {err[k]:.3e} tol {tol[k]:.3e}")
    return float((err / np.maximum(tol, 1e-300)).max())


def hessian_reference(O, lens, grid, istart=0, jstart=0, kmap=False):
    """Extended-precision Hessian planes and their per-pixel tolerances (a, b, c)."""
    ea, eb, ec, ed = O.hessian_planes(lens, grid, istart, jstart, kmap=kmap, extended=True)
    d = O.hessian_planes(lens, grid, istart, jstart, kmap=kmap, with_abs=True)
    sa, sb, sc = d[4], d[5], d[6]
    _, ch = O.cond_planes(lens, grid, istart, jstart)
    if kmap:
        w = float(np.abs(lens.dlsds).max())
        ch = ch * w
    ta, tb, tc = REL * sa + ch, REL * sb + ch, REL * sc + ch
    return (ea, eb, ec), (ta, tb, tc)


def print_case(name, lens, grid):
    """Name the case on stdout before a GPU step, so that a failing run says which one it was."""
    print(f"[case] {name}: {lens.n} halos, {grid.nx}x{grid.ny}", flush=True)
