"""Parity criteria shared by the GPU tests.

FP64 parity: a GPU value g of a halo sum is accepted when
    |g - x| <= 1e-12 * S + E
where x is the oracle's extended-precision (80-bit) value of the same formulas and S the sum over
halos of |per-halo term| (the conditioning of the sum). E is the error of x itself: the oracle's
closed form (the reference's) cancels next to the removable 0/0 point of the PIEMD terms, where
its first-order FP64 rounding estimate `cond` (orc_cond_planes) is large; in 80-bit arithmetic that
error shrinks by the ratio of the unit roundoffs, 2^-64 / 2^-53 = 2^-11.
  * The GPU Hessian uses a form without that cancellation, so against the extended oracle it gets
    E = 2^-10 cond (the oracle's own share, with a factor 2 to spare): EXT_SHARE.
  * The GPU FP64 gradient takes the I_{0.5} ratio in a form without that cancellation as well
    (Q / S, lt_terms64.h grad_piemd): E = EXT_SHARE too.
  * Comparisons with numbers of the reference's FP64 build (golden grids) keep E = cond.
"""
import numpy as np

REL = 1e-12
EXT_SHARE = 2.0 ** -10


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
    ch = ch * EXT_SHARE
    if kmap:
        w = float(np.abs(lens.dlsds).max())
        ch = ch * w
    ta, tb, tc = REL * sa + ch, REL * sb + ch, REL * sc + ch
    return (ea, eb, ec), (ta, tb, tc)


def print_case(name, lens, grid):
    """Name the case on stdout before a GPU step, so that a failing run says which one it was."""
    print(f"[case] {name}: {lens.n} halos, {grid.nx}x{grid.ny}", flush=True)
