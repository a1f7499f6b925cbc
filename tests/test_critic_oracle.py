"""The marching-squares restatement of the critical-line rules in the oracle (CPU), pinned by
geometry: the det A = 0 contour of a singular isothermal sphere is the circle r = b0 D_ls/D_s
(det A = 1 - theta_E / r in closed form), every crossing lies on a cell edge within the linear
interpolation's O(h^2) of that circle, the contour closes (every end point is the start of
another segment, bit for bit), it runs clockwise (positive det A, outside, on its left), and the
saddle and kind rules give the hand-derived segments of single cells. No reference counterpart
exists (the reference parses a `cline` block but has no critical-line code)."""
import math

import numpy as np
import pytest

import ltmodel as M
import pyoracle as O


def _grid(n, half):
    return M.Grid(M.grid_param(-half, half, -half, half), n, n)


def _sis_det(grid, theta_e):
    x = grid.frame.xmin + np.arange(grid.nx) * grid.dx
    y = grid.frame.ymin + np.arange(grid.ny) * grid.dy
    X, Y = np.meshgrid(x, y)
    R = np.hypot(X, Y)
    with np.errstate(divide="ignore"):
        return 1.0 - theta_e / R, theta_e / (2 * R)


@pytest.mark.parametrize("n", [64, 201])
def test_sis_circle(n):
    theta_e = 7.3
    grid = _grid(n, 20.0)
    det, kap = _sis_det(grid, theta_e)
    segs, kinds = O.critic_lines(det, kap, grid)
    assert len(segs) > 20
    h = grid.dx
    for px, py in ((segs[:, 0], segs[:, 1]), (segs[:, 2], segs[:, 3])):
        r = np.hypot(px, py)
        assert np.abs(r - theta_e).max() < h * h / theta_e
    # closed: the end points are the start points, exactly
    a = sorted(map(tuple, segs[:, :2]))
    b = sorted(map(tuple, segs[:, 2:]))
    assert a == b
    # clockwise around the origin (positive det outside, on the left); area and length of the circle
    area = 0.5 * np.sum(segs[:, 0] * segs[:, 3] - segs[:, 2] * segs[:, 1])
    assert area < 0 and abs(-area - math.pi * theta_e ** 2) < 2 * math.pi * theta_e * h * h / theta_e + 1e-9
    length = np.hypot(segs[:, 2] - segs[:, 0], segs[:, 3] - segs[:, 1]).sum()
    assert abs(length - 2 * math.pi * theta_e) < 0.01 * 2 * math.pi * theta_e
    # kappa = 1/2 on the SIS critical circle: all tangential
    assert (kinds == 0).all()


def _cell(c, k=None):
    g = M.Grid(M.grid_param(0.0, 1.0, 0.0, 1.0), 2, 2)
    det = np.array([[c[0], c[1]], [c[3], c[2]]], dtype=float)
    kap = None if k is None else np.full((2, 2), float(k))
    return O.critic_lines(det, kap, g)


def test_single_cells():
    # BL positive: one segment from the bottom edge to the left edge
    s, _ = _cell([1.0, -1.0, -1.0, -1.0])
    np.testing.assert_array_equal(s, [[0.5, 0.0, 0.0, 0.5]])
    # bottom positive (BL, BR): right edge to left edge, t = 1/(1+3) on the right, 1/(1+1) on the left
    s, _ = _cell([1.0, 1.0, -3.0, -1.0])
    np.testing.assert_array_equal(s, [[1.0, 0.25, 0.0, 0.5]])
    # saddle, corner sum 0: positives apart (around BL, then around TR)
    s, _ = _cell([1.0, -1.0, 1.0, -1.0])
    np.testing.assert_array_equal(s, [[0.5, 0.0, 0.0, 0.5], [0.5, 1.0, 1.0, 0.5]])
    # saddle, corner sum > 0: positives joined (around the negative BR, then around TL)
    s, _ = _cell([2.0, -1.0, 2.0, -1.0])
    np.testing.assert_array_equal(s, [[2.0 / 3.0, 0.0, 1.0, 1.0 / 3.0], [1.0 / 3.0, 1.0, 0.0, 2.0 / 3.0]])
    # the other saddle, positives apart: around BR, then around TL
    s, _ = _cell([-1.0, 1.0, -1.0, 1.0])
    np.testing.assert_array_equal(s, [[1.0, 0.5, 0.5, 0.0], [0.0, 0.5, 0.5, 1.0]])
    # no crossing, and zero counts as non-positive
    assert len(_cell([1.0, 2.0, 3.0, 4.0])[0]) == 0 and len(_cell([0.0, 0.0, 0.0, 0.0])[0]) == 0
    # kinds: 1 - kappa > 0 tangential, otherwise radial
    assert _cell([1.0, -1.0, -1.0, -1.0], 0.3)[1].tolist() == [0]
    assert _cell([1.0, -1.0, -1.0, -1.0], 1.7)[1].tolist() == [1]


def test_oriented_positive_left():
    """A smooth field with many crossings: in every cell crossed by one segment, the cell corners
    left of the directed segment are positive and those right of it are not."""
    grid = _grid(40, 10.0)
    x = grid.frame.xmin + np.arange(grid.nx) * grid.dx
    X, Y = np.meshgrid(x, x)
    det = np.cos(0.7 * X + 0.2 * Y) * np.sin(0.5 * Y - 0.3 * X) + 0.05
    segs, _ = O.critic_lines(det, None, grid)
    assert len(segs) > 50
    cells = {}
    for k, (x0, y0, x1, y1) in enumerate(segs):
        i = int(np.floor((0.5 * (x0 + x1) - grid.frame.xmin) / grid.dx))
        j = int(np.floor((0.5 * (y0 + y1) - grid.frame.ymin) / grid.dy))
        cells.setdefault((i, j), []).append(k)
    checked = 0
    for (i, j), ks in cells.items():
        if len(ks) != 1:
            continue
        x0, y0, x1, y1 = segs[ks[0]]
        for ci, cj in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
            cx, cy = x[ci], x[cj]
            side = (x1 - x0) * (cy - y0) - (y1 - y0) * (cx - x0)
            if side > 0:
                assert det[cj, ci] > 0
            elif side < 0:
                assert det[cj, ci] <= 0
        checked += 1
    assert checked > 40


def test_count_only_and_overflow():
    grid = _grid(101, 20.0)
    det, kap = _sis_det(grid, 5.0)
    full, kinds = O.critic_lines(det, kap, grid)
    part, pk = O.critic_lines(det, kap, grid, max_seg=10)
    np.testing.assert_array_equal(part, full[:10])
    np.testing.assert_array_equal(pk, kinds[:10])
