"""Batches of lens-model samples in one launch, and captured (graph) launch sequences.

Every pixel of every sample is computed by the same per-pixel code whatever the batch, so a batched
map must equal the single-model map bit for bit, in FP64 and in FP32C."""
import numpy as np
import pytest

import ltmodel as M
from parity import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _single(hip, lens, kind, mode, grid, r, precision, stream):
    m = hip.Model(lens, hip.MAP_HESSIAN)
    out = hip.DeviceArray(grid.npix)
    m.map(kind, mode, [out.ptr], grid, r=r, precision=precision, stream=stream)
    hip.sync(stream)
    v = out.download()
    m.close()
    out.close()
    return v


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single(hip, precision):
    base = M.group_lens(seed=4)
    samples = M.perturbed_samples(base, 6, seed=11)
    grid = M.group_grid(192)
    print_case(f"batch of 6 samples precision={precision}", base, grid)
    s = hip.stream_create()
    try:
        bm = hip.Model(samples, hip.MAP_HESSIAN)
        out = hip.DeviceArray(len(samples) * grid.npix)
        bm.map(hip.MAP_AMPLI, 1, [out.ptr], grid, r=0.7, precision=precision, stream=s)
        hip.sync(s)
        got = out.download().reshape(len(samples), grid.npix)
        for k, L in enumerate(samples):
            np.testing.assert_array_equal(got[k], _single(hip, L, hip.MAP_AMPLI, 1, grid, 0.7, precision, s))
        # update in place with the samples reversed: the maps follow
        bm.update_batch(samples[::-1], stream=s)
        bm.map(hip.MAP_AMPLI, 1, [out.ptr], grid, r=0.7, precision=precision, stream=s)
        hip.sync(s)
        again = out.download().reshape(len(samples), grid.npix)
        np.testing.assert_array_equal(again, got[::-1])
        bm.close()
        out.close()
    finally:
        hip.stream_destroy(s)


def test_graph_replays_upload_and_map(hip):
    """Capture [upload staged parameters, fields map]; replay it for three different staged samples
    with eager work on the same stream between replays; every replay equals the eager map."""
    base = M.group_lens(seed=4)
    samples = M.perturbed_samples(base, 3, seed=12)
    grid = M.group_grid(160)
    print_case("graph replay upload + fields", base, grid)
    s = hip.stream_create()
    try:
        m = hip.Model(base, hip.MAP_HESSIAN)
        outs = [hip.DeviceArray(grid.npix) for _ in range(5)]
        g = hip.Graph(s)
        m.upload(s)
        m.map(hip.MAP_FIELDS, 0, [o.ptr for o in outs], grid, r=0.7, stream=s)
        g.end()
        scratch = hip.DeviceArray(grid.npix)
        for L in samples:
            m.stage(L)
            g.launch()
            hip.sync(s)
            got = [o.download() for o in outs]
            # eager work between replays: another map on the same stream into other memory
            m2 = hip.Model(L, hip.MAP_HESSIAN)
            m2.map(hip.MAP_AMPLI, 1, [scratch.ptr], grid, r=0.7, stream=s)
            eager = [hip.DeviceArray(grid.npix) for _ in range(5)]
            m2.map(hip.MAP_FIELDS, 0, [e.ptr for e in eager], grid, r=0.7, stream=s)
            hip.sync(s)
            for a, e in zip(got, eager):
                np.testing.assert_array_equal(a, e.download())
            assert np.isfinite(scratch.download()).all()
            m2.close()
        g.close()
        m.close()
    finally:
        hip.stream_destroy(s)


def test_batch_rejects_mismatched_samples(hip):
    base = M.group_lens(seed=4)
    other = M.group_lens(seed=5)
    other.type[3] = 81
    other.rcore[3] = 0.2
    other.rcut[3] = 5.0
    other.epot[3] = 0.1
    other._soa = None
    with pytest.raises(hip.LtError):
        hip.Model([base, other], hip.MAP_HESSIAN)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_results(hip, precision):
    """The autotuner only changes the launch shape: maps before and after tuning are identical."""
    lens = M.cluster_lens(seed=8, ngal=30)
    grid = M.cluster_grid(384)
    print_case(f"autotune precision={precision}", lens, grid)
    s = hip.stream_create()
    try:
        m = hip.Model(lens, hip.MAP_HESSIAN)
        outs = [hip.DeviceArray(grid.npix) for _ in range(5)]
        m.map(hip.MAP_FIELDS, 0, [o.ptr for o in outs], grid, r=0.7, precision=precision, px=2, stream=s)
        hip.sync(s)
        before = [o.download() for o in outs]
        best, ms = m.autotune(hip.MAP_FIELDS, 0, grid, precision, s)
        assert best in (1, 2, 4) and all(t > 0 for t in ms)
        m.map(hip.MAP_FIELDS, 0, [o.ptr for o in outs], grid, r=0.7, precision=precision, stream=s)
        hip.sync(s)
        for b, o in zip(before, outs):
            np.testing.assert_array_equal(b, o.download())
        m.close()
    finally:
        hip.stream_destroy(s)


def test_memset_is_a_kernel_and_replays(hip):
    """lt_memset_d sets bytes by a fill kernel: ragged offsets and lengths are set exactly, bytes
    around them are untouched, and a captured [memset, map] sequence replays correctly (the graph
    holds no memset node)."""
    import ctypes as C
    n = 4099
    buf = hip.DeviceArray(n, dtype=np.uint8)
    lib = hip.lib()
    print("[case] lt_memset_d ragged ranges and a captured memset", flush=True)
    ref = np.full(n, 0x5A, np.uint8)
    buf.upload(ref)
    for off, ln, v in ((0, 1, 1), (1, 15, 2), (3, 40, 3), (17, 1000, 0xFF), (2049, 2050, 0), (4098, 1, 7), (5, 0, 9)):
        hip.check(lib.lt_memset_d(C.c_void_p(buf.ptr + off), v, ln, None))
        ref[off:off + ln] = v
        np.testing.assert_array_equal(buf.download(), ref)
    buf.close()
    grid = M.group_grid(96)
    s = hip.stream_create()
    try:
        m = hip.Model(M.group_lens(seed=4), hip.MAP_HESSIAN)
        out = hip.DeviceArray(grid.npix)
        flag = hip.DeviceArray(grid.npix)
        g = hip.Graph(s)
        hip.check(lib.lt_memset_d(C.c_void_p(flag.ptr), 0, flag.nbytes, C.c_void_p(s)))
        m.map(hip.MAP_AMPLI, 1, [out.ptr], grid, r=0.7, stream=s)
        g.end()
        want = _single(hip, M.group_lens(seed=4), hip.MAP_AMPLI, 1, grid, 0.7, 0, s)
        for rep in range(3):
            flag.upload(np.full(grid.npix, 1.5))
            out.upload(np.zeros(grid.npix))
            g.launch()
            hip.sync(s)
            assert (flag.download() == 0).all()
            np.testing.assert_array_equal(out.download(), want)
        g.close()
        m.close()
        out.close()
        flag.close()
    finally:
        hip.stream_destroy(s)
