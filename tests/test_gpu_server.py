"""The map-serving loop (lt_server_*): requests of mixed kind, size and precision on several
streams, admitted by HBM footprint. Each result must be identical alone and in the mix (the same
per-pixel code runs whatever else is on the device), and equal to the synchronous entry point."""
import numpy as np
import pytest

import ltmodel as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _pool():
    a = M.cluster_lens(seed=5, ngal=40)
    b = M.group_lens(seed=6)
    c = M.cluster_lens(seed=7, ngal=120)
    for L in (a, b, c):
        L.dlsds[:] = 0.7
        L._soa = None
    return [a, b, c]


def _requests(hip, rng, n):
    kinds = [(hip.MAP_AMPLI, 1), (hip.MAP_SHEAR, 2), (hip.MAP_MASS, 1), (hip.MAP_POT, 1), (hip.MAP_DPL, 1),
             (hip.MAP_FIELDS, 0)]
    out = []
    for _ in range(n):
        k, m = kinds[rng.integers(len(kinds))]
        side = int(rng.choice([64, 128, 192, 320]))
        grid = M.Grid(M.grid_param(-100.0, 100.0, -90.0, 110.0), side, side + 64)
        prec = int(rng.integers(2))
        out.append((int(rng.integers(3)), k, m, grid, prec))
    return out


def _run(server, mids, reqs, order=None):
    tickets = [server.submit(mids[mi], k, m, g, precision=p, r=0.7, mult=1.0) for mi, k, m, g, p in reqs]
    res = [None] * len(reqs)
    for i in (order if order is not None else range(len(reqs))):
        _, k, _, g, _ = reqs[i]
        res[i], _ = server.fetch(tickets[i], k, g.npix)
    return res


def test_alone_equals_mixed(hip):
    rng = np.random.default_rng(3)
    pool = _pool()
    reqs = _requests(hip, rng, 18)
    print(f"[case] server alone vs mixed: {len(reqs)} requests", flush=True)
    srv = hip.Server(nslots=4, hbm_budget_bytes=1 << 30)
    try:
        mids = [srv.add_model(L) for L in pool]
        alone = [_run(srv, mids, [q])[0] for q in reqs]
        mixed = _run(srv, mids, reqs, order=rng.permutation(len(reqs)))
        for i, (a, b) in enumerate(zip(alone, mixed)):
            for pa, pb in zip(a, b):
                np.testing.assert_array_equal(pa, pb, err_msg=f"request {i}")
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"]
    finally:
        srv.close()
    # and the server's maps equal the synchronous entry point's
    cosmo = M.flat_lcdm()
    for mi, k, m, g, p in reqs[:6]:
        if k in (hip.MAP_AMPLI, hip.MAP_SHEAR):
            r = hip.cosmo_dlsds(0.4, 2.0, cosmo)
            want = hip.map_grid(k, m, pool[mi], g, cosmo, 0.4, 2.0, precision=p)
            srv2 = hip.Server(nslots=1, hbm_budget_bytes=1 << 28)
            mid = srv2.add_model(pool[mi])
            t = srv2.submit(mid, k, m, g, precision=p, r=r)
            got, _ = srv2.fetch(t, k, g.npix)
            srv2.close()
            np.testing.assert_array_equal(got[0].reshape(g.ny, g.nx), want)


def test_admission_by_memory(hip):
    """A budget that holds one large request at a time: everything still completes, in submission
    order for equal sizes, and the device memory in use never exceeds the budget."""
    pool = _pool()
    g = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), 512, 512)
    need = 5 * g.npix * 8
    srv = hip.Server(nslots=4, hbm_budget_bytes=int(need * 1.5))
    try:
        mid = srv.add_model(pool[0])
        tickets = [srv.submit(mid, hip.MAP_FIELDS, 0, g, precision=1, r=0.7) for _ in range(5)]
        done = []
        for t in tickets:
            _, tm = srv.fetch(t, hip.MAP_FIELDS, g.npix)
            done.append(tm.start_ms)
        assert done == sorted(done)
        st = srv.stats()
        assert st["bytes_peak"] <= int(need * 1.5)
        assert st["completed"] == 5
        # a request larger than the whole budget is refused at submission
        big = M.Grid(g.frame, 1024, 1024)
        with pytest.raises(hip.LtError):
            srv.submit(mid, hip.MAP_FIELDS, 0, big)
    finally:
        srv.close()


def test_unsupported_use_is_refused(hip):
    """A lens with an external-shear halo has no gradient: gradient requests fail at submission,
    Hessian requests run."""
    lens = M.LensModel.from_halos([dict(type=81, x=0.0, y=0.0, sigma=600.0, e=0.3, theta=0.2, rcore=3.0,
                                        rcut=300.0),
                                   dict(type=14, x=0.0, y=0.0, b0=0.0, epot=0.05, theta=0.4)])
    srv = hip.Server(nslots=2, hbm_budget_bytes=1 << 28)
    try:
        mid = srv.add_model(lens)
        g = M.Grid(M.grid_param(-50.0, 50.0, -50.0, 50.0), 64, 64)
        with pytest.raises(hip.LtError):
            srv.submit(mid, hip.MAP_DPL, 1, g)
        t = srv.submit(mid, hip.MAP_AMPLI, 1, g, r=0.7)
        out, _ = srv.fetch(t, hip.MAP_AMPLI, g.npix)
        assert np.isfinite(out[0]).all()
    finally:
        srv.close()


def test_tight_budget_mixed_sizes_in_submission_order(hip):
    """Sizes of 50 %, 60 % and 30 % of the budget, then a mix, fetched in submission order: every
    request completes (a later request never takes memory that an earlier one needs for good), the
    arena is never exceeded, and results equal those of a server with room for all."""
    pool = _pool()
    sides = [(512, 400), (512, 480), (512, 240), (256, 256), (512, 480), (128, 128), (512, 400), (384, 384)]
    unit = 5 * 8  # fields: five FP64 planes
    budget = int(512 * 800 * unit * 1.0)
    reqs = [(M.Grid(M.grid_param(-100.0, 100.0, -90.0, 110.0), nx, ny)) for nx, ny in sides]
    print(f"[case] server tight budget {budget} B, {len(reqs)} fields requests fetched in order", flush=True)
    srv = hip.Server(nslots=3, hbm_budget_bytes=budget)
    big = hip.Server(nslots=3, hbm_budget_bytes=1 << 31)
    try:
        mid, bid = srv.add_model(pool[0]), big.add_model(pool[0])
        tickets = [srv.submit(mid, hip.MAP_FIELDS, 0, g, precision=0, r=0.7) for g in reqs]
        got = [srv.fetch(t, hip.MAP_FIELDS, g.npix)[0] for t, g in zip(tickets, reqs)]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["bytes_peak"] <= budget
        for g, planes in zip(reqs, got):
            want, _ = big.fetch(big.submit(bid, hip.MAP_FIELDS, 0, g, precision=0, r=0.7), hip.MAP_FIELDS, g.npix)
            for a, b in zip(planes, want):
                np.testing.assert_array_equal(a, b)
    finally:
        srv.close()
        big.close()


def test_wait_that_could_never_end_is_busy(hip):
    """A caller waiting for a request that only the fetch of an earlier result could admit gets
    LT_ERR_BUSY instead of waiting forever; after that fetch the request runs."""
    import ctypes as C
    pool = _pool()
    g = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), 256, 256)
    need = 5 * g.npix * 8
    srv = hip.Server(nslots=2, hbm_budget_bytes=int(need * 1.5))
    try:
        mid = srv.add_model(pool[1])
        t1 = srv.submit(mid, hip.MAP_FIELDS, 0, g, r=0.7)
        t2 = srv.submit(mid, hip.MAP_FIELDS, 0, g, r=0.7)
        print("[case] server stuck wait: second request behind an unfetched first", flush=True)
        st = hip.lib().lt_server_wait(srv.h, t2, C.c_double(-1.0))
        assert st == 10  # LT_ERR_BUSY
        assert b"fetched" in hip.lib().lt_last_error()
        srv.fetch(t1, hip.MAP_FIELDS, g.npix)
        out, _ = srv.fetch(t2, hip.MAP_FIELDS, g.npix)
        assert np.isfinite(out[3]).all()
    finally:
        srv.close()


def test_wait_while_another_thread_fetches(hip):
    """A thread waiting on a ticket that another thread fetches returns (LT_OK if it saw the request
    finish first, LT_ERR_NOT_FOUND if the ticket was released during its wait) instead of touching
    the released ticket; the fetch gets the result."""
    import ctypes as C
    import threading
    pool = _pool()
    g = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), 1024, 1024)
    srv = hip.Server(nslots=2, hbm_budget_bytes=1 << 30)
    try:
        mid = srv.add_model(pool[2])
        print("[case] server wait on a ticket fetched by another thread", flush=True)
        for _ in range(4):
            t = srv.submit(mid, hip.MAP_POT, 1, g, r=0.7)
            got = []
            th = threading.Thread(target=lambda: got.append(hip.lib().lt_server_wait(srv.h, t, C.c_double(-1.0))))
            th.start()
            out, _ = srv.fetch(t, hip.MAP_POT, g.npix)
            th.join(timeout=60)
            assert not th.is_alive()
            assert got[0] in (0, 8)  # LT_OK or LT_ERR_NOT_FOUND
            assert np.isfinite(out[0]).sum() > 0
            assert hip.lib().lt_server_wait(srv.h, t, C.c_double(0.0)) == 8
    finally:
        srv.close()


def test_serving_workload_sizes(hip):
    """Requests at the serving workload's sizes: 8192^2 fields (kappa, gamma, det A, mu) of a
    2000-halo model, 8192^2 potential of a 2000-halo model, 8192^2 FP32C amplification, and small
    requests of 10-300 halos on 256^2-1024^2 grids. Each result is bit-identical alone and in the mix
    (large maps run as many row tiles interleaved with the small ones), within the HBM budget."""
    rng = np.random.default_rng(21)
    h2000 = M.cluster_lens(seed=31, ngal=1997)
    small = [M.cluster_lens(seed=40 + k, ngal=n) for k, n in enumerate((7, 40, 297))]
    for L in [h2000] + small:
        L.dlsds[:] = 0.7
        L._soa = None
    frame = M.grid_param(-150.0, 150.0, -150.0, 150.0)
    reqs = [(0, hip.MAP_FIELDS, 0, 8192, 0), (0, hip.MAP_POT, 1, 8192, 0), (0, hip.MAP_AMPLI, 1, 8192, 1)]
    kinds = [(hip.MAP_DPL, 1), (hip.MAP_AMPLI, 1), (hip.MAP_SHEAR, 1), (hip.MAP_MASS, 1), (hip.MAP_POT, 1)]
    for _ in range(9):
        k, m = kinds[rng.integers(len(kinds))]
        p = int(rng.integers(2))
        reqs.append((1 + int(rng.integers(3)), k, m, int(rng.choice([256, 512, 1024])), p))
    budget = 12 << 30
    print(f"[case] serving workload sizes: {len(reqs)} requests, 8192^2 fields/pot/ampli of 2000 halos", flush=True)
    srv = hip.Server(nslots=8, hbm_budget_bytes=budget)
    try:
        mids = [srv.add_model(L) for L in [h2000] + small]
        grids = [M.Grid(frame, n, n) for _, _, _, n, _ in reqs]
        # mixed: everything submitted at once (large first), fetched in submission order
        tickets = [srv.submit(mids[mi], k, m, g, precision=p, r=0.7) for (mi, k, m, _, p), g in zip(reqs, grids)]
        mixed = [srv.fetch(t, k, g.npix)[0] for t, (_, k, _, _, _), g in zip(tickets, reqs, grids)]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["bytes_peak"] <= budget
        # small requests finished before the large ones that were submitted ahead of them
        for i, ((mi, k, m, n, p), g) in enumerate(zip(reqs, grids)):
            t = srv.submit(mids[mi], k, m, g, precision=p, r=0.7)
            alone, _ = srv.fetch(t, k, g.npix)
            for a, b in zip(alone, mixed[i]):
                np.testing.assert_array_equal(a, b, err_msg=f"request {i}")
            del alone
            mixed[i] = None
    finally:
        srv.close()
