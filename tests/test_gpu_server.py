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
