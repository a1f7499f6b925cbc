"""Delensing + chi^2 at its full size: 1024 lens samples of the 1 dPIE + 10 SIS lens against 40
families of 3 images (source z ~ U[1, 6], images from the oracle's delensing of the true model on a
1024^2 grid + N(0, 0.1") noise), delensing grid 512^2 over 200", all samples in one ChiBatch run.

Size-independent properties over the whole batch, checked members against the single-model path and
the oracle:
  * a member of the batch equals the single-model chi^2 path bitwise, and the oracle within the
    gradient tolerance (1e-9 relative on chi^2; delensed images equal in count, within 1e-12");
  * a permutation of the samples permutes the results bitwise (no cross-sample state);
  * the true model, placed in the batch, fits the catalogue better than the median perturbed one."""
import numpy as np
import pytest

import ltmodel as M

pytestmark = pytest.mark.gpu

NSAMPLES, N, NFAM = 1024, 512, 40


@pytest.fixture(scope="module")
def lt():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _catalogue(lt, oracle, lens, frame, cosmo):
    rng = np.random.default_rng(41)
    fine = 1024
    gx, gy = oracle.gradient_grid(lens, M.Grid(frame, fine, fine))
    fams = []
    while len(fams) < NFAM:
        src = M.galaxies_array(16)
        src["cx"] = rng.uniform(-20, 20, 16)
        src["cy"] = rng.uniform(-20, 20, 16)
        zs = rng.uniform(1.0, 6.0, 16)
        src["dr"] = [lt.cosmo_dlsds(lens.z[0], z, cosmo) for z in zs]
        _, pos, loc, _ = oracle.delense_barycenter(16, fine, frame, src, gx.ravel(), gy.ravel())
        for s in range(16):
            if loc[s] >= 3 and len(fams) < NFAM:
                fams.append((pos[s, :3].copy(), src["dr"][s], zs[s]))
    nimages = np.full(NFAM, 3, dtype=np.int32)
    images = M.galaxies_array(3 * NFAM)
    for f, (p, dr, z) in enumerate(fams):
        for k in range(3):
            images[3 * f + k]["cx"] = p[k, 0] + rng.normal(0, 0.1)
            images[3 * f + k]["cy"] = p[k, 1] + rng.normal(0, 0.1)
            images[3 * f + k]["dr"] = dr
            images[3 * f + k]["redshift"] = z
    return nimages, images


def test_chi_fullsize_batch(lt, oracle):
    base = M.group_lens(seed=1)
    cosmo = M.flat_lcdm(70.0, 0.3)
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    nimages, images = _catalogue(lt, oracle, base, frame, cosmo)
    samples = M.perturbed_samples(base, NSAMPLES, seed=9)
    true_at = 777
    samples[true_at] = base
    print(f"[case] chi full size: {NSAMPLES} samples, {NFAM} families, {N}^2 delensing grid", flush=True)
    b = lt.ChiBatch(NSAMPLES, base.type, N, frame, nimages, images)
    try:
        chi, found, status = b.run(samples)
        assert np.isfinite(chi).all() and (status == 0).all()
        assert chi[true_at] < np.median(chi)
        for s in (0, 1, 300, true_at, 1022, 1023):
            st1, chi1 = lt.chi_barycentersource(samples[s], N, frame, nimages, images)
            assert st1 == status[s] and chi1 == chi[s], s
            ost, ochi = oracle.chi_barycentersource(samples[s], N, frame, nimages, images)
            assert ost == status[s]
            assert abs(chi[s] - ochi) <= 1e-9 * abs(ochi), s
        for s in (5, true_at):
            pos, loc = b.images(s)
            src = oracle.source_barycenters(nimages, images, samples[s])
            gx, gy = oracle.gradient_grid(samples[s], M.Grid(frame, N, N))
            _, opos, oloc, _ = oracle.delense_barycenter(NFAM, N, frame, src, gx.ravel(), gy.ravel())
            np.testing.assert_array_equal(loc, oloc)
            np.testing.assert_allclose(pos, opos, rtol=0, atol=1e-12)
        perm = np.random.default_rng(3).permutation(NSAMPLES)
        print("[case] chi full size: permuted samples", flush=True)
        chi_p, found_p, status_p = b.run([samples[i] for i in perm])
        np.testing.assert_array_equal(chi_p, chi[perm])
        np.testing.assert_array_equal(found_p, found[perm])
        np.testing.assert_array_equal(status_p, status[perm])
    finally:
        b.close()
