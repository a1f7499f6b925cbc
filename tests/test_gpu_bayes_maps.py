"""The reference's bayes-map configuration end to end: utils/maps' m1931.par (3 dPIE potentials with
limit blocks, 203 + 3 potfile clumps) and its bayes.dat (2 MCMC samples), through lenstool_maps
(read, apply samples, batched maps of every runmode kind, FITS files).

Each FITS image must equal, as float32, the map the library computes in this process for the
sample's model read back with lt_par_bayes_apply (the batched launch computes every pixel of every
sample with the same code as a single-model launch), and the Hessian of sample 0 must pass the FP64
parity criterion against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import ltmodel as M
from conftest import GOLDEN
from parity import EXT_SHARE, REL

pytestmark = pytest.mark.gpu

MAPS = os.path.join(GOLDEN, "maps")
TOOL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "lenstool-hpc_amd", "lenstool_maps")


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _fits(path):
    raw = open(path, "rb").read()
    cards = []
    while not cards or not cards[-1].startswith("END "):
        cards.append(raw[80 * len(cards):80 * len(cards) + 80].decode("ascii"))
    end = len(cards) - 1
    kv = {c[:8].strip(): c[10:].split(" / ")[0].strip() for c in cards[:end] if c[8:10] == "= "}
    nx, ny = int(kv["NAXIS1"]), int(kv["NAXIS2"])
    off = ((end + 1) * 80 + 2879) // 2880 * 2880
    return kv, np.frombuffer(raw[off:off + 4 * nx * ny], dtype=">f4").reshape(ny, nx)


def test_maps_tool_bayes_samples(hip, oracle, tmp_path):
    print("[case] lenstool_maps m1931.par + bayes.dat: 209 halos, 2 samples, 200x200 maps", flush=True)
    assert os.path.exists(TOOL), "lenstool_maps is not built"
    r = subprocess.run([TOOL, os.path.join(MAPS, "m1931.par"), "--bayes", os.path.join(MAPS, "bayes.dat"),
                        "--out", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    with hip.ParFile(os.path.join(MAPS, "m1931.par")) as p:
        d = p.get()
        info, frame, cosmo = d["info"], d["frame"], d["cosmo"]
        p.bayes_load(os.path.join(MAPS, "bayes.dat"))
        models = p.bayes_models()
    assert len(models) == 2
    jobs = [("mass_4", hip.MAP_MASS, info["mass"], info["mass_cells"], info["mass_z"], info["z_source"]),
            ("ampli_5", hip.MAP_AMPLI, info["ampli"], info["ampli_cells"], None, info["ampli_z"]),
            ("shear_1", hip.MAP_SHEAR, info["shear"], info["shear_cells"], None, info["shear_z"])]
    for k, lens in enumerate(models):
        for name, kind, mode, cells, zl, zs in jobs:
            kv, img = _fits(tmp_path / f"{name}_{k}.fits")
            assert img.shape == (cells, cells) and kv["CTYPE1"] == "'RA---TAN'"
            grid = M.Grid(frame, cells, cells)
            want = hip.map_grid(kind, mode, lens, grid, cosmo, zl if zl is not None else lens.z[0], zs)
            np.testing.assert_array_equal(img, want.astype(np.float32), err_msg=f"{name} sample {k}")
    # the sample-0 model on the GPU against the oracle (FP64 criterion), 200 x 200 over the frame
    lens = models[0]
    grid = M.Grid(frame, 200, 200)
    a, b, c, _ = hip.hessian_grid(lens, grid)
    ea, eb, ec, _ = oracle.hessian_planes(lens, grid, extended=True)
    _, _, _, _, aa, ab, ac = oracle.hessian_planes(lens, grid, with_abs=True)
    _, ch = oracle.cond_planes(lens, grid)
    for got, ex, sc in ((a, ea, aa), (b, eb, ab), (c, ec, ac)):
        assert (np.abs(got - ex) <= REL * sc + EXT_SHARE * ch).all()
