"""bayes.dat samples applied to a parameter file's model (lt_par_bayes_*; CPU only).

Pinned by the column headers Lenstool wrote into the reference's utils/maps/bayes.dat: each
'#O<n> : <name>' line names the potential (file order) and parameter its column holds, and
'#Pot<n> ...' the potfile scaling parameter. The checks read the mapping from those names, not from
the reader's rule, so a shifted, skipped or mis-ordered column fails. The 23rd column ('#Chi2') is
what the reference's rule hands to the second potfile's freed sigma (the .par frees it, the run
that wrote the file did not): that is kept, as the reference does it. The ordering of potentials
of different types (slots by ascending type) against limit blocks (file order) is pinned on a
synthetic file whose potentials are identified by their positions."""
import math
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN

MAPS = os.path.join(GOLDEN, "maps")
PI_C2 = 7.209970e-06
DTR = math.pi / 180


@pytest.fixture(scope="module")
def lt():
    import lthip
    return lthip


def _columns():
    """(name of each column, rows) of the bayes.dat fixture, parsed here from its text."""
    names, rows = [], []
    for s in open(os.path.join(MAPS, "bayes.dat")):
        if s.startswith("#"):
            names.append(s[1:].strip())
        elif s.strip():
            rows.append([float(v) for v in s.split()])
    return names, np.array(rows)


def test_bayes_columns_follow_the_headers(lt):
    names, rows = _columns()
    with lt.ParFile(os.path.join(MAPS, "m1931.par")) as p:
        base = p.get()
        ns, nc = p.bayes_load(os.path.join(MAPS, "bayes.dat"))
        assert (ns, nc) == (len(rows), len(names)) == (2, 23)
        info = base["info"]
        assert info["nhalos_main"] == 3 and info["npotfiles"] == 2
        # catalogue clumps: every line not starting with '#' in its first column (the reference
        # also reads the indented '  #1009.0 ...' lines as clumps), 3 group halos in the second
        cat = os.path.join(MAPS, "m1931_CM_zCLASH-MUSE.cat")
        n1 = sum(1 for s in open(cat) if not s.startswith("#") and s.strip())
        assert n1 == 203 and info["nhalos"] == 3 + n1 + 3
        for k in range(ns):
            p.bayes_apply(k)
            d = p.get()
            L, e = d["lens"], d["extra"]
            row = rows[k]
            seen = set()
            for c, name in enumerate(names):
                if not name.startswith("O"):
                    continue
                halo = int(name[1:name.index(" ")]) - 1   # all three potentials are dPIE: slot = file order
                field = name.split(":")[1].strip().split(" ")[0]
                v = row[c]
                got = {"x": L.x, "y": L.y, "emass": e["ellipticity"], "rc": L.rcore, "sigma": e["vdisp"],
                       "rcut": L.rcut}.get(field)
                if field == "theta":
                    assert L.angle[halo] == v * DTR and e["anglecos"][halo] == math.cos(v * DTR)
                else:
                    assert got[halo] == v, (name, got[halo], v)
                seen.add((halo, field))
            assert len(seen) == 14
            for h in range(3):
                assert L.b0[h] == pytest.approx(6 * PI_C2 * e["vdisp"][h] ** 2, rel=1e-15)
                ell = e["ellipticity"][h]
                assert L.epot[h] == pytest.approx((1 - math.sqrt(1 - ell * ell)) / ell, rel=1e-14)
            # parameters without a limit keep their read values
            assert L.rcut[0] == base["lens"].rcut[0] and L.rcore[2] == base["lens"].rcore[2]
            # potfile 0: rcut (arcsec) and sigma from the 'Pot0' columns, through the scaling relations
            cut0, sig0 = row[names.index("Pot0 rcut (arcsec)")], row[names.index("Pot0 sigma (km/s)")]
            mag = e["mag"]
            first = slice(3, 3 + n1)
            f10 = 10 ** (0.4 * (19.65 - mag[first]))
            np.testing.assert_allclose(e["vdisp"][first], sig0 * f10 ** (1 / 4.0), rtol=1e-14)
            np.testing.assert_allclose(L.rcut[first], cut0 * f10 ** (2 / 4.0), rtol=1e-14)
            np.testing.assert_allclose(L.b0[first], 6 * PI_C2 * e["vdisp"][first] ** 2, rtol=1e-14)
            # cores are not free: unchanged
            np.testing.assert_array_equal(L.rcore[first], base["lens"].rcore[first])
            # potfile 1 frees only sigma: the 'Chi2' column; its cut stays the catalogue's kpc value
            sig1 = row[names.index("Chi2")]
            second = slice(3 + n1, 6 + n1)
            f27 = 10 ** (0.4 * (27.0 - mag[second]))
            np.testing.assert_allclose(e["vdisp"][second], sig1 * f27 ** (1 / 4.0), rtol=1e-14)
            np.testing.assert_array_equal(L.rcut[second], base["lens"].rcut[second])
            # positions, angles and D_ls/D_s of the clumps do not move
            np.testing.assert_array_equal(L.x[3:], base["lens"].x[3:])
            np.testing.assert_array_equal(L.dlsds, base["lens"].dlsds)


def test_samples_in_place_like_the_reference(lt):
    """The model is updated in place and b0 / the ellipticity pair are recomputed from the current
    values at every sample, as the reference does. For the round clumps of the second potfile
    (a = b in the catalogue, so ellipticity 0) that walks: read epot = 1e-5; first sample
    ellipticity = 2 epot / (1 + epot^2); from the second sample on epot = (1 - sqrt(1 - e^2)) / e,
    then stationary. Free parameters follow the sample whatever came before."""
    with lt.ParFile(os.path.join(MAPS, "m1931.par")) as p:
        d0 = p.get()
        rnd = slice(d0["info"]["nhalos"] - 3, d0["info"]["nhalos"])
        assert (d0["extra"]["ellipticity"][rnd] == 0).all() and (d0["lens"].epot[rnd] == 1e-5).all()
        p.bayes_load(os.path.join(MAPS, "bayes.dat"))
        p.bayes_apply(0)
        d1 = p.get()
        e1 = 2 * 1e-5 / (1 + 1e-5 * 1e-5)
        np.testing.assert_array_equal(d1["extra"]["ellipticity"][rnd], e1)
        np.testing.assert_array_equal(d1["lens"].epot[rnd], 1e-5)
        p.bayes_apply(1)
        d2 = p.get()
        np.testing.assert_array_equal(d2["lens"].epot[rnd], (1 - math.sqrt(1 - e1 * e1)) / e1)
        p.bayes_apply(0)
        a = p.get()["lens"]
        p.bayes_apply(1)
        p.bayes_apply(0)
        b = p.get()["lens"]
        for f in ("x", "y", "b0", "angle", "epot", "rcore", "rcut"):
            np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
        # the free parameters of sample 0 are the same after any history
        np.testing.assert_array_equal(a.x[:3], d1["lens"].x[:3])
        np.testing.assert_array_equal(a.b0[:3], d1["lens"].b0[:3])


SYNTH = """runmode
\tnbgridcells 50
\tend
image
\tz_m_limit 1 2.0 1 1.0 6.0 0.1
\tend
source
\tz_source 2.0
\tend
cosmologie
\tH0 70.
\tomega 0.3
\tlambda 0.7
\tend
potentiel A
\tprofil 81
\tx_centre 1.0
\ty_centre 0.0
\tellipticite 0.2
\tangle_pos 10.0
\tcore_radius 0.5
\tcut_radius 100.0
\tv_disp 900.
\tz_lens 0.3
\tend
limit A
\tv_disp 1 500. 1500. 0.1
\tend
potentiel B
\tprofil 5
\tx_centre 2.0
\ty_centre 0.0
\tellipticite 0.1
\tv_disp 300.
\tz_lens 0.3
\tend
limit B
\tx_centre 1 0. 5. 0.1
\tv_disp 1 100. 500. 0.1
\tend
potentiel C
\tprofil 81
\tx_centre 3.0
\ty_centre 0.0
\tellipticite 0.0
\tcore_radius 0.1
\tcut_radius 10.0
\tv_disp 200.
\tz_lens 0.3
\tend
limit C
\tellipticite 1 0. 0.5 0.1
\tangle_pos 1 0. 180. 0.1
\trcut 1 1. 50. 0.1
\tend
fini
"""


def test_limits_follow_file_order_not_slots(lt, tmp_path):
    """Potential B (SIS, type 5) takes slot 0 ahead of A and C (dPIE 81); the columns still go
    limit block by limit block in file order: A's sigma, B's x and sigma, C's e, angle, rcut."""
    par = tmp_path / "synth.par"
    par.write_text(SYNTH)
    cols = ["Nsample", "ln(Lhood)", "A sigma", "B x", "B sigma", "C emass", "C theta", "C rcut"]
    vals = [[1, -3.0, 1100.0, 2.5, 350.0, 0.3, 40.0, 20.0], [2, -2.0, 1200.0, 2.25, 250.0, 0.4, 50.0, 25.0]]
    bay = tmp_path / "bayes.dat"
    bay.write_text("".join("#" + c + "\n" for c in cols) + "".join(" ".join(map(str, v)) + "\n" for v in vals))
    with lt.ParFile(str(par)) as p:
        base = p.get()
        assert list(base["lens"].type) == [5, 81, 81]
        assert list(base["lens"].x) == [2.0, 1.0, 3.0]   # B, A, C
        assert p.bayes_load(str(bay)) == (2, 8)
        for k, v in enumerate(vals):
            p.bayes_apply(k)
            d = p.get()
            L, e = d["lens"], d["extra"]
            # slot 1 = A
            assert e["vdisp"][1] == v[2] and L.x[1] == 1.0
            assert L.b0[1] == pytest.approx(6 * PI_C2 * v[2] ** 2, rel=1e-15)
            # slot 0 = B (SIS): x and sigma; b0 = 4 pi_c2 sigma^2; after a bayes update the SIS
            # mass ellipticity follows the potential one (ell = 3 epot)
            assert L.x[0] == v[3] and e["vdisp"][0] == v[4]
            assert L.b0[0] == pytest.approx(4 * PI_C2 * v[4] ** 2, rel=1e-15)
            assert e["ellipticity"][0] == pytest.approx(3 * L.epot[0], rel=1e-15)
            # slot 2 = C: ellipticity, angle, rcut
            assert e["ellipticity"][2] == v[5] and L.angle[2] == v[6] * DTR and L.rcut[2] == v[7]
            assert L.epot[2] == pytest.approx((1 - math.sqrt(1 - v[5] ** 2)) / v[5], rel=1e-14)
            assert L.rcore[2] == 0.1


def test_bayes_errors(lt, tmp_path):
    par = tmp_path / "synth.par"
    par.write_text(SYNTH)
    with lt.ParFile(str(par)) as p:
        with pytest.raises(lt.LtError):
            p.bayes_load(str(tmp_path / "absent.dat"))
        # more values than columns
        bad = tmp_path / "wide.dat"
        bad.write_text("#a\n#b\n1 2 3\n")
        with pytest.raises(lt.LtError):
            p.bayes_load(str(bad))
        # a sample too short for the free parameters is refused and leaves the model alone
        short = tmp_path / "short.dat"
        short.write_text("#a\n#b\n#c\n#d\n1 -1 1000 2.2\n")
        assert p.bayes_load(str(short)) == (1, 4)
        before = p.get()["lens"]
        with pytest.raises(lt.LtError):
            p.bayes_apply(0)
        after = p.get()["lens"]
        np.testing.assert_array_equal(before.b0, after.b0)
        np.testing.assert_array_equal(before.x, after.x)
        with pytest.raises(lt.LtError):
            p.bayes_apply(5)
        # not a number
        nan = tmp_path / "text.dat"
        nan.write_text("#a\n#b\n1 x\n")
        with pytest.raises(lt.LtError):
            p.bayes_load(str(nan))


def test_catalogue_beside_the_par(lt, tmp_path):
    """A relative catalogue name is found next to the .par, wherever the caller runs."""
    for f in os.listdir(MAPS):
        shutil.copyfile(os.path.join(MAPS, f), tmp_path / f)
    d = lt.load_par(str(tmp_path / "m1931.par"))
    assert d["info"]["nhalos"] == 209


def test_cline_block(lt):
    """utils/maps' m1931.par: `nplan 0 1.49 1.894 2.497`, `dmax 150.0`, `pas .1` -- no plane asked,
    the redshifts still read, as the reference's reader reads them."""
    with lt.ParFile(os.path.join(MAPS, "m1931.par")) as p:
        k = p.cline()
    assert k["nplan"] == 0 and k["z"][:3] == [1.49, 1.894, 2.497] and k["dmax"] == 150.0 and k["step"] == 0.1
    assert k["limit_high"] == 10 and k["nbgridcells"] == 0
