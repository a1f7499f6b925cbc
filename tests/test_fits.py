"""FITS map writer (lt_fits_write; CPU only).

The absolute-coordinate header is pinned byte for byte, DATE aside, against the header Lenstool
wrote with cfitsio for a 200 x 200 mass map of A2744 (reference point 3.586259 -30.400174, frame
+-100 arcsec: Benchmarks/ConfigFiles/hff_slmodels/A2744/ellipticite/mass.fits, first record). The
data records are read back here with numpy as big-endian float32, the FITS standard's layout."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

HDR = os.path.join(GOLDEN, "fits", "a2744_mass_header.bin")


@pytest.fixture(scope="module")
def lt():
    import lthip
    return lthip


def _cards(raw):
    out = []
    for k in range(0, len(raw), 80):
        c = raw[k:k + 80].decode("ascii")
        out.append(c)
        if c.startswith("END "):
            break
    return out


def _read(path):
    """(cards, data) of a single-image FITS file, parsed with the standard's rules."""
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    cards = _cards(raw)
    hdr_len = ((len(cards) * 80 + 2879) // 2880) * 2880
    assert raw[len(cards) * 80:hdr_len] == b" " * (hdr_len - len(cards) * 80)
    kv = {}
    for c in cards:
        if c[8:10] == "= ":
            v = c[10:].split(" / ")[0].strip()
            kv[c[:8].strip()] = v.strip("'").rstrip() if v.startswith("'") else v
    nx, ny = int(kv["NAXIS1"]), int(kv["NAXIS2"])
    assert kv["BITPIX"] == "-32" and kv["NAXIS"] == "2" and kv["SIMPLE"] == "T"
    data = np.frombuffer(raw[hdr_len:hdr_len + 4 * nx * ny], dtype=">f4").reshape(ny, nx)
    assert raw[hdr_len + 4 * nx * ny:] == b"\0" * (len(raw) - hdr_len - 4 * nx * ny)
    return cards, kv, data


def test_absolute_header_matches_lenstool(lt, tmp_path):
    ref = _cards(open(HDR, "rb").read())
    img = np.random.default_rng(1).normal(size=(200, 200))
    out = tmp_path / "mass.fits"
    lt.fits_write(out, img, -100.0, 100.0, -100.0, 100.0, ra=3.586259, dec=-30.400174)
    cards, kv, data = _read(out)
    assert len(cards) == len(ref)
    for a, b in zip(cards, ref):
        if a.startswith("DATE    "):
            assert b.startswith("DATE    ") and a[30:] == b[30:] and len(a) == 80
            continue
        assert a == b, (a, b)
    np.testing.assert_array_equal(data, img.astype(np.float32))


def test_relative_header_and_data(lt, tmp_path):
    ny, nx = 30, 50   # not square: NAXIS1 is x, rows are y
    img = np.arange(ny * nx, dtype=np.float64).reshape(ny, nx) * 1.1e-3 - 0.5
    out = tmp_path / "rel.fits"
    lt.fits_write(out, img, -136.5, 136.5, -20.0, 40.0)
    cards, kv, data = _read(out)
    assert (kv["NAXIS1"], kv["NAXIS2"]) == ("50", "30")
    assert kv["CTYPE1"] == "LINEAR" and kv["CUNIT2"] == "arcsec" and "EQUINOX" not in kv
    # float keywords, 7 significant digits, pixel 1 at (xmin, ymin)
    assert kv["CRVAL1"] == "-136.5" and kv["CRPIX1"] == "1." and kv["CRVAL2"] == "-20."
    assert float(kv["CDELT1"]) == pytest.approx(273.0 / 49, rel=1e-6) and kv["CDELT1"] == "%.7G" % np.float32(273.0 / 49)
    assert float(kv["CDELT2"]) == pytest.approx(60.0 / 29, rel=1e-6)
    np.testing.assert_array_equal(data, img.astype(np.float32))
    order = [c[:8].strip() for c in cards]
    assert order[:6] == ["SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "EXTEND"] and order[-1] == "END"


def test_fits_errors(lt, tmp_path):
    with pytest.raises(lt.LtError):
        lt.fits_write(tmp_path / "x.fits", np.zeros((1, 5)), 0, 1, 0, 1)
    with pytest.raises(lt.LtError):
        lt.fits_write(tmp_path / "x.fits", np.zeros((4, 4)), 1, 1, 0, 1)
    with pytest.raises(lt.LtError):
        lt.fits_write(tmp_path / "no" / "such" / "dir.fits", np.zeros((4, 4)), 0, 1, 0, 1)
