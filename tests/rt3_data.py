"""The reference's result test (Benchmarks/ResultTest): a 6-halo PIEMD (type 8) lens
(TestResultParameter.par, z_lens 0.4), 30 images of 10 families at z = 1.5
(theoretical_images_time.txt), and what the original Lenstool computed for them: the source position
of every image (source.dat, 7 decimals), the amplification of every image found again (dist.dat,
AMP column, 3 decimals; positions at 3 decimals) and chi^2 = 0 (chires.dat). Data fixtures under
tests/golden/rt3, read here as text."""
import os

import numpy as np

from conftest import GOLDEN

RT3 = os.path.join(GOLDEN, "rt3")
Z_LENS = 0.4


def load(lthip):
    """(parameter-file dict, images (n, 3): family, x, y, source.dat (n, 2), dist.dat rows (family, x, y, AMP))."""
    d = lthip.load_par(os.path.join(RT3, "TestResultParameter.par"))
    img = np.loadtxt(os.path.join(RT3, "theoretical_images_time.txt"))[:, :3]
    src = np.loadtxt(os.path.join(RT3, "source.dat"), comments="#")[:, 1:3]
    rows = []
    for line in open(os.path.join(RT3, "dist.dat")):
        t = line.split()
        if t and not t[0].startswith("#"):
            rows.append((int(t[0]), float(t[1]), float(t[2]), float(t[6])))
    return d, img, src, np.array(rows)


def amp_rows(rows, fam, x, y):
    """AMP of the dist.dat rows of family `fam` whose 3-decimal position is within 1e-3 of (x, y)."""
    m = (rows[:, 0] == fam) & (np.abs(rows[:, 1] - x) <= 1.0e-3 + 1e-12) & (np.abs(rows[:, 2] - y) <= 1.0e-3 + 1e-12)
    return rows[m, 3]


def amp_tolerance(mu):
    """dist.dat's AMP is printed at 3 decimals for images found again at slightly different positions
    (3-decimal coordinates): half a unit of the last digit plus the change of mu over such a shift,
    which grows as mu^2 next to critical lines."""
    return 5e-4 + 1e-6 * mu * mu
