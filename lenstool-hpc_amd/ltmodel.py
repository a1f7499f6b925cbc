"""ctypes mirror of the C data model (include/lenstool_hip/lt_types.h) and synthetic lens models.

The structures have the field order of the reference's ``src/structure_hpc.hpp`` (double build).
``LensModel`` owns numpy arrays for every field the lens-map path reads and hands out an
``LtPotentialSOA`` view of them; halos are kept grouped by ascending type as the reference's
parameter reader slots them (``src/module_readParameters.cpp:2455-2478``).

The synthetic workloads follow the five benchmark configurations of this project (single FP64
gradient grid, cluster Hessian maps, batched Bayes maps, delensing + chi^2, map serving); every
generator takes an explicit seed and uses numpy's PCG64.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

PI_C2 = 7.209970e-06
MAXIMPERSOURCE = 20

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class LtPoint(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


class LtMatrix(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("d", C.c_double)]


class LtEllipse(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("theta", C.c_double)]


class LtGalaxy(C.Structure):
    _fields_ = [("center", LtPoint), ("shape", LtEllipse), ("mag", C.c_double),
                ("redshift", C.c_double), ("dls", C.c_double), ("dos", C.c_double),
                ("dr", C.c_double)]


class LtGridParam(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("xmin", "xmax", "ymin", "ymax", "lmin", "lmax", "rmax")]


class LtCosmoParam(C.Structure):
    _fields_ = [("model", C.c_int)] + [(n, C.c_double) for n in
                                       ("omegaM", "omegaX", "curvature", "wX", "wa", "H0", "h")]


class LtPotentialSOA(C.Structure):
    _fields_ = [("type", _ip), ("type_name", C.c_char * 10), ("name", _dp),
                ("position_x", _dp), ("position_y", _dp), ("weight", _dp), ("b0", _dp),
                ("vdisp", _dp), ("ellipticity_angle", _dp), ("ellipticity", _dp),
                ("ellipticity_potential", _dp), ("rcore", _dp), ("rcut", _dp), ("rscale", _dp),
                ("exponent", _dp), ("alpha", _dp), ("einasto_kappacritic", _dp), ("z", _dp),
                ("mag", _dp), ("lum", _dp), ("theta", _dp), ("anglecos", _dp), ("anglesin", _dp),
                ("sigma", _dp), ("dlsds", _dp), ("SOA_index", _ip), ("N_types", _ip)]


GALAXY_DTYPE = np.dtype([("cx", "f8"), ("cy", "f8"), ("sa", "f8"), ("sb", "f8"), ("st", "f8"),
                         ("mag", "f8"), ("redshift", "f8"), ("dls", "f8"), ("dos", "f8"),
                         ("dr", "f8")])
assert GALAXY_DTYPE.itemsize == C.sizeof(LtGalaxy)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def grid_param(xmin, xmax, ymin, ymax) -> LtGridParam:
    g = LtGridParam()
    g.xmin, g.xmax, g.ymin, g.ymax = float(xmin), float(xmax), float(ymin), float(ymax)
    return g


def flat_lcdm(H0=70.0, omegaM=0.3) -> LtCosmoParam:
    """Flat Lambda-CDM as the reference reads it: omegaX = 1 - omegaM, h = H0/50
    (src/module_readParameters.cpp:347-352)."""
    c = LtCosmoParam()
    c.model = 1
    c.omegaM = omegaM
    c.omegaX = 1.0 - omegaM
    c.curvature = 0.0
    c.wX = -1.0
    c.wa = 0.0
    c.H0 = H0
    c.h = H0 / 50.0
    return c


def epot_from_e(t: int, e: float) -> float:
    """Potential ellipticity from the mass ellipticity (src/module_readParameters.cpp:2872-2930)."""
    if t == 5:
        return e / 3.0
    if e == 0.0:
        return 1e-5
    return (1.0 - math.sqrt(1.0 - e * e)) / e


def b0_from_sigma(t: int, sigma: float) -> float:
    """Impact parameter (src/module_readParameters.cpp:2880-2920): 4 pi_c2 s^2 (SIS), 6 pi_c2 s^2."""
    return (4.0 if t == 5 else 6.0) * PI_C2 * sigma * sigma


@dataclass
class LensModel:
    """SOA lens model: halos grouped by type; all arrays float64 / int32, C-contiguous."""
    type: np.ndarray
    x: np.ndarray
    y: np.ndarray
    b0: np.ndarray
    angle: np.ndarray
    epot: np.ndarray
    rcore: np.ndarray
    rcut: np.ndarray
    z: np.ndarray
    dlsds: np.ndarray
    _soa: LtPotentialSOA = field(default=None, repr=False)

    def __post_init__(self):
        self.type = np.ascontiguousarray(self.type, dtype=np.int32)
        for n in ("x", "y", "b0", "angle", "epot", "rcore", "rcut", "z", "dlsds"):
            setattr(self, n, np.ascontiguousarray(getattr(self, n), dtype=np.float64))
        self.cos = np.cos(self.angle)
        self.sin = np.sin(self.angle)
        n = self.n
        for a in (self.x, self.y, self.b0, self.angle, self.epot, self.rcore, self.rcut, self.z,
                  self.dlsds):
            assert a.shape == (n,)
        self._soa = None

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def soa(self) -> LtPotentialSOA:
        if self._soa is None:
            s = LtPotentialSOA()
            s.type = iptr(self.type)
            s.position_x = dptr(self.x)
            s.position_y = dptr(self.y)
            s.b0 = dptr(self.b0)
            s.ellipticity_angle = dptr(self.angle)
            s.ellipticity_potential = dptr(self.epot)
            s.rcore = dptr(self.rcore)
            s.rcut = dptr(self.rcut)
            s.z = dptr(self.z)
            s.anglecos = dptr(self.cos)
            s.anglesin = dptr(self.sin)
            s.dlsds = dptr(self.dlsds)
            self._soa = s
        return self._soa

    def refresh(self):
        """Recompute cos/sin after editing arrays in place and drop the cached SOA view."""
        self.cos = np.cos(self.angle)
        self.sin = np.sin(self.angle)
        self._soa = None

    def subset(self, idx) -> "LensModel":
        idx = np.asarray(idx)
        return LensModel(self.type[idx], self.x[idx], self.y[idx], self.b0[idx], self.angle[idx],
                         self.epot[idx], self.rcore[idx], self.rcut[idx], self.z[idx],
                         self.dlsds[idx])

    @staticmethod
    def from_halos(halos: list[dict], zl: float = 0.4) -> "LensModel":
        """Build from halo dicts {type, x, y, sigma or b0, e or epot, theta, rcore, rcut}; stable
        sort by type (ascending), as the reference slots them."""
        order = sorted(range(len(halos)), key=lambda i: halos[i]["type"])
        hs = [halos[i] for i in order]
        t = np.array([h["type"] for h in hs], dtype=np.int32)
        b0 = np.array([h["b0"] if "b0" in h else b0_from_sigma(h["type"], h["sigma"]) for h in hs])
        ep = np.array([h["epot"] if "epot" in h else epot_from_e(h["type"], h["e"]) for h in hs])
        return LensModel(
            type=t,
            x=np.array([h["x"] for h in hs]), y=np.array([h["y"] for h in hs]), b0=b0,
            angle=np.array([h.get("theta", 0.0) for h in hs]), epot=ep,
            rcore=np.array([h.get("rcore", 0.0) for h in hs]),
            rcut=np.array([h.get("rcut", 0.0) for h in hs]),
            z=np.full(len(hs), zl), dlsds=np.array([h.get("dlsds", 1.0) for h in hs]))

    def halo_flops(self, kind: str) -> float:
        """Algorithmic flops per pixel of one map: sum over halos of F(type) (fixed convention:
        add/sub/mul 1, div 8, sqrt 8, log 40, atan2 60)."""
        table = {"gradient": {5: 41.0, 8: 195.0, 81: 355.0, 14: 0.0},
                 "hessian": {5: 49.0, 8: 140.0, 81: 262.0, 14: 30.0}}[kind]
        return float(sum(table[int(t)] for t in self.type))


# ------------------------------------------------------------------------------------------------
# synthetic workloads
# ------------------------------------------------------------------------------------------------

@dataclass
class Grid:
    """Pixel (col, row) sits at (xmin + (col + istart) dx, ymin + (row + jstart) dy); dx, dy default
    to (xmax - xmin)/(nx - 1), (ymax - ymin)/(ny - 1) as in src/grid_gradient_CPU.cpp:35-36."""
    frame: LtGridParam
    nx: int
    ny: int
    step: tuple = None

    @property
    def dx(self):
        if self.step is not None:
            return float(self.step[0])
        return (self.frame.xmax - self.frame.xmin) / (self.nx - 1)

    @property
    def dy(self):
        if self.step is not None:
            return float(self.step[1])
        return (self.frame.ymax - self.frame.ymin) / (self.ny - 1)

    @property
    def npix(self):
        return self.nx * self.ny


def group_lens(seed: int = 1) -> LensModel:
    """One cluster dPIE (sigma 1000 km/s, e ~ U[0,0.5], theta ~ U[0,pi), rcore 10", rcut 1000") at
    the origin plus ten SIS galaxies (sigma lognormal, median 150 km/s, 0.3 dex; positions uniform
    in the 200" field)."""
    rng = np.random.default_rng(seed)
    halos = [dict(type=81, x=0.0, y=0.0, sigma=1000.0, e=float(rng.uniform(0, 0.5)),
                  theta=float(rng.uniform(0, math.pi)), rcore=10.0, rcut=1000.0)]
    for _ in range(10):
        s = 150.0 * 10.0 ** (0.3 * rng.standard_normal())
        halos.append(dict(type=5, x=float(rng.uniform(-100, 100)), y=float(rng.uniform(-100, 100)),
                          sigma=float(s), e=0.0, theta=float(rng.uniform(0, math.pi)),
                          rcore=0.0, rcut=0.0))
    return LensModel.from_halos(halos)


def group_grid(n: int = 512) -> Grid:
    return Grid(grid_param(-100.0, 100.0, -100.0, 100.0), n, n)


def cluster_lens(seed: int = 2, ngal: int = 500) -> LensModel:
    """Two cluster dPIEs + one BCG dPIE (sigma 1000/800/300 km/s, rcore 10/8/0.5", rcut
    1000/1000/100", e ~ U[0,0.6]) plus `ngal` galaxy dPIEs scaled from magnitudes m ~ U[18,24]
    with mag0 = 18, slopes 4: sigma = 180 x 10^(0.4(18-m)/4), rcut = 40" x 10^(0.4(18-m)/2),
    rcore = 0.15" x 10^(0.4(18-m)/2) (src/module_readParameters.cpp:1603-1625)."""
    rng = np.random.default_rng(seed)
    halos = []
    centres = [(0.0, 0.0), (42.0, 28.0), (1.5, -0.8)]
    for (cx, cy), s, rc, rct in zip(centres, (1000.0, 800.0, 300.0), (10.0, 8.0, 0.5),
                                    (1000.0, 1000.0, 100.0)):
        halos.append(dict(type=81, x=cx, y=cy, sigma=s, e=float(rng.uniform(0, 0.6)),
                          theta=float(rng.uniform(0, math.pi)), rcore=rc, rcut=rct))
    mag0 = 18.0
    for _ in range(ngal):
        m = float(rng.uniform(18.0, 24.0))
        f = 0.4 * (mag0 - m)
        halos.append(dict(type=81, x=float(rng.uniform(-150, 150)), y=float(rng.uniform(-150, 150)),
                          sigma=180.0 * 10.0 ** (f / 4.0), e=float(rng.uniform(0, 0.6)),
                          theta=float(rng.uniform(0, math.pi)),
                          rcore=0.15 * 10.0 ** (f / 2.0), rcut=40.0 * 10.0 ** (f * 2.0 / 4.0)))
    return LensModel.from_halos(halos)


def cluster_grid(n: int = 4096) -> Grid:
    return Grid(grid_param(-150.0, 150.0, -150.0, 150.0), n, n)


def perturbed_samples(base: LensModel, nsamples: int, seed: int = 3) -> list[LensModel]:
    """Bayes-sample lens models: the cluster halo (the type-81 halo, first of its run) moves by
    N(0, 0.5") in x and y (5% of rcore = 10") and its e, theta, rcore, sigma get 5% Gaussian
    relative perturbations; galaxy sigma is scaled by a common N(1, 0.05) factor (rcut* has no
    effect on SIS galaxies)."""
    rng = np.random.default_rng(seed)
    out = []
    ci = int(np.nonzero(base.type == 81)[0][0])
    for _ in range(nsamples):
        m = base.subset(np.arange(base.n))
        m.x[ci] += 0.5 * rng.standard_normal()
        m.y[ci] += 0.5 * rng.standard_normal()
        # sigma enters b0 quadratically; epot follows e through the reference's relation
        e0 = _e_from_epot(81, base.epot[ci])
        e = min(max(e0 * (1.0 + 0.05 * rng.standard_normal()), 0.0), 0.95)
        m.epot[ci] = epot_from_e(81, e)
        m.angle[ci] = base.angle[ci] * (1.0 + 0.05 * rng.standard_normal())
        m.rcore[ci] = base.rcore[ci] * (1.0 + 0.05 * rng.standard_normal())
        m.b0[ci] = base.b0[ci] * (1.0 + 0.05 * rng.standard_normal()) ** 2
        gs = 1.0 + 0.05 * rng.standard_normal()
        gal = base.type == 5
        m.b0[gal] = base.b0[gal] * gs * gs
        m.refresh()
        out.append(m)
    return out


def _e_from_epot(t: int, epot: float) -> float:
    if t == 5:
        return 3.0 * epot
    return 2.0 * epot / (1.0 + epot * epot)


def galaxies_array(n: int) -> np.ndarray:
    return np.zeros(n, dtype=GALAXY_DTYPE)


def galaxy_ptr(a: np.ndarray):
    assert a.dtype == GALAXY_DTYPE and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(LtGalaxy))
