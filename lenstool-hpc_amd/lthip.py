"""ctypes binding of the lens-map C-ABI (include/lenstool_hip/lt_capi.h, liblenstool_hip.so).

This is the Python side of the drop-in boundary: every call goes to the HIP library; there is no
CPU fallback. If the shared library is missing the first call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import ltmodel as M

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblenstool_hip.so")

LT_OK = 0
LT_ERR_TYPE = 2
LT_ERR_OVERFLOW = 5
PREC_FP64 = 0
PREC_FP32C = 1
PX_HALO_LANES = 16
MAP_AMPLI, MAP_SHEAR, MAP_MASS, MAP_POT, MAP_DPL, MAP_FIELDS, MAP_HESSIAN, MAP_GRADIENT = range(1, 9)

_lib = None

_dp = C.POINTER(C.c_double)
_soa = C.POINTER(M.LtPotentialSOA)
_gp = C.POINTER(M.LtGridParam)
_cp = C.POINTER(M.LtCosmoParam)


class LtGridDesc(C.Structure):
    _fields_ = [("xmin", C.c_double), ("ymin", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int), ("istart", C.c_int), ("jstart", C.c_int),
                ("pitch", C.c_longlong)]


class LtRequest(C.Structure):
    _fields_ = [("model", C.c_int), ("kind", C.c_int), ("mode", C.c_int), ("precision", C.c_int),
                ("grid", LtGridDesc), ("r", C.c_double), ("mult", C.c_double)]


class LtRequestTimes(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("fetched_ms", C.c_double)]


class LtServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
                ("queued", C.c_longlong), ("running", C.c_longlong), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("p50_fetch_ms", C.c_double), ("p99_fetch_ms", C.c_double), ("halo_px", C.c_double),
                ("elapsed_ms", C.c_double), ("bytes_in_use", C.c_longlong), ("bytes_peak", C.c_longlong)]


class LtParInfo(C.Structure):
    _fields_ = [("nhalos", C.c_int), ("nhalos_main", C.c_int), ("npotfiles", C.c_int), ("nbgridcells", C.c_int),
                ("ampli", C.c_int), ("ampli_cells", C.c_int), ("ampli_z", C.c_double),
                ("mass", C.c_int), ("mass_cells", C.c_int), ("mass_z", C.c_double),
                ("shear", C.c_int), ("shear_cells", C.c_int), ("shear_z", C.c_double),
                ("poten", C.c_int), ("poten_cells", C.c_int), ("poten_z", C.c_double),
                ("dpl", C.c_int), ("dpl_cells", C.c_int), ("dpl_z", C.c_double),
                ("inverse", C.c_int), ("ref_mode", C.c_int), ("ref_ra", C.c_double), ("ref_dec", C.c_double),
                ("z_source", C.c_double), ("nsets", C.c_int), ("nimagestot", C.c_int)]


class LtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"lenstool_hip status {status}: {msg}")
        self.status = status


# every symbol include/lenstool_hip/*.h declares, with its ctypes signature
SIGNATURES = {
    "lt_last_error": ([], C.c_char_p),
    "lt_version": ([], C.c_int),
    "lt_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "lt_init": ([C.c_int], C.c_int),
    "lt_sync": ([], C.c_int),
    "lt_dev_alloc": ([C.POINTER(C.c_void_p), C.c_longlong], C.c_int),
    "lt_dev_free": ([C.c_void_p], C.c_int),
    "lt_memcpy_h2d": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p], C.c_int),
    "lt_memcpy_d2h": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p], C.c_int),
    "lt_memset_d": ([C.c_void_p, C.c_int, C.c_longlong, C.c_void_p], C.c_int),
    "lt_memcpy_d2d": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p], C.c_int),
    "lt_stream_create": ([C.POINTER(C.c_void_p)], C.c_int),
    "lt_stream_destroy": ([C.c_void_p], C.c_int),
    "lt_stream_sync": ([C.c_void_p], C.c_int),
    "lt_event_create": ([C.POINTER(C.c_void_p)], C.c_int),
    "lt_event_destroy": ([C.c_void_p], C.c_int),
    "lt_event_record": ([C.c_void_p, C.c_void_p], C.c_int),
    "lt_event_sync": ([C.c_void_p], C.c_int),
    "lt_event_elapsed": ([C.POINTER(C.c_float), C.c_void_p, C.c_void_p], C.c_int),
    "lt_gradient_grid": ([_dp, _dp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                          C.c_int, C.c_int], C.c_int),
    "lt_gradient2_grid": ([_dp, _dp, _dp, _dp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                           C.c_int, C.c_int, C.c_int], C.c_int),
    "lt_kmap_grid": ([_dp, _dp, _dp, _dp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                      C.c_int, C.c_int, C.c_int], C.c_int),
    "lt_potential_grid": ([_dp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                           C.c_int], C.c_int),
    "lt_map_grid": ([C.c_int, C.c_int, _dp, _dp, _cp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_double,
                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "lt_map_fields": ([_dp, _dp, _dp, _dp, _dp, _cp, _gp, _soa, C.c_int, C.c_double, C.c_double, C.c_double,
                       C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "lt_cosmo_observer_object": ([C.c_double, _cp], C.c_double),
    "lt_cosmo_object_object": ([C.c_double, C.c_double, _cp], C.c_double),
    "lt_cosmo_dlsds": ([C.c_double, C.c_double, _cp], C.c_double),
    "lt_model_create": ([C.POINTER(C.c_void_p), _soa, C.c_int, C.c_int], C.c_int),
    "lt_model_update": ([C.c_void_p, _soa, C.c_int, C.c_void_p], C.c_int),
    "lt_model_create_batch": ([C.POINTER(C.c_void_p), _soa, C.c_int, C.c_int, C.c_int], C.c_int),
    "lt_model_update_batch": ([C.c_void_p, _soa, C.c_int, C.c_void_p], C.c_int),
    "lt_model_stage": ([C.c_void_p, _soa, C.c_int], C.c_int),
    "lt_model_upload": ([C.c_void_p, C.c_void_p], C.c_int),
    "lt_server_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_longlong], C.c_int),
    "lt_server_add_model": ([C.c_void_p, _soa, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "lt_server_submit": ([C.c_void_p, C.POINTER(LtRequest), C.POINTER(C.c_longlong)], C.c_int),
    "lt_server_wait": ([C.c_void_p, C.c_longlong, C.c_double], C.c_int),
    "lt_server_fetch": ([C.c_void_p, C.c_longlong] + [C.c_void_p] * 5 + [C.POINTER(LtRequestTimes)], C.c_int),
    "lt_server_get_stats": ([C.c_void_p, C.POINTER(LtServerStats)], C.c_int),
    "lt_server_reset_stats": ([C.c_void_p], C.c_int),
    "lt_server_destroy": ([C.c_void_p], C.c_int),
    "lt_autotune": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(LtGridDesc), C.c_int, C.c_void_p,
                     C.POINTER(C.c_int), C.POINTER(C.c_float)], C.c_int),
    "lt_par_load": ([C.POINTER(C.c_void_p), C.c_char_p], C.c_int),
    "lt_par_get": ([C.c_void_p, C.POINTER(LtParInfo), C.POINTER(M.LtPotentialSOA), C.POINTER(M.LtGridParam),
                    C.POINTER(M.LtCosmoParam)], C.c_int),
    "lt_par_images": ([C.c_void_p, C.POINTER(C.c_int), C.POINTER(M.LtGalaxy)], C.c_int),
    "lt_par_bayes_load": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "lt_par_bayes_apply": ([C.c_void_p, C.c_int], C.c_int),
    "lt_par_map_name": ([C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "lt_par_cline": ([C.c_void_p, C.c_void_p], C.c_int),
    "lt_map_factors": ([C.c_int, C.c_int, C.POINTER(M.LtCosmoParam), C.c_double, C.c_double, C.c_double, C.c_double,
                        C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "lt_par_destroy": ([C.c_void_p], C.c_int),
    "lt_critic_device": ([C.c_void_p, C.c_void_p, C.POINTER(LtGridDesc), C.c_void_p, C.c_void_p, C.c_longlong,
                          C.POINTER(C.c_longlong), C.c_void_p], C.c_int),
    "lt_critic_device_noalloc": ([C.c_void_p, C.c_void_p, C.POINTER(LtGridDesc), C.c_void_p, C.c_void_p,
                                  C.c_longlong, C.POINTER(C.c_longlong), C.c_void_p], C.c_int),
    "lt_critic_lines": ([C.POINTER(M.LtCosmoParam), C.POINTER(M.LtGridParam), C.POINTER(M.LtPotentialSOA), C.c_int,
                         C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_longlong, C.POINTER(C.c_longlong)], C.c_int),
    "lt_fits_write": ([C.c_char_p, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                       C.c_double, C.c_int, C.c_double, C.c_double], C.c_int),
    "lt_graph_begin": ([C.c_void_p], C.c_int),
    "lt_graph_end": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "lt_graph_launch": ([C.c_void_p, C.c_void_p], C.c_int),
    "lt_graph_destroy": ([C.c_void_p], C.c_int),
    "lt_model_destroy": ([C.c_void_p], C.c_int),
    "lt_model_bytes": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "lt_map_device": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_double, C.c_double, C.POINTER(LtGridDesc), C.c_int, C.c_int,
                       C.c_void_p], C.c_int),
    "lt_point_gradients": ([_dp, _dp, _dp, _dp, C.c_int, _soa, C.c_int], C.c_int),
    "lt_delense_barycenter": ([C.POINTER(M.LtPoint), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int, _gp,
                               C.POINTER(M.LtGalaxy), _dp, _dp], C.c_int),
    "lt_chi_barycentersource": ([_dp, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, _soa, _gp, C.POINTER(C.c_int),
                                 C.POINTER(M.LtGalaxy)], C.c_int),
    "lt_chi_computation": ([_dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(M.LtPoint), C.POINTER(C.c_int),
                            C.POINTER(M.LtGalaxy), C.c_int], C.c_int),
    "lt_chi_batch_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, _gp, C.c_int,
                             C.POINTER(C.c_int), C.POINTER(M.LtGalaxy)], C.c_int),
    "lt_chi_batch_run": ([C.c_void_p, _soa, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p], C.c_int),
    "lt_chi_batch_images": ([C.c_void_p, C.c_int, C.POINTER(M.LtPoint), C.POINTER(C.c_int)], C.c_int),
    "lt_chi_batch_destroy": ([C.c_void_p], C.c_int),
    "lt_scale_device": ([C.c_void_p, C.c_double, C.c_longlong, C.c_void_p], C.c_int),
    "lt_map_epilogue_device": ([C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_double, C.c_double,
                                C.c_void_p], C.c_int),
}


def lib():
    """Load liblenstool_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C lenstool-hpc_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status):
    if status != LT_OK:
        raise LtError(status, lib().lt_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int(0)
    check(lib().lt_device_count(C.byref(n)))
    return n.value


# ------------------------------------------------------------------------------------------------
# synchronous host-array entry points
# ------------------------------------------------------------------------------------------------

def gradient_grid(lens: M.LensModel, grid: M.Grid, istart=0, jstart=0, precision=PREC_FP64):
    gx = np.empty(grid.npix)
    gy = np.empty(grid.npix)
    check(lib().lt_gradient_grid(M.dptr(gx), M.dptr(gy), C.byref(grid.frame), C.byref(lens.soa()), lens.n,
                                 grid.dx, grid.dy, grid.nx, grid.ny, istart, jstart, precision))
    s = (grid.ny, grid.nx)
    return gx.reshape(s), gy.reshape(s)


def hessian_grid(lens: M.LensModel, grid: M.Grid, istart=0, jstart=0, precision=PREC_FP64, kmap=False):
    planes = [np.empty(grid.npix) for _ in range(4)]
    f = lib().lt_kmap_grid if kmap else lib().lt_gradient2_grid
    check(f(*[M.dptr(p) for p in planes], C.byref(grid.frame), C.byref(lens.soa()), lens.n, grid.dx, grid.dy,
            grid.nx, grid.ny, istart, jstart, precision))
    return [p.reshape(grid.ny, grid.nx) for p in planes]


def potential_grid(lens: M.LensModel, grid: M.Grid, istart=0, jstart=0):
    phi = np.empty(grid.npix)
    check(lib().lt_potential_grid(M.dptr(phi), C.byref(grid.frame), C.byref(lens.soa()), lens.n, grid.dx,
                                  grid.dy, grid.nx, grid.ny, istart, jstart))
    return phi.reshape(grid.ny, grid.nx)


def map_grid(kind, mode, lens: M.LensModel, grid: M.Grid, cosmo, z_lens, z_source, istart=0, jstart=0,
             precision=PREC_FP64):
    m0 = np.empty(grid.npix)
    m1 = np.empty(grid.npix) if kind == MAP_DPL else None
    check(lib().lt_map_grid(kind, mode, M.dptr(m0), M.dptr(m1) if m1 is not None else None, C.byref(cosmo),
                            C.byref(grid.frame), C.byref(lens.soa()), lens.n, z_lens, z_source, grid.dx,
                            grid.dy, grid.nx, grid.ny, istart, jstart, precision))
    s = (grid.ny, grid.nx)
    return (m0.reshape(s), m1.reshape(s)) if m1 is not None else m0.reshape(s)


def map_fields(lens: M.LensModel, grid: M.Grid, cosmo, z_lens, z_source, istart=0, jstart=0,
               precision=PREC_FP64):
    outs = [np.empty(grid.npix) for _ in range(5)]
    check(lib().lt_map_fields(*[M.dptr(o) for o in outs], C.byref(cosmo), C.byref(grid.frame),
                              C.byref(lens.soa()), lens.n, z_lens, z_source, grid.dx, grid.dy, grid.nx, grid.ny,
                              istart, jstart, precision))
    return [o.reshape(grid.ny, grid.nx) for o in outs]


def point_gradients(lens: M.LensModel, px, py):
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    gx = np.empty_like(px)
    gy = np.empty_like(py)
    check(lib().lt_point_gradients(M.dptr(gx), M.dptr(gy), M.dptr(px), M.dptr(py), px.size, C.byref(lens.soa()),
                                   lens.n))
    return gx, gy


def delense_barycenter(nsets, n, frame, sources: np.ndarray, gx, gy):
    """Returns (status, image_pos[nsets, 20, 2], locimagesfound[nsets], numimg)."""
    pos = np.zeros((nsets, M.MAXIMPERSOURCE, 2))
    loc = np.zeros(nsets, dtype=np.int32)
    numimg = C.c_int(0)
    gx = np.ascontiguousarray(gx, dtype=np.float64).ravel()
    gy = np.ascontiguousarray(gy, dtype=np.float64).ravel()
    st = lib().lt_delense_barycenter(pos.ctypes.data_as(C.POINTER(M.LtPoint)), loc.ctypes.data_as(C.POINTER(C.c_int)),
                                     C.byref(numimg), nsets, n, C.byref(frame), M.galaxy_ptr(sources), M.dptr(gx),
                                     M.dptr(gy))
    if st not in (LT_OK, 5):
        check(st)
    return st, pos, loc, numimg.value


def chi_barycentersource(lens: M.LensModel, n, frame, nimages, images):
    nimages = np.ascontiguousarray(nimages, dtype=np.int32)
    chi = C.c_double(0.0)
    err = C.c_int(0)
    st = lib().lt_chi_barycentersource(C.byref(chi), C.byref(err), n, len(nimages), lens.n, C.byref(lens.soa()),
                                       C.byref(frame), nimages.ctypes.data_as(C.POINTER(C.c_int)),
                                       M.galaxy_ptr(images))
    return st, chi.value


def chi_computation(loc, pos, nimages, images):
    nimages = np.ascontiguousarray(nimages, dtype=np.int32)
    loc = np.ascontiguousarray(loc, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    chi = C.c_double(0.0)
    found = C.c_int(0)
    check(lib().lt_chi_computation(C.byref(chi), C.byref(found), loc.ctypes.data_as(C.POINTER(C.c_int)),
                                   pos.ctypes.data_as(C.POINTER(M.LtPoint)), nimages.ctypes.data_as(C.POINTER(C.c_int)),
                                   M.galaxy_ptr(images), len(nimages)))
    return chi.value, found.value


class ChiBatch:
    """chi^2 of many lens-model samples (same halo types) against one set of image families."""

    def __init__(self, nsamples, types, n, frame, nimages, images):
        self.h = C.c_void_p()
        self.nsamples = int(nsamples)
        self.nsets = len(nimages)
        self._keep = (np.ascontiguousarray(types, dtype=np.int32), np.ascontiguousarray(nimages, dtype=np.int32),
                      np.ascontiguousarray(images))
        t, ni, im = self._keep
        check(lib().lt_chi_batch_create(C.byref(self.h), self.nsamples, t.size, t.ctypes.data_as(C.POINTER(C.c_int)),
                                        n, C.byref(frame), ni.size, ni.ctypes.data_as(C.POINTER(C.c_int)),
                                        M.galaxy_ptr(im)))

    def run(self, lenses, stream=0):
        assert len(lenses) == self.nsamples
        arr = (M.LtPotentialSOA * self.nsamples)(*[l.soa() for l in lenses])
        chi = np.empty(self.nsamples)
        found = np.empty(self.nsamples, dtype=np.int32)
        status = np.empty(self.nsamples, dtype=np.int32)
        check(lib().lt_chi_batch_run(self.h, arr, M.dptr(chi), found.ctypes.data_as(C.POINTER(C.c_int)),
                                     status.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(stream)))
        return chi, found, status

    def images(self, sample):
        pos = np.zeros((self.nsets, M.MAXIMPERSOURCE, 2))
        loc = np.zeros(self.nsets, dtype=np.int32)
        check(lib().lt_chi_batch_images(self.h, sample, pos.ctypes.data_as(C.POINTER(M.LtPoint)),
                                        loc.ctypes.data_as(C.POINTER(C.c_int))))
        return pos, loc

    def close(self):
        if self.h:
            lib().lt_chi_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def map_factors(kind, mode, cosmo, z_lens, z_source, dx=1.0, dy=1.0):
    """(r, mult) arguments of Model.map for a reference map kind and mode (lt_map_factors)."""
    r, mult = C.c_double(0), C.c_double(0)
    check(lib().lt_map_factors(kind, mode, C.byref(cosmo), z_lens, z_source, dx, dy, C.byref(r), C.byref(mult)))
    return r.value, mult.value


def cosmo_dlsds(zl, zs, cosmo):
    return lib().lt_cosmo_dlsds(zl, zs, C.byref(cosmo))


# ------------------------------------------------------------------------------------------------
# device memory, streams, events
# ------------------------------------------------------------------------------------------------

class DeviceArray:
    """A device allocation of `n` elements of `dtype` (owned; freed on close / garbage collection)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().lt_dev_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    def upload(self, host, stream=0):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes <= self.nbytes
        check(lib().lt_memcpy_h2d(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), host.nbytes,
                                  C.c_void_p(stream)))

    def download(self, n=None, stream=0):
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=self.dtype)
        check(lib().lt_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes,
                                  C.c_void_p(stream)))
        if stream:
            sync(stream)
        return out

    def download_range(self, offset, n, stream=0):
        """Elements [offset, offset + n) to a new host array."""
        assert 0 <= offset and offset + n <= self.n
        out = np.empty(int(n), dtype=self.dtype)
        check(lib().lt_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset * out.itemsize),
                                  out.nbytes, C.c_void_p(stream)))
        if stream:
            sync(stream)
        return out

    def fill_bytes(self, value: int, stream=0):
        check(lib().lt_memset_d(C.c_void_p(self.ptr), int(value), self.nbytes, C.c_void_p(stream)))

    def close(self):
        if getattr(self, "ptr", None):
            lib().lt_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Launches captured on a stream between begin() and end(), replayed by launch()."""

    def __init__(self, stream):
        self.h = C.c_void_p()
        self.stream = stream
        check(lib().lt_graph_begin(C.c_void_p(stream)))

    def end(self):
        check(lib().lt_graph_end(C.c_void_p(self.stream), C.byref(self.h)))
        return self

    def launch(self, stream=None):
        check(lib().lt_graph_launch(self.h, C.c_void_p(self.stream if stream is None else stream)))

    def close(self):
        if self.h:
            lib().lt_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Map-serving loop (lt_server_*): resident models, requests admitted by HBM footprint."""

    PLANES = {MAP_HESSIAN: 4, MAP_FIELDS: 5, MAP_GRADIENT: 2, MAP_DPL: 2}

    def __init__(self, nslots=4, hbm_budget_bytes=32 << 30):
        self.h = C.c_void_p()
        check(lib().lt_server_create(C.byref(self.h), nslots, hbm_budget_bytes))
        self._lenses = []

    def add_model(self, lens: M.LensModel) -> int:
        mid = C.c_int()
        check(lib().lt_server_add_model(self.h, C.byref(lens.soa()), lens.n, C.byref(mid)))
        self._lenses.append(lens)
        return mid.value

    def submit(self, model, kind, mode, grid: M.Grid, precision=PREC_FP64, r=1.0, mult=1.0) -> int:
        g = LtGridDesc(grid.frame.xmin, grid.frame.ymin, grid.dx, grid.dy, grid.nx, grid.ny, 0, 0, grid.nx)
        req = LtRequest(model, kind, mode, precision, g, r, mult)
        t = C.c_longlong()
        check(lib().lt_server_submit(self.h, C.byref(req), C.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms=-1.0) -> int:
        return lib().lt_server_wait(self.h, ticket, timeout_ms)

    def fetch(self, ticket, kind, npix, into=None):
        """Copy the result planes to host arrays (or into the given ones); returns (planes, times)."""
        n = self.PLANES.get(kind, 1)
        outs = into if into is not None else [np.empty(npix) for _ in range(n)]
        ptrs = [o.ctypes.data_as(C.c_void_p) if o is not None else None for o in outs] + [None] * (5 - len(outs))
        tm = LtRequestTimes()
        check(lib().lt_server_fetch(self.h, ticket, *ptrs, C.byref(tm)))
        return outs, tm

    def stats(self) -> dict:
        st = LtServerStats()
        check(lib().lt_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in LtServerStats._fields_}

    def reset_stats(self):
        check(lib().lt_server_reset_stats(self.h))

    def close(self):
        if self.h:
            lib().lt_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LtClineInfo(C.Structure):
    _fields_ = [("nplan", C.c_int), ("z", C.c_double * 9), ("dmax", C.c_double), ("step", C.c_double),
                ("limit_high", C.c_double), ("nbgridcells", C.c_int)]


class ParFile:
    """A Lenstool .par read by the library (lt_par_*): run settings, lens model, frame, cosmology,
    images, and the bayes.dat samples of the model (`bayes_load`, `bayes_apply`, in place like the
    reference's setbayesmapmodels)."""

    def __init__(self, path):
        h = C.c_void_p()
        check(lib().lt_par_load(C.byref(h), str(path).encode()))
        self.h = h
        self.nsamples = 0

    def get(self):
        """dict with `info`, `lens` (LensModel, halos in the reference's SOA order), `extra` (vdisp,
        ellipticity, mag, anglecos, anglesin), `frame`, `cosmo`, `nimages`, `images` (GALAXY_DTYPE)."""
        info, soa, frame, cosmo = LtParInfo(), M.LtPotentialSOA(), M.LtGridParam(), M.LtCosmoParam()
        check(lib().lt_par_get(self.h, C.byref(info), C.byref(soa), C.byref(frame), C.byref(cosmo)))
        n = info.nhalos

        def arr(ptr, dtype=np.float64):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype).copy() if n else np.zeros(0, dtype)
        lens = M.LensModel(type=arr(soa.type, np.int32), x=arr(soa.position_x), y=arr(soa.position_y),
                           b0=arr(soa.b0), angle=arr(soa.ellipticity_angle), epot=arr(soa.ellipticity_potential),
                           rcore=arr(soa.rcore), rcut=arr(soa.rcut), z=arr(soa.z), dlsds=arr(soa.dlsds))
        extra = {"vdisp": arr(soa.vdisp), "ellipticity": arr(soa.ellipticity), "mag": arr(soa.mag),
                 "anglecos": arr(soa.anglecos), "anglesin": arr(soa.anglesin)}
        nimages = np.zeros(info.nsets, dtype=np.int32)
        images = M.galaxies_array(info.nimagestot)
        check(lib().lt_par_images(self.h, nimages.ctypes.data_as(C.POINTER(C.c_int)), M.galaxy_ptr(images)))
        return {"info": {f: getattr(info, f) for f, _ in LtParInfo._fields_}, "lens": lens, "extra": extra,
                "frame": frame, "cosmo": cosmo, "nimages": nimages, "images": images}

    def bayes_load(self, path):
        """Read a bayes.dat; returns (samples, columns)."""
        ns, nc = C.c_int(0), C.c_int(0)
        check(lib().lt_par_bayes_load(self.h, str(path).encode(), C.byref(ns), C.byref(nc)))
        self.nsamples = ns.value
        return ns.value, nc.value

    def bayes_apply(self, sample):
        check(lib().lt_par_bayes_apply(self.h, int(sample)))

    def map_name(self, kind, which=0):
        """File name of the runmode line of a map kind ('' if absent)."""
        buf = C.create_string_buffer(4096)
        check(lib().lt_par_map_name(self.h, kind, which, buf, 4096))
        return buf.value.decode()

    def cline(self):
        """The `cline` block: nplan, z (list), dmax, step, limit_high, nbgridcells."""
        k = LtClineInfo()
        check(lib().lt_par_cline(self.h, C.byref(k)))
        return {"nplan": k.nplan, "z": list(k.z), "dmax": k.dmax, "step": k.step, "limit_high": k.limit_high,
                "nbgridcells": k.nbgridcells}

    def bayes_models(self, samples=None):
        """LensModel of each sample, applied in order (as the reference's map loop does)."""
        out = []
        for k in (range(self.nsamples) if samples is None else samples):
            self.bayes_apply(k)
            out.append(self.get()["lens"])
        return out

    def close(self):
        if self.h:
            lib().lt_par_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def critic_device(det_ptr, kappa_ptr, grid: M.Grid, max_segments, stream=0, istart=0, jstart=0, pitch=None,
                  noalloc=False):
    """Contour det A = 0 of device planes; returns (segments (n, 4), kinds or None, total, status).
    noalloc: through lt_critic_device_noalloc (fixed scratch array) instead of lt_critic_device."""
    g = LtGridDesc(grid.frame.xmin, grid.frame.ymin, grid.dx, grid.dy, grid.nx, grid.ny, istart, jstart,
                   pitch or grid.nx)
    total = C.c_longlong(0)
    room = int(max_segments)
    # room 0: count only (no output arrays)
    dsegs = DeviceArray(4 * room) if room > 0 else None
    dkinds = DeviceArray(room, dtype=np.int32) if (kappa_ptr and room > 0) else None
    try:
        fn = lib().lt_critic_device_noalloc if noalloc else lib().lt_critic_device
        st = fn(C.c_void_p(det_ptr), C.c_void_p(kappa_ptr) if dkinds is not None else None,
                                    C.byref(g), C.c_void_p(dsegs.ptr) if dsegs is not None else None,
                                    C.c_void_p(dkinds.ptr) if dkinds is not None else None,
                                    room, C.byref(total), C.c_void_p(stream))
        if st not in (LT_OK, LT_ERR_OVERFLOW):
            check(st)
        got = min(total.value, room)
        segs = dsegs.download(4 * got).reshape(got, 4) if got > 0 else np.zeros((0, 4))
        kinds = (dkinds.download(got) if got > 0 else np.zeros(0, np.int32)) if dkinds is not None else None
        return segs, kinds, total.value, st
    finally:
        if dsegs is not None:
            dsegs.close()
        if dkinds is not None:
            dkinds.close()


def critic_lines(lens: M.LensModel, frame, nx, ny, cosmo, z_lens, z_source, precision=PREC_FP64,
                 max_segments=1 << 20, caustics=True):
    """Critical lines (n, 4), caustics (n, 4) or None, kinds (n,) of a lens; total segment count.
    A model with external-shear (type 14) halos gets its lines and NaN caustics (status LT_ERR_TYPE
    from the library, not raised here)."""
    crit = np.zeros((max_segments, 4))
    caus = np.zeros((max_segments, 4)) if caustics else None
    kinds = np.zeros(max_segments, dtype=np.int32)
    total = C.c_longlong(0)
    st = lib().lt_critic_lines(C.byref(cosmo), C.byref(frame), C.byref(lens.soa()), lens.n, z_lens, z_source, nx, ny,
                               precision, M.dptr(crit), M.dptr(caus) if caus is not None else None,
                               kinds.ctypes.data_as(C.POINTER(C.c_int)), max_segments, C.byref(total))
    if st not in (LT_OK, LT_ERR_OVERFLOW, LT_ERR_TYPE):
        check(st)
    got = min(total.value, max_segments)
    return crit[:got], (caus[:got] if caus is not None else None), kinds[:got], total.value


def fits_write(path, image, xmin, xmax, ymin, ymax, ra=0.0, dec=0.0, absolute=None):
    """Write a 2-D map (rows = y) as a float32 FITS image with the frame's coordinates; absolute
    RA/DEC axes unless ra == dec == 0 (as the reference's module_writeFits picks)."""
    img = np.ascontiguousarray(image, dtype=np.float64)
    ny, nx = img.shape
    if absolute is None:
        absolute = not (ra == 0.0 and dec == 0.0)
    check(lib().lt_fits_write(str(path).encode(), img.ctypes.data_as(C.POINTER(C.c_double)), nx, ny,
                              xmin, xmax, ymin, ymax, int(bool(absolute)), ra, dec))


def load_par(path):
    """Read a Lenstool .par (and its potfiles / image file): ParFile(path).get()."""
    with ParFile(path) as p:
        return p.get()


def stream_create() -> int:
    s = C.c_void_p()
    check(lib().lt_stream_create(C.byref(s)))
    return s.value


def stream_destroy(s: int):
    check(lib().lt_stream_destroy(C.c_void_p(s)))


def sync(stream=0):
    check(lib().lt_stream_sync(C.c_void_p(stream)))


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().lt_event_create(C.byref(e)))
        self.h = e.value

    def record(self, stream=0):
        check(lib().lt_event_record(C.c_void_p(self.h), C.c_void_p(stream)))

    def synchronize(self):
        check(lib().lt_event_sync(C.c_void_p(self.h)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        check(lib().lt_event_elapsed(C.byref(ms), C.c_void_p(self.h), C.c_void_p(end.h)))
        return ms.value

    def __del__(self):
        try:
            lib().lt_event_destroy(C.c_void_p(self.h))
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# resident models (device pointers are plain integers, e.g. torch.Tensor.data_ptr())
# ------------------------------------------------------------------------------------------------

class Model:
    """A resident lens model, or a batch of samples of one (lenses = list of LensModel)."""

    def __init__(self, lens, use=MAP_HESSIAN):
        self.h = C.c_void_p()
        self.lens = lens
        if isinstance(lens, (list, tuple)):
            arr = (M.LtPotentialSOA * len(lens))(*[l.soa() for l in lens])
            check(lib().lt_model_create_batch(C.byref(self.h), arr, len(lens), lens[0].n, use))
            self.nsamples = len(lens)
        else:
            check(lib().lt_model_create(C.byref(self.h), C.byref(lens.soa()), lens.n, use))
            self.nsamples = 1

    def update_batch(self, lenses, stream=0):
        arr = (M.LtPotentialSOA * len(lenses))(*[l.soa() for l in lenses])
        check(lib().lt_model_update_batch(self.h, arr, len(lenses), C.c_void_p(stream)))
        self.lens = list(lenses)

    @staticmethod
    def soa_array(lenses):
        """The lt_potential_soa array of a list of lens models, for stage() (built once, it can be
        staged many times without rebuilding it)."""
        lenses = lenses if isinstance(lenses, (list, tuple)) else [lenses]
        return (M.LtPotentialSOA * len(lenses))(*[l.soa() for l in lenses])

    def stage(self, lenses):
        """Pack new sample parameters into the staging buffer: a list of lens models, or an array from
        soa_array() (the lens models it was built from must stay alive)."""
        arr = lenses if isinstance(lenses, C.Array) else self.soa_array(lenses)
        check(lib().lt_model_stage(self.h, arr, len(arr)))

    def upload(self, stream=0):
        check(lib().lt_model_upload(self.h, C.c_void_p(stream)))

    def update(self, lens: M.LensModel, stream=0):
        self.lens = lens
        check(lib().lt_model_update(self.h, C.byref(lens.soa()), lens.n, C.c_void_p(stream)))

    def bytes(self) -> int:
        b = C.c_longlong(0)
        check(lib().lt_model_bytes(self.h, C.byref(b)))
        return b.value

    def map(self, kind, mode, outs, grid: M.Grid, r=1.0, mult=1.0, precision=PREC_FP64, px=0, stream=0,
            istart=0, jstart=0, pitch=None):
        g = LtGridDesc(grid.frame.xmin, grid.frame.ymin, grid.dx, grid.dy, grid.nx, grid.ny, istart, jstart,
                       pitch or grid.nx)
        ptrs = [C.c_void_p(p) if p else None for p in (list(outs) + [0] * 5)[:5]]
        check(lib().lt_map_device(self.h, kind, mode, *ptrs, r, mult, C.byref(g), precision, px,
                                  C.c_void_p(stream)))

    def autotune(self, kind, mode, grid: M.Grid, precision=PREC_FP64, stream=0):
        """Pick the fastest pixels-per-thread for this map shape; returns (best_px, ms for px 1, 2, 4)."""
        g = LtGridDesc(grid.frame.xmin, grid.frame.ymin, grid.dx, grid.dy, grid.nx, grid.ny, 0, 0, grid.nx)
        best = C.c_int()
        ms = (C.c_float * 3)()
        check(lib().lt_autotune(self.h, kind, mode, C.byref(g), precision, C.c_void_p(stream), C.byref(best), ms))
        return best.value, list(ms)

    def close(self):
        if self.h:
            lib().lt_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
