"""ctypes binding of libbifhip.so (the gfx950 HIP library).

Mirrors the reference's host API (Library/hostLibrary.cuh) one to one:
the reference-named functions below take the same arguments in the same
order and raise BifhipError where the reference would print and exit().
Device buffers are passed as integer addresses (DeviceArray.ptr, memory
from the library's own HIP runtime).  There is no CPU fallback: a missing
library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbifhip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "bifhip.h")

MODEL_ROSSLER = 0
MODEL_LINEAR = 1
MODEL_MEMRISTIVE = 2
PREC_FP64 = 0
PREC_FP32 = 1
READING_CLASS_AND_CHAOS, READING_CHAOS, READING_CLASS = 0, 1, 2
PERT_REFERENCE = 0
PERT_AXIS_X = 1

E_ARG, E_FILE, E_NODEV, E_LIMIT = -1, -2, -3, -4


class BifhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bifhip error {code}: {msg}")
        self.code = code


_D, _I, _LL = C.c_double, C.c_int, C.c_longlong
_PD, _PI, _VP = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
_lib = None


def load():
    """Load libbifhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with "
                "python cudabifurcation3.0_amd/build.py")
        lib = C.CDLL(LIB_PATH)
        _declare(lib)
        _lib = lib
    return _lib


def _declare(L):
    drv_bif1 = [_D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _D, _PD, _I, _I]
    drv_kde = [_D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _I, _D, _PD, _I, _I, _I, _D, _D, _D]
    drv_db = [_D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _I, _D, _PD, _I, _I, _D]
    drv_ly = [_D, _D, _I, _D, _D, _PD, _I, _PD, _PI, _I, _D, _D, _PD, _I]
    sig = {
        "bifhip_set_output_path": [C.c_char_p],
        "bifhip_set_model": [_I], "bifhip_set_precision": [_I],
        "bifhip_set_perturbation": [_I], "bifhip_set_direction_chunk": [_LL],
        "bifhip_set_max_cells_per_launch": [_LL], "bifhip_set_peak_capacity": [_I],
        "bifhip_set_cluster_clamp": [_I], "bifhip_set_verbose": [_I],
        "bifhip_set_timing": [_I],
        "bifhip_kernel_time": [C.c_char_p, _PD, C.POINTER(_LL)],
        "bifhip_kernel_cells": [C.c_char_p, _PD],
        "bifhip_release": [], "bifhip_device_count": [],
        "bifhip_bifurcation1D": drv_bif1,
        "bifhip_bifurcation2DKDE": drv_kde,
        "bifhip_bifurcation2DDBSCAN": drv_db,
        "bifhip_bifurcation1DIC": drv_bif1,
        "bifhip_bifurcation2DKDEIC": drv_kde,
        "bifhip_bifurcation2DDBSCANIC": drv_db,
        "bifhip_LLE1DIC": drv_ly, "bifhip_LLE2DIC": drv_ly,
        "bifhip_set_sweep": [_I],
        "bifhip_calculationOfPeriodicityByOstrovsky": drv_db + [_D],
        "bifhip_ostrovsky_to_buffer": drv_db + [_D, _PD],
        "bifhip_modelingOneSystemDenis": [_D, _D, _D, _I, _PD, _I, _PD, _I],
        "bifhip_denis_to_buffer": [_D, _D, _D, _I, _PD, _I, _PD, _I, _PD, _LL],
        "bifhip_LLE1D": drv_ly, "bifhip_LLE2D": drv_ly,
        "bifhip_LS1D": drv_ly, "bifhip_LS2D": drv_ly,
        "bifhip_bifurcation1D_to_buffer": drv_bif1 + [_PI, _PD, _I],
        "bifhip_bifurcation2DKDE_to_buffer": drv_kde + [_PI],
        "bifhip_bifurcation2DDBSCAN_to_buffer": drv_db + [_PI],
        "bifhip_LLE_to_buffer": [_I] + drv_ly + [_PD],
        "bifhip_LS_to_buffer": [_I] + drv_ly + [_PD],
        "bifhip_bifurcation2DDBSCAN_device": drv_db + [_VP],
        "bifhip_bifurcation2DKDE_device": drv_kde + [_VP],
        "bifhip_LLE_device": [_I] + drv_ly + [_VP],
        "bifhip_peak_finder_device": [_VP, _I, _I, _VP, _VP, _VP, _D],
        "bifhip_dbscan_device": [_VP, _VP, _I, _I, _VP, _I, _D, _VP],
        "bifhip_dbscan_device_noalloc": [_VP, _VP, _I, _I, _VP, _I, _D, _VP],
        "bifhip_kde_device": [_VP, _I, _I, _VP, _I, _I, _D, _D, _D, _VP],
        "bifhip_gram_schmidt_device": [_VP, _VP, _VP, _I, _I],
        "bifhip_series_device": [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _D, _PD, _I,
                                 _I, _VP, _VP],
        "bifhip_write_map_int": [C.c_char_p, _I, _PD, _I, _I, _PI, _LL],
        "bifhip_write_map_double": [C.c_char_p, _I, _PD, _I, _I, _PD, _LL, _I, _I],
        "bifhip_write_bif1d": [C.c_char_p, _I, _PD, _PI, _PD, _I],
        "bifhip_write_1d_values": [C.c_char_p, _I, _PD, _PD, _I],
        "bifhip_device_malloc": [C.POINTER(_VP), C.c_size_t],
        "bifhip_device_free": [_VP],
        "bifhip_memcpy_h2d": [_VP, _VP, C.c_size_t],
        "bifhip_memcpy_d2h": [_VP, _VP, C.c_size_t],
        "bifhip_memcpy_d2d": [_VP, _VP, C.c_size_t],
        "bifhip_device_memset": [_VP, _I, C.c_size_t],
        "bifhip_synchronize": [],
        "bifhip_tiles_generate": [_I, _I, _I, _I, C.c_ulonglong, C.c_ulonglong, _D, _D, _D,
                                  _PI, _PI, _PI],
        "bifhip_ls_tiles_device": [_I, _I, _I, _PI, _PI, _PI, _I, _PD, _PI, _PD, _I, _PD, _I,
                                   _D, _D, _D, _VP],
        "bifhip_ls_tiles_autotune": [_I, _I, _I, _PI, _PI, _PI, _I, _PD, _PI, _PD, _I, _PD,
                                     _I, _D, _D, _D, _VP],
        "bifhip_set_tile_config": [_I, _I, _I],
        "bifhip_get_tile_config": [_PI, _PI, _PI],
        "bifhip_set_tile_lanes": [_I],
        "bifhip_get_tile_lanes": [_PI],
        "bifhip_tile_autotune_log": [_PD, _I],
        "bifhip_set_tile_precision": [_I],
        "bifhip_set_graph": [_I],
        "bifhip_set_partial_resolve": [_I],
        "bifhip_set_chaos_confidence": [_D],
        "bifhip_set_lle32_margin": [_D],
        "bifhip_set_regime_reading": [_I],
        "bifhip_set_quad_confidence": [_D],
        "bifhip_lle_peak_map_device": [_I, _D, _I, _PD, _PD, _PI, _PD, _I, _D, _D, _I, _I, _D,
                                       _D, _D, _D, _D, _VP, _VP],
        "bifhip_surrogate_map_device": [_I, _I, _D, _I, _PD, _PD, _PI, _PD, _I, _D, _D, _I,
                                        _I, _D, _D, _D, _D, _D, _I, _I, _D, C.c_ulonglong,
                                        _D, _D, _D, _VP, _VP, _VP, _PD],
        "bifhip_mlp_param_count": [],
        "bifhip_mlp_train_device": [_I, _VP, _VP, _I, _I, _D, _D, C.c_ulonglong, _D, _D, _I,
                                    C.POINTER(C.c_float)],
        "bifhip_mlp_infer_device": [_I, _D, _D, _D, _VP, _VP, _VP, _VP, _VP],
        "bifhip_mlp_get_params": [C.POINTER(C.c_float), _I, C.POINTER(C.c_float), _I],
        "bifhip_mlp_adam": [_I, C.POINTER(C.c_float), C.POINTER(C.c_float),
                            C.POINTER(C.c_float), C.POINTER(C.c_float), _I, _I,
                            C.POINTER(C.c_ubyte), _D, _D, _D, _D, C.POINTER(C.c_ushort)],
        "bifhip_mlp_gradient": [_I, C.POINTER(C.c_float), C.POINTER(C.c_float), _PI, _PI,
                                C.POINTER(C.c_float), _I, _D, _D, C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.POINTER(C.c_float)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _I
    L.bifhip_get_output_path.restype = C.c_char_p
    L.bifhip_last_error.restype = C.c_char_p
    L.bifhip_version.restype = C.c_char_p
    L.bifhip_value_by_idx.argtypes = [_I, _I, _D, _D, _I]
    L.bifhip_value_by_idx.restype = _D


def _check(rc):
    if rc != 0:
        raise BifhipError(rc, load().bifhip_last_error().decode(errors="replace"))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


# ----------------------------------------------------------- device memory

class DeviceArray:
    """A device buffer on the library's own HIP runtime (numpy dtype/shape)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(np.asarray(shape, dtype=np.int64)).tolist()) \
            if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(load().bifhip_device_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        _check(load().bifhip_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    @classmethod
    def full(cls, shape, value, dtype):
        return cls.from_host(np.full(shape, value, dtype=dtype))

    def to_host(self):
        a = np.empty(self.shape, dtype=self.dtype)
        _check(load().bifhip_memcpy_d2h(a.ctypes.data, self.ptr, self.nbytes))
        return a

    def zero_(self):
        _check(load().bifhip_device_memset(self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if getattr(self, "ptr", None):
            load().bifhip_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(load().bifhip_synchronize())


# ------------------------------------------------------------ context

def set_output_path(p):
    _check(load().bifhip_set_output_path(str(p).encode()))


def set_model(m):
    _check(load().bifhip_set_model(m))


def set_precision(p):
    _check(load().bifhip_set_precision(p))


def set_perturbation(m):
    _check(load().bifhip_set_perturbation(m))


def set_direction_chunk(n):
    _check(load().bifhip_set_direction_chunk(n))


def set_max_cells_per_launch(n):
    _check(load().bifhip_set_max_cells_per_launch(n))


def set_peak_capacity(n):
    _check(load().bifhip_set_peak_capacity(n))


def set_cluster_clamp(n):
    _check(load().bifhip_set_cluster_clamp(n))


def set_timing(on):
    _check(load().bifhip_set_timing(1 if on else 0))


def kernel_time(name):
    ms = C.c_double(0)
    n = C.c_longlong(0)
    _check(load().bifhip_kernel_time(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


def kernel_cells(name):
    """Cells covered by the timed launches of a kernel family."""
    c = C.c_double(0)
    _check(load().bifhip_kernel_cells(name.encode(), C.byref(c)))
    return c.value


def reset_kernel_times():
    _check(load().bifhip_kernel_time(None, None, None))


def release():
    _check(load().bifhip_release())


def reset_context():
    """Restore every knob to its default."""
    set_model(MODEL_ROSSLER)
    set_precision(PREC_FP64)
    set_perturbation(PERT_REFERENCE)
    set_direction_chunk(-1)
    set_max_cells_per_launch(0)
    set_peak_capacity(256)
    set_cluster_clamp(0)
    set_timing(False)
    set_sweep(SWEEP_PARAMETERS)
    set_graph(True)
    set_partial_resolve(True)
    set_chaos_confidence(0.975)
    set_lle32_margin(0.02)
    set_regime_reading(READING_CLASS_AND_CHAOS)
    set_quad_confidence(0.0)
    set_tile_precision(PREC_FP32)
    set_tile_lanes(1)


def device_count():
    return load().bifhip_device_count()


def value_by_idx(idx, nPts, s, f, k):
    return load().bifhip_value_by_idx(idx, nPts, s, f, k)


# ---------------------------------------- reference-named drivers (files)

def _bif1(fn, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
          indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
          preScaller):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(fn(tMax, nPts, h, amountOfInitialConditions, _pd(ic), _pd(r), _pi(m), writableVar,
              maxValue, transientTime, _pd(v), amountOfValues, preScaller))


def _kde2(fn, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
          indicesOfMutVars, writableVar, maxValue, maxAmountOfPeaks, transientTime, values,
          amountOfValues, preScaller, kdeSampling, kdeSamplesInterval1, kdeSamplesInterval2,
          kdeSamplesSmooth):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(fn(tMax, nPts, h, amountOfInitialConditions, _pd(ic), _pd(r), _pi(m), writableVar,
              maxValue, maxAmountOfPeaks, transientTime, _pd(v), amountOfValues, preScaller,
              kdeSampling, kdeSamplesInterval1, kdeSamplesInterval2, kdeSamplesSmooth))


def _db2(fn, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
         indicesOfMutVars, writableVar, maxValue, maxAmountOfPeaks, transientTime, values,
         amountOfValues, preScaller, eps):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(fn(tMax, nPts, h, amountOfInitialConditions, _pd(ic), _pd(r), _pi(m), writableVar,
              maxValue, maxAmountOfPeaks, transientTime, _pd(v), amountOfValues, preScaller, eps))


def bifurcation1D(*a):
    _bif1(load().bifhip_bifurcation1D, *a)


def bifurcation2DKDE(*a):
    _kde2(load().bifhip_bifurcation2DKDE, *a)


def bifurcation2DDBSCAN(*a):
    _db2(load().bifhip_bifurcation2DDBSCAN, *a)


def bifurcation1DIC(*a):
    _bif1(load().bifhip_bifurcation1DIC, *a)


def bifurcation2DKDEIC(*a):
    _kde2(load().bifhip_bifurcation2DKDEIC, *a)


def bifurcation2DDBSCANIC(*a):
    _db2(load().bifhip_bifurcation2DDBSCANIC, *a)


def calculationOfPeriodicityByOstrovsky(tMax, nPts, h, amountOfInitialConditions,
                                       initialConditions, ranges, indicesOfMutVars, writableVar,
                                       maxValue, maxAmountOfPeaks, transientTime, values,
                                       amountOfValues, preScaller, eps, ostrovskyThreshold):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_calculationOfPeriodicityByOstrovsky(
        tMax, nPts, h, amountOfInitialConditions, _pd(ic), _pd(r), _pi(m), writableVar, maxValue,
        maxAmountOfPeaks, transientTime, _pd(v), amountOfValues, preScaller, eps,
        ostrovskyThreshold))


def ostrovsky_map(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar,
                  maxValue, maxAmountOfPeaks, transientTime, values, preScaller, eps, threshold):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros(nPts * nPts)
    _check(load().bifhip_ostrovsky_to_buffer(tMax, nPts, h, len(ic), _pd(ic), _pd(r), _pi(m),
                                             writableVar, maxValue, maxAmountOfPeaks,
                                             transientTime, _pd(v), len(v), preScaller, eps,
                                             threshold, _pd(out)))
    return out


def modelingOneSystemDenis(tMax, h, hSpecial, amountOfInitialConditions, initialConditions,
                           writableVar, values, amountOfValues):
    ic, v = _f64(initialConditions), _f64(values)
    _check(load().bifhip_modelingOneSystemDenis(tMax, h, hSpecial, amountOfInitialConditions,
                                                _pd(ic), writableVar, _pd(v), amountOfValues))


def denis_series(tMax, h, hSpecial, initialConditions, writableVar, values):
    ic, v = _f64(initialConditions), _f64(values)
    out = np.zeros(max(int(tMax / h), 1))
    _check(load().bifhip_denis_to_buffer(tMax, h, hSpecial, len(ic), _pd(ic), writableVar,
                                         _pd(v), len(v), _pd(out), len(out)))
    return out[:int(tMax / h)]


SWEEP_PARAMETERS, SWEEP_INITIAL_CONDITIONS = 0, 1


def set_sweep(target):
    """Sweep target of the buffer and device entry points."""
    _check(load().bifhip_set_sweep(int(target)))


def _lyap_file(fn, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
               indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(fn(tMax, NT, nPts, h, eps, _pd(ic), amountOfInitialConditions, _pd(r), _pi(m),
              writableVar, maxValue, transientTime, _pd(v), amountOfValues))


def LLE1D(*a):
    _lyap_file(load().bifhip_LLE1D, *a)


def LLE2D(*a):
    _lyap_file(load().bifhip_LLE2D, *a)


def LS1D(*a):
    _lyap_file(load().bifhip_LS1D, *a)


def LS2D(*a):
    _lyap_file(load().bifhip_LS2D, *a)


def LLE1DIC(*a):
    _lyap_file(load().bifhip_LLE1DIC, *a)


def LLE2DIC(*a):
    _lyap_file(load().bifhip_LLE2DIC, *a)


# ------------------------------------------------ in-memory (host) twins

def bifurcation1D_to_buffer(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars,
                            writableVar, maxValue, transientTime, values, preScaller, stride=None):
    """Counts and peaks per point; peaks[k, :counts[k]].  Without a stride,
    rows of 1024 slots are tried first; when a point has more peaks the
    library refuses them with every count filled and keeps the diagram, and
    the retry with the stride max(counts) copies it (no second
    integration)."""
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    st = stride or 1024
    for attempt in range(2):
        counts = np.zeros(nPts, dtype=np.int32)
        peaks = np.zeros(nPts * st)
        rc = load().bifhip_bifurcation1D_to_buffer(tMax, nPts, h, len(ic), _pd(ic), _pd(r),
                                                   _pi(m), writableVar, maxValue, transientTime,
                                                   _pd(v), len(v), preScaller, _pi(counts),
                                                   _pd(peaks), st)
        if rc == E_LIMIT and not stride and attempt == 0:
            st = int(counts.max())
            continue
        _check(rc)
        return counts, peaks.reshape(nPts, st)


def dbscan_map(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue,
               maxAmountOfPeaks, transientTime, values, preScaller, eps):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros(nPts * nPts, dtype=np.int32)
    _check(load().bifhip_bifurcation2DDBSCAN_to_buffer(tMax, nPts, h, len(ic), _pd(ic), _pd(r),
                                                       _pi(m), writableVar, maxValue,
                                                       maxAmountOfPeaks, transientTime, _pd(v),
                                                       len(v), preScaller, eps, _pi(out)))
    return out


def kde_map(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars, writableVar, maxValue,
            maxAmountOfPeaks, transientTime, values, preScaller, kdeSampling, i1, i2, smooth):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    out = np.zeros(nPts * nPts, dtype=np.int32)
    _check(load().bifhip_bifurcation2DKDE_to_buffer(tMax, nPts, h, len(ic), _pd(ic), _pd(r),
                                                    _pi(m), writableVar, maxValue,
                                                    maxAmountOfPeaks, transientTime, _pd(v),
                                                    len(v), preScaller, kdeSampling, i1, i2,
                                                    smooth, _pi(out)))
    return out


def lle_map(dimension, tMax, NT, nPts, h, eps, initialConditions, ranges, indicesOfMutVars,
            maxValue, transientTime, values):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    n = nPts if dimension == 1 else nPts * nPts
    out = np.zeros(n)
    _check(load().bifhip_LLE_to_buffer(dimension, tMax, NT, nPts, h, eps, _pd(ic), len(ic),
                                       _pd(r), _pi(m), 0, maxValue, transientTime, _pd(v), len(v),
                                       _pd(out)))
    return out


def ls_map(dimension, tMax, NT, nPts, h, eps, initialConditions, ranges, indicesOfMutVars,
           maxValue, transientTime, values):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    n = nPts if dimension == 1 else nPts * nPts
    out = np.zeros(n * len(ic))
    _check(load().bifhip_LS_to_buffer(dimension, tMax, NT, nPts, h, eps, _pd(ic), len(ic), _pd(r),
                                      _pi(m), 0, maxValue, transientTime, _pd(v), len(v),
                                      _pd(out)))
    return out.reshape(n, len(ic))


# ------------------------------------------------- device-level entry points

def kde_map_device(d_out, tMax, nPts, h, initialConditions, ranges, indicesOfMutVars,
                   writableVar, maxValue, maxAmountOfPeaks, transientTime, values, preScaller,
                   kdeSampling, i1, i2, smooth):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_bifurcation2DKDE_device(tMax, nPts, h, len(ic), _pd(ic), _pd(r), _pi(m),
                                                 writableVar, maxValue, maxAmountOfPeaks,
                                                 transientTime, _pd(v), len(v), preScaller,
                                                 kdeSampling, i1, i2, smooth, d_out))


def dbscan_map_device(d_out, tMax, nPts, h, initialConditions, ranges, indicesOfMutVars,
                      writableVar, maxValue, maxAmountOfPeaks, transientTime, values, preScaller,
                      eps):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_bifurcation2DDBSCAN_device(tMax, nPts, h, len(ic), _pd(ic), _pd(r),
                                                    _pi(m), writableVar, maxValue,
                                                    maxAmountOfPeaks, transientTime, _pd(v),
                                                    len(v), preScaller, eps, d_out))


def lle_map_device(d_out, dimension, tMax, NT, nPts, h, eps, initialConditions, ranges,
                   indicesOfMutVars, maxValue, transientTime, values):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_LLE_device(dimension, tMax, NT, nPts, h, eps, _pd(ic), len(ic), _pd(r),
                                    _pi(m), 0, maxValue, transientTime, _pd(v), len(v), d_out))


def peak_finder_device(d_data, sizeOfBlock, nBlocks, d_counts, d_peaks, d_times, h=0.01):
    _check(load().bifhip_peak_finder_device(d_data, sizeOfBlock, nBlocks, d_counts, d_peaks,
                                            d_times, h))


def dbscan_device(d_peaks, d_ivls, sizeOfBlock, nBlocks, d_counts, maxPeaks, eps, d_out):
    _check(load().bifhip_dbscan_device(d_peaks, d_ivls, sizeOfBlock, nBlocks, d_counts, maxPeaks,
                                       eps, d_out))


def dbscan_device_noalloc(d_peaks, d_ivls, sizeOfBlock, nBlocks, d_counts, maxPeaks, eps,
                          d_out):
    """dbscan_device with the library's fixed per-GPU scratch array: no
    device allocation; at most NOALLOC_MAX_BLOCKS blocks and rows of at
    most NOALLOC_MAX_ROW peaks per call."""
    _check(load().bifhip_dbscan_device_noalloc(d_peaks, d_ivls, sizeOfBlock, nBlocks, d_counts,
                                               maxPeaks, eps, d_out))


NOALLOC_MAX_BLOCKS = 1 << 20
NOALLOC_MAX_ROW = 1 << 20


def kde_device(d_peaks, sizeOfBlock, nBlocks, d_counts, maxPeaks, sampling, i1, i2, smooth,
               d_out):
    _check(load().bifhip_kde_device(d_peaks, sizeOfBlock, nBlocks, d_counts, maxPeaks, sampling,
                                    i1, i2, smooth, d_out))


def gram_schmidt_device(d_a, d_b, d_den, n, count):
    _check(load().bifhip_gram_schmidt_device(d_a, d_b, d_den or None, n, count))


def series_device(d_data, d_flags, dimension, tMax, nPts, h, initialConditions, ranges,
                  indicesOfMutVars, writableVar, maxValue, transientTime, values, preScaller):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_series_device(dimension, tMax, nPts, h, len(ic), _pd(ic), _pd(r), _pi(m),
                                       writableVar, maxValue, transientTime, _pd(v), len(v),
                                       preScaller, d_data, d_flags))


def lle_peak_map_device(d_peak, d_lle, nPts, h, initialConditions, ranges, indicesOfMutVars,
                        values, peak_tmax, peak_transient, writableVar, maxAmountOfPeaks,
                        dbscan_eps, lle_tmax, NT, lle_transient, lle_eps):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_lle_peak_map_device(nPts, h, len(ic), _pd(ic), _pd(r), _pi(m), _pd(v),
                                             len(v), peak_tmax, peak_transient, writableVar,
                                             maxAmountOfPeaks, dbscan_eps, lle_tmax, NT,
                                             lle_transient, lle_eps, d_peak, d_lle))


# ------------------------------------------------------- surrogate

SSTAT = ["ms_total", "ms_coarse", "ms_train", "ms_infer", "ms_resolve", "resolved",
         "resolved_frac", "validation", "flagged_class_acc", "flagged_regime_acc",
         "flagged_lle_mae", "valid_class_acc", "valid_regime_acc", "valid_lle_mae",
         "loss_ce", "loss_mse", "lle_mean", "lle_sd", "peak_only", "peak_only_class_acc", "ms_flag",
         "ms_solve", "ms_score", "lle_only", "lle_only_regime_acc", "lle_only_recounted",
         "lle_only_fp32"]


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def surrogate_map_device(d_peak, d_lle, d_resolved, n_fine, n_coarse, h, initialConditions,
                         ranges, indicesOfMutVars, values, peak_tmax, peak_transient,
                         writableVar, maxAmountOfPeaks, dbscan_eps, lle_tmax, NT, lle_transient,
                         lle_eps, epochs=20, batch=1024, lr=1e-3, seed=0, conf=0.95,
                         chaos=0.01, validation=0.0):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    st = np.zeros(len(SSTAT))
    _check(load().bifhip_surrogate_map_device(n_fine, n_coarse, h, len(ic), _pd(ic), _pd(r),
                                              _pi(m), _pd(v), len(v), peak_tmax, peak_transient,
                                              writableVar, maxAmountOfPeaks, dbscan_eps, lle_tmax,
                                              NT, lle_transient, lle_eps, epochs, batch, lr, seed,
                                              conf, chaos, validation, d_peak, d_lle, d_resolved,
                                              _pd(st)))
    return dict(zip(SSTAT, st.tolist()))


def mlp_param_count():
    return load().bifhip_mlp_param_count()


def mlp_train_device(n_lat, d_cls, d_lle, epochs, batch, lr, w_lle, seed, sigma, chaos_std=0.0,
                     graph=True):
    loss = np.zeros(3, dtype=np.float32)
    _check(load().bifhip_mlp_train_device(n_lat, d_cls, d_lle, epochs, batch, lr, w_lle, seed,
                                          sigma, chaos_std, 1 if graph else 0, _pf(loss)))
    return loss


def mlp_infer_device(n_lat, mean, sd, chaos_thr, d_cls, d_chaos, d_conf, d_lle,
                     d_conf_chaos=None):
    _check(load().bifhip_mlp_infer_device(n_lat, mean, sd, chaos_thr, d_cls, d_chaos, d_conf,
                                          d_conf_chaos, d_lle))


def mlp_get_params():
    p = np.zeros(mlp_param_count(), dtype=np.float32)
    f = np.zeros(64, dtype=np.float32)
    _check(load().bifhip_mlp_get_params(_pf(p), len(p), _pf(f), len(f)))
    return p, f


def mlp_adam(params, m, v, grads, keep, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """k_adam applied len(grads) times (steps t = 1..) to FP32 params/m/v;
    grads[k] is (splits, n) partials.  Returns (params, m, v, bf16 bits)."""
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    m = np.ascontiguousarray(m, dtype=np.float32).copy()
    v = np.ascontiguousarray(v, dtype=np.float32).copy()
    g = np.ascontiguousarray(grads, dtype=np.float32)
    steps, splits, n = g.shape
    assert n == len(p) == len(m) == len(v)
    k = np.ascontiguousarray(keep, dtype=np.uint8)
    w = np.zeros(n, dtype=np.uint16)
    _check(load().bifhip_mlp_adam(n, _pf(p), _pf(m), _pf(v), _pf(g), splits, steps,
                                  k.ctypes.data_as(C.POINTER(C.c_ubyte)), lr, b1, b2, eps,
                                  w.ctypes.data_as(C.POINTER(C.c_ushort))))
    return p, m, v, w


def mlp_gradient(n_lat, params, fourier, cells, cls, lle, w_lle=1.0, chaos_std=0.0):
    params = np.ascontiguousarray(params, dtype=np.float32)
    fourier = np.ascontiguousarray(fourier, dtype=np.float32)
    cells, cls = _i32(cells), _i32(cls)
    lle = np.ascontiguousarray(lle, dtype=np.float32)
    g = np.zeros_like(params)
    loss = np.zeros(3, dtype=np.float32)
    out = np.zeros((len(cells), 32), dtype=np.float32)
    _check(load().bifhip_mlp_gradient(n_lat, _pf(params), _pf(fourier), _pi(cells), _pi(cls),
                                      _pf(lle), len(cells), w_lle, chaos_std, _pf(g), _pf(loss),
                                      _pf(out)))
    return g, loss, out


# ------------------------------------------------------- uneven LS tiles

def tiles_generate(n_tiles, tile_w, tile_h, n_pts, seed_origins, seed_lengths, t_min, t_max,
                   NT):
    oi = np.zeros(n_tiles, dtype=np.int32)
    oj = np.zeros(n_tiles, dtype=np.int32)
    nb = np.zeros(n_tiles, dtype=np.int32)
    _check(load().bifhip_tiles_generate(n_tiles, tile_w, tile_h, n_pts, seed_origins,
                                        seed_lengths, t_min, t_max, NT, _pi(oi), _pi(oj),
                                        _pi(nb)))
    return oi, oj, nb


def ls_tiles_device(d_out, tiles, tile_w, tile_h, n_pts, ranges, indicesOfMutVars,
                    initialConditions, values, h, NT, transient, autotune=False):
    oi, oj, nb = (_i32(t) for t in tiles)
    r, m, ic, v = _f64(ranges), _i32(indicesOfMutVars), _f64(initialConditions), _f64(values)
    fn = load().bifhip_ls_tiles_autotune if autotune else load().bifhip_ls_tiles_device
    _check(fn(len(oi), tile_w, tile_h, _pi(oi), _pi(oj), _pi(nb), n_pts, _pd(r), _pi(m), _pd(ic),
              len(ic), _pd(v), len(v), h, NT, transient, d_out))


def set_tile_config(block, cells_per_lane=1, blocks_per_cu=0):
    _check(load().bifhip_set_tile_config(block, cells_per_lane, blocks_per_cu))


def get_tile_config():
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    _check(load().bifhip_get_tile_config(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def set_tile_lanes(lanes):
    """Lanes per cell of the FP32 tile kernel: 1 or 3."""
    _check(load().bifhip_set_tile_lanes(int(lanes)))


def get_tile_lanes():
    a = C.c_int()
    _check(load().bifhip_get_tile_lanes(C.byref(a)))
    return a.value


def tile_autotune_log():
    """Candidates of the last autotuner run: rows (block, lanes, cells per
    lane, ms on the sample)."""
    n = load().bifhip_tile_autotune_log(None, 0)
    out = np.zeros(max(n, 1) * 4)
    load().bifhip_tile_autotune_log(_pd(out), n)
    return out[:4 * n].reshape(n, 4)


def set_tile_precision(p):
    """PREC_FP32 (default) or PREC_FP64 state for the tile path."""
    _check(load().bifhip_set_tile_precision(int(p)))


def set_graph(on):
    _check(load().bifhip_set_graph(1 if on else 0))


def set_partial_resolve(on):
    _check(load().bifhip_set_partial_resolve(1 if on else 0))


def set_chaos_confidence(threshold):
    _check(load().bifhip_set_chaos_confidence(float(threshold)))


def set_lle32_margin(margin):
    """Regime-only re-solves of the surrogate map: FP32 bulk LLE where it
    lies at least `margin` from the chaos threshold, FP64 elsewhere; a
    negative margin keeps them all in FP64."""
    _check(load().bifhip_set_lle32_margin(float(margin)))


def set_quad_confidence(threshold):
    """Coarse-quad chaos rule of the surrogate: 0 off (default); cells in a
    coarse quad with mixed exact chaos bits re-solve their LLE when their
    chaos confidence is below the threshold."""
    _check(load().bifhip_set_quad_confidence(float(threshold)))


def set_regime_reading(reading):
    """What a cell's regime is for the surrogate's re-solve decision:
    READING_CLASS_AND_CHAOS (default), READING_CHAOS or READING_CLASS."""
    _check(load().bifhip_set_regime_reading(int(reading)))


# --------------------------------------------------------------- writers

def write_map_int(path, nPts, ranges, values, header=True, precision=6):
    r, v = _f64(ranges), _i32(values)
    _check(load().bifhip_write_map_int(str(path).encode(), nPts, _pd(r), int(header), precision,
                                       _pi(v), len(v)))


def write_map_double(path, nPts, ranges, values, header=True, precision=6, stride=1, offset=0):
    r, v = _f64(ranges), _f64(values)
    _check(load().bifhip_write_map_double(str(path).encode(), nPts, _pd(r), int(header),
                                          precision, _pd(v), len(v) // stride, stride, offset))


def write_bif1d(path, nPts, ranges, counts, peaks):
    r, c = _f64(ranges), _i32(counts)
    p = _f64(peaks)
    _check(load().bifhip_write_bif1d(str(path).encode(), nPts, _pd(r), _pi(c), _pd(p),
                                     p.shape[1] if p.ndim == 2 else 1))


def write_1d_values(path, nPts, ranges, values, per_cell=1):
    r, v = _f64(ranges), _f64(values)
    _check(load().bifhip_write_1d_values(str(path).encode(), nPts, _pd(r), _pd(v), per_cell))


def header_symbols():
    """Function names declared in include/bifhip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bifhip_[A-Za-z0-9_]+)\s*\(", txt)))
