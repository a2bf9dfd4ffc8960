"""ctypes binding of libbifhip.so (the gfx950 HIP library).

Mirrors the reference's host API (Library/hostLibrary.cuh) one to one:
the reference-named functions below take the same arguments in the same
order and raise BifhipError where the reference would print and exit().
Device buffers are passed as integer addresses (e.g. torch tensor
data_ptr()).  There is no CPU fallback: a missing library raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbifhip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "bifhip.h")

MODEL_ROSSLER = 0
MODEL_LINEAR = 1
PREC_FP64 = 0
PREC_FP32 = 1
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
        "bifhip_release": [], "bifhip_device_count": [],
        "bifhip_bifurcation1D": drv_bif1,
        "bifhip_bifurcation2DKDE": drv_kde,
        "bifhip_bifurcation2DDBSCAN": drv_db,
        "bifhip_LLE1D": drv_ly, "bifhip_LLE2D": drv_ly,
        "bifhip_LS1D": drv_ly, "bifhip_LS2D": drv_ly,
        "bifhip_bifurcation1D_to_buffer": drv_bif1 + [_PI, _PD, _I],
        "bifhip_bifurcation2DKDE_to_buffer": drv_kde + [_PI],
        "bifhip_bifurcation2DDBSCAN_to_buffer": drv_db + [_PI],
        "bifhip_LLE_to_buffer": [_I] + drv_ly + [_PD],
        "bifhip_LS_to_buffer": [_I] + drv_ly + [_PD],
        "bifhip_bifurcation2DDBSCAN_device": drv_db + [_VP],
        "bifhip_LLE_device": [_I] + drv_ly + [_VP],
        "bifhip_peak_finder_device": [_VP, _I, _I, _VP, _VP, _VP, _D],
        "bifhip_dbscan_device": [_VP, _VP, _I, _I, _VP, _I, _D, _VP],
        "bifhip_kde_device": [_VP, _I, _I, _VP, _I, _I, _D, _D, _D, _VP],
        "bifhip_gram_schmidt_device": [_VP, _VP, _VP, _I, _I],
        "bifhip_series_device": [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _D, _PD, _I,
                                 _I, _VP, _VP],
        "bifhip_write_map_int": [C.c_char_p, _I, _PD, _I, _I, _PI, _LL],
        "bifhip_write_map_double": [C.c_char_p, _I, _PD, _I, _I, _PD, _LL, _I, _I],
        "bifhip_write_bif1d": [C.c_char_p, _I, _PD, _PI, _PD, _I],
        "bifhip_write_1d_values": [C.c_char_p, _I, _PD, _PD, _I],
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


def device_count():
    return load().bifhip_device_count()


def value_by_idx(idx, nPts, s, f, k):
    return load().bifhip_value_by_idx(idx, nPts, s, f, k)


# ---------------------------------------- reference-named drivers (files)

def bifurcation1D(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                  indicesOfMutVars, writableVar, maxValue, transientTime, values,
                  amountOfValues, preScaller):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_bifurcation1D(tMax, nPts, h, amountOfInitialConditions, _pd(ic), _pd(r),
                                       _pi(m), writableVar, maxValue, transientTime, _pd(v),
                                       amountOfValues, preScaller))


def bifurcation2DKDE(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                     indicesOfMutVars, writableVar, maxValue, maxAmountOfPeaks, transientTime,
                     values, amountOfValues, preScaller, kdeSampling, kdeSamplesInterval1,
                     kdeSamplesInterval2, kdeSamplesSmooth):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_bifurcation2DKDE(tMax, nPts, h, amountOfInitialConditions, _pd(ic),
                                          _pd(r), _pi(m), writableVar, maxValue,
                                          maxAmountOfPeaks, transientTime, _pd(v),
                                          amountOfValues, preScaller, kdeSampling,
                                          kdeSamplesInterval1, kdeSamplesInterval2,
                                          kdeSamplesSmooth))


def bifurcation2DDBSCAN(tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                        indicesOfMutVars, writableVar, maxValue, maxAmountOfPeaks,
                        transientTime, values, amountOfValues, preScaller, eps):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    _check(load().bifhip_bifurcation2DDBSCAN(tMax, nPts, h, amountOfInitialConditions, _pd(ic),
                                             _pd(r), _pi(m), writableVar, maxValue,
                                             maxAmountOfPeaks, transientTime, _pd(v),
                                             amountOfValues, preScaller, eps))


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


# ------------------------------------------------ in-memory (host) twins

def bifurcation1D_to_buffer(tMax, nPts, h, initialConditions, ranges, indicesOfMutVars,
                            writableVar, maxValue, transientTime, values, preScaller, stride=None):
    ic, r, m, v = _f64(initialConditions), _f64(ranges), _i32(indicesOfMutVars), _f64(values)
    N = int(tMax / h / preScaller)
    stride = stride or max(1, N // 2 + 1)
    counts = np.zeros(nPts, dtype=np.int32)
    peaks = np.zeros(nPts * stride)
    _check(load().bifhip_bifurcation1D_to_buffer(tMax, nPts, h, len(ic), _pd(ic), _pd(r), _pi(m),
                                                 writableVar, maxValue, transientTime, _pd(v),
                                                 len(v), preScaller, _pi(counts), _pd(peaks),
                                                 stride))
    return counts, peaks.reshape(nPts, stride)


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
