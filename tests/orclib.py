"""ctypes binding of the CPU oracle (oracle/_build/liborc.so).

Test infrastructure only: the oracle is the checker for the HIP path and
the timed CPU baseline in bench.py, never part of the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liborc.so")

MODEL_ROSSLER = 0
MODEL_LINEAR = 1
MODEL_MEMRISTIVE = 2
PERT_REFERENCE = 0
PERT_AXIS_X = 1

_D = C.c_double
_I = C.c_int
_L = C.c_long
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class use_library:
    """Context manager: route every call of this module through another
    build of the oracle (a copy with a planted mistake, for checking that
    the pinning tests catch it)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        lib()
        self.saved = _lib
        other = C.CDLL(self.path)
        _declare(other)
        _lib = other
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _declare(L):
    L.orc_set_sweep_ic.argtypes = [_I]
    L.orc_set_sweep_ic.restype = None
    L.orc_ostrovsky_cells.argtypes = [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _I, _D, _PD, _I,
                                      _I, _D, _D, _L, _L, _PD]
    L.orc_ostrovsky_cells.restype = None
    L.orc_denis.argtypes = [_I, _D, _D, _D, _I, _PD, _I, _PD, _I, _PD, _L]
    L.orc_denis.restype = None
    L.orc_step.argtypes = [_I, _PD, _PD, _D]
    L.orc_loop.argtypes = [_I, _PD, _PD, _D, _I, _I, _I, _D, _PD, _L, _I]
    L.orc_loop.restype = _I
    L.orc_value_by_idx.argtypes = [_I, _I, _D, _D, _I]
    L.orc_value_by_idx.restype = _D
    L.orc_peak_finder.argtypes = [_PD, _L, _I, _PD, _PD, _D]
    L.orc_peak_finder.restype = _I
    L.orc_kde.argtypes = [_PD, _L, _I, _I, _I, _D, _D, _D]
    L.orc_kde.restype = _I
    L.orc_kde_sums.argtypes = [_PD, _I, _I, _D, _D, _D, _PD, _I]
    L.orc_kde_sums.restype = _I
    L.orc_distance.argtypes = [_D, _D, _D, _D]
    L.orc_distance.restype = _D
    L.orc_dbscan.argtypes = [_PD, _PD, _PD, _L, _I, _I, _I, _D]
    L.orc_dbscan.restype = _I
    L.orc_gram_schmidt.argtypes = [_PD, _PD, _I, _PD]
    L.orc_bif_cell.argtypes = [_I, _I, _L, _I, _PD, _PI, _PD, _I, _PD, _I,
                               _D, _I, _I, _I, _I, _D, _PD]
    L.orc_bif_cell.restype = _I
    L.orc_bif_cell32.argtypes = L.orc_bif_cell.argtypes
    L.orc_bif_cell32.restype = _I
    L.orc_lle_cell.argtypes = [_I, _I, _L, _L, _I, _PD, _PI, _PD, _I, _PD, _I,
                               _D, _D, _D, _D, _I, _D, _I]
    L.orc_lle_cell.restype = _D
    L.orc_ls_cell.argtypes = [_I, _I, _L, _L, _I, _PD, _PI, _PD, _I, _PD, _I,
                              _D, _D, _D, _D, _I, _D, _I, _PD]
    L.orc_peaks_cells.argtypes = [_I, _I, _D, _I, _D, _I, _PD, _PD, _PI, _I,
                                  _D, _D, _PD, _I, _I, _L, _L, _PI, _PD, _PD,
                                  _I]
    L.orc_bifurcation1d.argtypes = [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D,
                                    _D, _PD, _I, _I, _PI, _PD, _I]
    L.orc_dbscan_cells.argtypes = [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D,
                                   _I, _D, _PD, _I, _I, _D, _L, _L, _PI]
    L.orc_peaks_cells32.argtypes = L.orc_peaks_cells.argtypes
    L.orc_dbscan_cells32.argtypes = L.orc_dbscan_cells.argtypes
    L.orc_kde_cells.argtypes = [_I, _D, _I, _D, _I, _PD, _PD, _PI, _I, _D, _I,
                                _D, _PD, _I, _I, _I, _D, _D, _D, _L, _L, _PI]
    L.orc_lle_cells.argtypes = [_I, _I, _D, _D, _I, _D, _D, _PD, _I, _PD, _PI,
                                _D, _D, _PD, _I, _I, _L, _L, _L, _PD]
    L.orc_ls_cells.argtypes = [_I, _I, _D, _D, _I, _D, _D, _PD, _I, _PD, _PI,
                               _D, _D, _PD, _I, _I, _L, _L, _L, _PD]
    L.orc_threads.restype = _I
    L.orc_ls_var_cell.argtypes = [_D, _D, _D, _PD, _D, _I, _I, _I, _PD]
    L.orc_ls_var_cells.argtypes = [_PD, _PD, _PD, _PD, _D, _I, _I, _PI, _L, _PD]


def pd(a):
    return a.ctypes.data_as(_PD)


def pi(a):
    return a.ctypes.data_as(_PI)


def f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


# ---------------- small wrappers ----------------

def set_sweep_ic(on):
    """Swept entries go into the initial conditions (True) or the parameters."""
    lib().orc_set_sweep_ic(1 if on else 0)


def step(x, a, h, model=MODEL_ROSSLER):
    x = f64(x).copy()
    lib().orc_step(model, pd(x), pd(f64(a)), h)
    return x


def value_by_idx(idx, n, s, f, k):
    return lib().orc_value_by_idx(idx, n, s, f, k)


def peak_finder(series, h=0.01):
    d = f64(series).copy()
    n = len(d)
    pk = np.zeros(max(n, 1))
    tp = np.zeros(max(n, 1))
    P = lib().orc_peak_finder(pd(d), 0, n, pd(pk), pd(tp), h)
    return P, pk[:P].copy(), tp[:P].copy()


def dbscan(amps, ivls, eps, max_peaks=1 << 30, size=None):
    a = f64(amps)
    iv = f64(ivls)
    P = len(a)
    size = size if size is not None else max(2 * P + 2, 1)
    hp = np.zeros(size)
    a2 = np.zeros(max(P, 1)); a2[:P] = a
    i2 = np.zeros(max(P, 1)); i2[:P] = iv
    return lib().orc_dbscan(pd(a2), pd(i2), pd(hp), 0, P, size, max_peaks, eps)


def kde(amps, max_peaks, sampling, i1, i2, smooth):
    a = f64(amps)
    a2 = np.zeros(max(len(a), 1)); a2[:len(a)] = a
    return lib().orc_kde(pd(a2), 0, len(a), max_peaks, sampling, i1, i2, smooth)


def kde_sums(amps, sampling, i1, i2, smooth):
    """The k1-1 KDE values orc_kde compares (k1 = sampling * P)."""
    a = f64(amps)
    a2 = np.zeros(max(len(a), 1)); a2[:len(a)] = a
    cap = max(int(sampling) * len(a), 1)
    out = np.zeros(cap)
    n = lib().orc_kde_sums(pd(a2), len(a), sampling, i1, i2, smooth, pd(out), cap)
    return out[:n].copy()


def distance(x1, y1, x2, y2):
    return lib().orc_distance(x1, y1, x2, y2)


def gram_schmidt(a):
    a = f64(a)
    n = a.shape[0]
    b = np.zeros_like(a)
    den = np.zeros(n)
    lib().orc_gram_schmidt(pd(a), pd(b), n, pd(den))
    return b, den


def bif_cell(cell, nPts, ranges, mut, ic, values, h, skip, N, pre=1, wv=0,
             max_value=0.0, dimension=None, model=MODEL_ROSSLER, fp32=False):
    """fp32: the FP32 form (float state, fused FP32 step)."""
    ranges = f64(ranges); mut = i32(mut); ic = f64(ic); values = f64(values)
    dimension = dimension if dimension is not None else len(mut)
    s = np.zeros(N)
    fn = lib().orc_bif_cell32 if fp32 else lib().orc_bif_cell
    ok = fn(model, nPts, cell, dimension, pd(ranges), pi(mut), pd(ic), len(ic), pd(values),
            len(values), h, skip, N, pre, wv, max_value, pd(s))
    return ok, s


def peaks_cells(dimension, tMax, nPts, h, ic, ranges, mut, wv, max_value,
                transient, values, pre, c0, nc, cap, model=MODEL_ROSSLER, fp32=False):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    counts = np.zeros(nc, dtype=np.int32)
    peaks = np.zeros(nc * cap)
    ivl = np.zeros(nc * cap)
    fn = lib().orc_peaks_cells32 if fp32 else lib().orc_peaks_cells
    fn(model, dimension, tMax, nPts, h, len(ic), pd(ic), pd(ranges), pi(mut), wv, max_value,
       transient, pd(values), len(values), pre, c0, nc, pi(counts), pd(peaks), pd(ivl), cap)
    return counts, peaks.reshape(nc, cap), ivl.reshape(nc, cap)


def dbscan_cells(tMax, nPts, h, ic, ranges, mut, wv, max_value, max_peaks,
                 transient, values, pre, eps, c0, nc, model=MODEL_ROSSLER, fp32=False):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    out = np.zeros(nc, dtype=np.int32)
    fn = lib().orc_dbscan_cells32 if fp32 else lib().orc_dbscan_cells
    fn(model, tMax, nPts, h, len(ic), pd(ic), pd(ranges), pi(mut), wv, max_value, max_peaks,
       transient, pd(values), len(values), pre, eps, c0, nc, pi(out))
    return out


def ostrovsky_cells(tMax, nPts, h, ic, ranges, mut, wv, max_value, max_peaks, transient,
                    values, pre, eps, threshold, c0, nc, model=MODEL_ROSSLER):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    out = np.zeros(nc)
    lib().orc_ostrovsky_cells(model, tMax, nPts, h, len(ic), pd(ic), pd(ranges), pi(mut), wv,
                              max_value, max_peaks, transient, pd(values), len(values), pre,
                              eps, threshold, c0, nc, pd(out))
    return out


def denis(tMax, h, hS, ic, wv, values, model=MODEL_ROSSLER):
    ic = f64(ic); values = f64(values)
    n = int(tMax / h)
    out = np.zeros(max(n, 0))
    lib().orc_denis(model, tMax, h, hS, len(ic), pd(ic), wv, pd(values), len(values), pd(out),
                    len(out))
    return out


def kde_cells(tMax, nPts, h, ic, ranges, mut, wv, max_value, max_peaks,
              transient, values, pre, sampling, i1, i2, smooth, c0, nc,
              model=MODEL_ROSSLER):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    out = np.zeros(nc, dtype=np.int32)
    lib().orc_kde_cells(model, tMax, nPts, h, len(ic), pd(ic), pd(ranges),
                        pi(mut), wv, max_value, max_peaks, transient,
                        pd(values), len(values), pre, sampling, i1, i2, smooth,
                        c0, nc, pi(out))
    return out


def lle_cells(dimension, tMax, NT, nPts, h, eps, ic, ranges, mut, max_value,
              transient, values, pert, dir_chunk, c0, nc, model=MODEL_ROSSLER):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    out = np.zeros(nc)
    lib().orc_lle_cells(model, dimension, tMax, NT, nPts, h, eps, pd(ic),
                        len(ic), pd(ranges), pi(mut), max_value, transient,
                        pd(values), len(values), pert, dir_chunk, c0, nc,
                        pd(out))
    return out


def ls_cells(dimension, tMax, NT, nPts, h, eps, ic, ranges, mut, max_value,
             transient, values, pert, dir_chunk, c0, nc, model=MODEL_ROSSLER):
    ic = f64(ic); ranges = f64(ranges); mut = i32(mut); values = f64(values)
    out = np.zeros(nc * len(ic))
    lib().orc_ls_cells(model, dimension, tMax, NT, nPts, h, eps, pd(ic),
                       len(ic), pd(ranges), pi(mut), max_value, transient,
                       pd(values), len(values), pert, dir_chunk, c0, nc,
                       pd(out))
    return out.reshape(nc, len(ic))


def ls_var_cells(a, b, c, ic, h, skip, nNT, nBlocks):
    a, b, c, ic = f64(a), f64(b), f64(c), f64(ic)
    nb = i32(nBlocks)
    out = np.zeros((len(a), 3))
    lib().orc_ls_var_cells(pd(a), pd(b), pd(c), pd(ic), h, skip, nNT, pi(nb), len(a), pd(out))
    return out
