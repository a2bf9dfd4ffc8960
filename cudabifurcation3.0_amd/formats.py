"""Readers for the CSV files the drivers write, equivalent to the reference's
MATLAB scripts (no GPU, numpy only).

File layouts (written by Library/hostLibrary.cu and by libbifhip.so):

  1D diagram   "param, peak" lines, one per peak (bifurcation1D, -IC);
               MatlabScripts/bifurcation1D.m reads columns 1 and 2.
  1D Lyapunov  "param, v1[, v2, ...]" lines, one per lattice point (LLE1D,
               LS1D, LLE1DIC); MatlabScripts/LS.m reads columns 1..4.
  2D map       two header lines "start end" (absent in bifurcation2DKDEIC and
               LLE2DIC), then nPts values per row separated by ", ", each
               row but the last ending in ", " (so MATLAB's csvread sees an
               empty last column); row j holds the second swept value j,
               column i the first swept value i.  MatlabScripts/
               bifuraction2D.m:13-18 drops the two header rows and the last
               column; LS2D writes one such file per exponent, <path>k.csv
               (MatlabScripts/LS2D.m:1-14).
  3D map       three header lines, then n blocks of n rows of n values
               (MatlabScripts/bifurcation3D.m:1-16; no driver writes it in
               the reference, the reader is here for files from elsewhere).
"""
import numpy as np


def _rows(path):
    with open(path) as f:
        text = f.read()
    return text.split("\n")


def _values(line):
    """Numbers of one map row: ", "-separated, empty fields (the trailing
    separator) dropped; "nan"/"-nan"/"inf" as written by C++ streams."""
    out = []
    for tok in line.split(","):
        tok = tok.strip()
        if tok:
            out.append(float(tok))
    return out


def lattice(start, end, n):
    """The swept values of one axis, as getValueByIdx
    (Library/cudaLibrary.cu:210-216): start + i * ((end - start) / (n - 1))."""
    if n == 1:
        return np.array([float(start)])
    step = (float(end) - float(start)) / (n - 1)
    return float(start) + np.arange(n) * step


def read_map2d(path, header=True):
    """2D map -> (ranges, values).  ranges is [[start0, end0], [start1,
    end1]] (None without a header); values[j, i] is the cell of second-axis
    index j and first-axis index i."""
    lines = _rows(path)
    ranges = None
    if header:
        ranges = np.array([[float(v) for v in lines[k].split()] for k in (0, 1)])
        lines = lines[2:]
    data = [_values(ln) for ln in lines if ln.strip()]
    if not data:
        return ranges, np.zeros((0, 0))
    n = len(data[0])
    if any(len(r) != n for r in data):
        raise ValueError(f"{path}: rows of unequal length")
    return ranges, np.array(data)


def read_ls2d(prefix, n_exponents, header=True):
    """LS2D output, files <prefix>1.csv .. <prefix>n.csv -> (ranges,
    values[k][j, i])."""
    maps = []
    ranges = None
    for k in range(1, n_exponents + 1):
        ranges, m = read_map2d(f"{prefix}{k}.csv", header)
        maps.append(m)
    return ranges, np.stack(maps)


def read_lines(path):
    """1D diagram or 1D Lyapunov file -> 2D array, one row per line."""
    rows = [_values(ln) for ln in _rows(path) if ln.strip()]
    if not rows:
        return np.zeros((0, 2))
    return np.array(rows)


def read_bif1d(path):
    """1D diagram -> (params, peaks), one entry per recorded peak."""
    a = read_lines(path)
    return a[:, 0], a[:, 1]


def read_map3d(path):
    """3D map -> (ranges [3 x 2], values[k, j, i])."""
    lines = _rows(path)
    ranges = np.array([[float(v) for v in lines[k].split()] for k in (0, 1, 2)])
    data = [_values(ln) for ln in lines[3:] if ln.strip()]
    n = len(data[0]) if data else 0
    if n == 0 or len(data) != n * n or any(len(r) != n for r in data):
        raise ValueError(f"{path}: expected {n} blocks of {n} rows of {n} values")
    return ranges, np.array(data).reshape(n, n, n)
