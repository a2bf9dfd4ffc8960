"""CPU tests of the file readers (cudabifurcation3.0_amd/formats.py): files
written by the library's writers, and text in the reference's layout,
read back to the values and ranges that went in."""
import numpy as np

import bifhip as B
import formats as F
from test_abi import _ref_map_text


def test_map2d_round_trip_int_and_double(tmp_path):
    rng = np.random.default_rng(1)
    ranges = [0.05, 0.45, 2.0, 10.0]
    n = 7
    ints = rng.integers(0, 40, size=n * n).astype(np.int32)
    p = tmp_path / "m.csv"
    B.write_map_int(p, n, ranges, ints)
    r, m = F.read_map2d(p)
    np.testing.assert_array_equal(r, [[0.05, 0.45], [2, 10]])
    np.testing.assert_array_equal(m, ints.reshape(n, n))
    # doubles at precision 17 round-trip exactly
    vals = rng.normal(size=n * n) * 1e-3
    B.write_map_double(p, n, ranges, vals, precision=17)
    _, m = F.read_map2d(p)
    np.testing.assert_array_equal(m, vals.reshape(n, n))


def test_map2d_reads_reference_layout_and_matlab_orientation(tmp_path):
    """Row j holds the second-axis index j: the cell of flat index j*n + i
    is values[j, i], what imagesc(x, y, a) shows at (x_i, y_j)."""
    n = 4
    flat = np.arange(n * n)
    p = tmp_path / "m.csv"
    p.write_text(_ref_map_text(n, [1, 2, 3, 4], flat, 6))
    r, m = F.read_map2d(p)
    assert m[2, 1] == 2 * n + 1
    assert m.shape == (n, n)
    # header-less files (bifurcation2DKDEIC, LLE2DIC)
    p.write_text(_ref_map_text(n, [1, 2, 3, 4], flat, 6, header=False))
    r, m = F.read_map2d(p, header=False)
    assert r is None and m[3, 0] == 3 * n


def test_map2d_nan_cells(tmp_path):
    p = tmp_path / "m.csv"
    p.write_text("0 1\n0 1\n1, nan, \n-nan, 2")
    _, m = F.read_map2d(p)
    assert m[0, 0] == 1 and np.isnan(m[0, 1]) and np.isnan(m[1, 0]) and m[1, 1] == 2


def test_ls2d_files(tmp_path):
    n = 3
    vals = np.arange(n * n * 3, dtype=float).reshape(n * n, 3)
    for k in range(3):
        B.write_map_double(tmp_path / f"mat.csv{k + 1}.csv", n, [0, 1, 0, 1], vals.ravel(),
                           precision=12, stride=3, offset=k)
    r, maps = F.read_ls2d(str(tmp_path / "mat.csv"), 3)
    for k in range(3):
        np.testing.assert_array_equal(maps[k], vals[:, k].reshape(n, n))


def test_one_dimensional_files(tmp_path):
    p = tmp_path / "b.csv"
    counts = np.array([2, 0, 1], dtype=np.int32)
    peaks = np.array([[1.5, 2.5], [0, 0], [3.25, 0]])
    B.write_bif1d(p, 3, [2.0, 7.0], counts, peaks)
    par, pk = F.read_bif1d(p)
    np.testing.assert_array_equal(par, [2.0, 2.0, 7.0])
    np.testing.assert_array_equal(pk, [1.5, 2.5, 3.25])
    B.write_1d_values(p, 3, [2.0, 7.0], np.arange(9, dtype=float), per_cell=3)
    a = F.read_lines(p)
    assert a.shape == (3, 4)
    np.testing.assert_array_equal(a[:, 0], F.lattice(2.0, 7.0, 3))


def test_lattice_is_value_by_idx():
    for n, s, e in ((1000, 2.0, 7.0), (512, 0.05, 0.45), (17, -6.0, 6.0)):
        want = [B.value_by_idx(i, n, s, e, 0) for i in range(n)]
        np.testing.assert_array_equal(F.lattice(s, e, n), want)


def test_map3d_layout(tmp_path):
    n = 3
    vol = np.arange(n ** 3).reshape(n, n, n)
    lines = ["0 1", "2 3", "4 5"]
    for k in range(n):
        for j in range(n):
            lines.append(", ".join(str(v) for v in vol[k, j]) + ", ")
    p = tmp_path / "v.csv"
    p.write_text("\n".join(lines))
    r, v = F.read_map3d(p)
    np.testing.assert_array_equal(r, [[0, 1], [2, 3], [4, 5]])
    np.testing.assert_array_equal(v, vol)


def _mixed_doubles(rng, n):
    """Magnitudes over the whole double range, subnormals, signed zeros,
    infinities and NaN, short decimals and exact integers."""
    v = rng.normal(size=n) * 10.0 ** rng.integers(-300, 300, size=n)
    k = rng.integers(0, 10, size=n)
    v[k == 1] = np.round(v[k == 1], 3)
    v[k == 2] = rng.integers(-10 ** 6, 10 ** 6, size=int((k == 2).sum()))
    v[k == 3] = rng.normal(size=int((k == 3).sum())) * 1e-310
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308])
    v[k == 4] = special[rng.integers(0, len(special), size=int((k == 4).sum()))]
    return v


def test_map_writers_threaded_text_equals_the_reference_layout(tmp_path):
    """Maps large enough to be formatted by several host threads (1024^2
    values; 1000 values per row, so row breaks fall inside the threads'
    ranges): the file equals the reference writer's row logic with printf
    %.*g values, at precisions 6 and 12, and %d for counts."""
    rng = np.random.default_rng(5)
    ranges = [0.05, 0.45, 2.0, 10.0]
    p = tmp_path / "big.csv"
    for nPts, count in ((1024, 1024 * 1024), (1000, 1000 * 1000)):
        v = _mixed_doubles(rng, count)
        for prec in (6, 12):
            B.write_map_double(p, nPts, ranges, v, precision=prec)
            want = _ref_map_text(nPts, ranges, v, prec, fmt=lambda x, q=prec: "%.*g" % (q, x))
            assert p.read_text() == want, (nPts, prec)
    ints = rng.integers(-2 ** 31, 2 ** 31, size=1024 * 1024).astype(np.int32)
    ints[:3] = [0, -2 ** 31, 2 ** 31 - 1]
    B.write_map_int(p, 1024, ranges, ints, header=True, precision=6)
    assert p.read_text() == _ref_map_text(1024, ranges, ints, 6)


def test_bif1d_writer_threaded_text(tmp_path):
    """A 1D diagram large enough to be formatted by several host threads
    (1000 points, about 50 000 peak lines, some points with none): one line
    "parameter, peak" per peak at %.12g, points in order."""
    rng = np.random.default_rng(9)
    n, stride = 1000, 120
    counts = rng.integers(0, 100, size=n).astype(np.int32)
    counts[::37] = 0
    peaks = _mixed_doubles(rng, n * stride).reshape(n, stride)
    p = tmp_path / "bif.csv"
    B.write_bif1d(p, n, [2.0, 7.0], counts, peaks)
    want = []
    for k in range(n):
        pv = "%.12g" % B.value_by_idx(k, n, 2.0, 7.0, 0)
        for j in range(counts[k]):
            want.append(f"{pv}, {'%.12g' % peaks[k, j]}\n")
    assert p.read_text() == "".join(want)
