"""GPU parity of the drivers that sweep initial conditions (the reference's
*IC drivers, Library/hostLibrary.cu:243, 929, 1094, 1397, 1666): the swept
lattice values go into the initial state, LLE directions follow
LLEKernelICCUDA, and the files keep each driver's layout.

Standards as in test_gpu_parity: peak counts and 1D peaks bit-exact, LLE
within 1e-12 (device and glibc `log` may differ in the last ulp), KDE by
count agreement.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IC = [0.1, 0.1, 0.1]
VALS = [0.2, 0.2, 5.7]
# (x0, z0) box of initial conditions around the attractor
RIC = [-4.0, 4.0, 0.0, 2.0]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _body(text, header):
    lines = text.split("\n")
    if header:
        assert lines[:2] == ["-4 4", "0 2"], lines[:2]
        lines = lines[2:]
    else:
        assert lines[0] != "-4 4"
    return "".join(lines).split(", ")


def test_dbscan_ic_file_bit_exact(orc, bh, tmp_path):
    from test_abi import _ref_map_text
    n = 12
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    say("bifurcation2DDBSCANIC", n)
    bh.bifurcation2DDBSCANIC(80.0, n, 0.01, 3, IC, RIC, [0, 2], 0, 0.0, 100000, 50.0, VALS, 3,
                             1, 0.01)
    orc.set_sweep_ic(True)
    try:
        want = orc.dbscan_cells(80.0, n, 0.01, IC, RIC, [0, 2], 0, 0.0, 100000, 50.0, VALS, 1,
                                0.01, 0, n * n)
    finally:
        orc.set_sweep_ic(False)
    assert out.read_text() == _ref_map_text(n, RIC, want, 6)
    assert len(set(want.tolist())) > 1


def test_dbscan_ic_buffer_matches_driver(orc, bh):
    """The sweep knob routes the buffer entry points through the same path."""
    n = 16
    bh.set_sweep(bh.SWEEP_INITIAL_CONDITIONS)
    say("dbscan map, initial-condition sweep", n)
    m = bh.dbscan_map(60.0, n, 0.01, IC, RIC, [0, 2], 0, 0.0, 100000, 40.0, VALS, 1, 0.01)
    orc.set_sweep_ic(True)
    try:
        want = orc.dbscan_cells(60.0, n, 0.01, IC, RIC, [0, 2], 0, 0.0, 100000, 40.0, VALS, 1,
                                0.01, 0, n * n)
    finally:
        orc.set_sweep_ic(False)
    np.testing.assert_array_equal(m, want)
    bh.set_sweep(bh.SWEEP_PARAMETERS)
    m2 = bh.dbscan_map(60.0, n, 0.01, IC, [0.05, 0.45, 2.0, 10.0], [0, 2], 0, 0.0, 100000, 40.0,
                       VALS, 1, 0.01)
    want2 = orc.dbscan_cells(60.0, n, 0.01, IC, [0.05, 0.45, 2.0, 10.0], [0, 2], 0, 0.0, 100000,
                             40.0, VALS, 1, 0.01, 0, n * n)
    np.testing.assert_array_equal(m2, want2)


def test_lle2d_ic_file_matches_oracle(orc, bh, tmp_path):
    n = 10
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    say("LLE2DIC", n)
    bh.LLE2DIC(40.0, 0.5, n, 0.01, 1e-8, IC, 3, RIC, [0, 2], 0, 0.0, 20.0, VALS, 3)
    orc.set_sweep_ic(True)
    try:
        want = orc.lle_cells(2, 40.0, 0.5, n, 0.01, 1e-8, IC, RIC, [0, 2], 0.0, 20.0, VALS,
                             orc.PERT_REFERENCE, 0, 0, n * n)
    finally:
        orc.set_sweep_ic(False)
    body = _body(out.read_text(), header=False)
    assert len(body) == n * n
    for b_, w_ in zip(body, want):
        w = f"{w_:g}"
        # 6 significant digits; a last-ulp difference may round the other way
        assert b_ == w or (b_ in ("nan", "-nan") and w == "nan") or \
            abs(float(b_) - float(w)) <= 1e-5 * abs(float(w)) + 1e-12, (b_, w)


def test_lle_ic_buffer_within_ulps(orc, bh):
    n = 12
    bh.set_sweep(bh.SWEEP_INITIAL_CONDITIONS)
    say("LLE buffer, initial-condition sweep", n)
    got = bh.lle_map(2, 60.0, 0.5, n, 0.01, 1e-8, IC, RIC, [0, 2], 0.0, 20.0, VALS)
    orc.set_sweep_ic(True)
    try:
        want = orc.lle_cells(2, 60.0, 0.5, n, 0.01, 1e-8, IC, RIC, [0, 2], 0.0, 20.0, VALS,
                             orc.PERT_REFERENCE, 0, 0, n * n)
    finally:
        orc.set_sweep_ic(False)
    fin = np.isfinite(want)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    assert np.abs(got[fin] - want[fin]).max() <= 1e-12
    say("max err", float(np.abs(got[fin] - want[fin]).max()))


def test_one_dimensional_ic_files(orc, bh, tmp_path):
    n = 12
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    say("bifurcation1DIC / LLE1DIC", n)
    bh.bifurcation1DIC(100.0, n, 0.01, 3, IC, [-3.0, 3.0], [1], 0, 0.0, 200.0, VALS, 3, 1)
    N = int(100.0 / 0.01)
    orc.set_sweep_ic(True)
    try:
        wc, wp, _ = orc.peaks_cells(1, 100.0, n, 0.01, IC, [-3.0, 3.0], [1], 0, 0.0, 200.0, VALS,
                                    1, 0, n, N)
        lle = orc.lle_cells(1, 40.0, 0.5, n, 0.01, 1e-8, IC, [-3.0, 3.0], [1], 0.0, 20.0, VALS,
                            orc.PERT_REFERENCE, 0, 0, n)
    finally:
        orc.set_sweep_ic(False)
    lines = []
    for k in range(n):
        pv = orc.value_by_idx(k, n, -3.0, 3.0, 0)
        for j in range(max(wc[k], 0)):
            lines.append(f"{pv:.12g}, {wp[k, j]:.12g}\n")
    assert out.read_text() == "".join(lines)

    bh.LLE1DIC(40.0, 0.5, n, 0.01, 1e-8, IC, 3, [-3.0, 3.0], [1], 0, 0.0, 20.0, VALS, 3)
    got = out.read_text().splitlines()
    assert len(got) == n
    for k, line in enumerate(got):
        a, b = line.split(", ")
        assert a == f"{orc.value_by_idx(k, n, -3.0, 3.0, 0):g}"
        w = f"{lle[k]:g}"
        assert b == w or abs(float(b) - float(w)) <= 1e-5 * abs(float(w)) + 1e-12, (b, w)


def test_kde_ic_file_layout_and_agreement(orc, bh, tmp_path):
    n = 10
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    say("bifurcation2DKDEIC", n)
    bh.bifurcation2DKDEIC(80.0, n, 0.01, 3, IC, RIC, [0, 2], 0, 0.0, 1000, 50.0, VALS, 3, 1, 10,
                          -15.0, 15.0, 0.05)
    got = [int(v) for v in _body(out.read_text(), header=False)]
    orc.set_sweep_ic(True)
    try:
        want = orc.kde_cells(80.0, n, 0.01, IC, RIC, [0, 2], 0, 0.0, 1000, 50.0, VALS, 1, 10,
                             -15.0, 15.0, 0.05, 0, n * n)
    finally:
        orc.set_sweep_ic(False)
    agree = float(np.mean(np.array(got) == want))
    say("KDE IC agreement", agree)
    assert agree >= 0.995


def test_ic_sweep_rejects_out_of_range_index(bh):
    """indicesOfMutVars index the initial conditions in a sweep over them."""
    bh.set_sweep(bh.SWEEP_INITIAL_CONDITIONS)
    with pytest.raises(bh.BifhipError):
        bh.dbscan_map(10.0, 4, 0.01, IC, RIC, [0, 3], 0, 0.0, 100, 5.0, VALS, 1, 0.01)
