"""GPU parity of the five-variable memristive model (the active body of the
reference's calculateDiscreteModel, Library/cudaLibrary.cu:19-43) through
the same drivers: the composition step is polynomial up to IEEE division,
float rounding and |x|, all correctly rounded on both sides, so series,
peak counts and DBSCAN maps are bit-exact; LLE / LS within 1e-12 (`log`).
The setting is the reference's active main.cpp call (main.cpp:115-131):
h = 0.005, x0 = (0.001, 0, 0, 0, 0), values {8, 35, 15, 0.01, -1.6676,
0.0676, 0.3682, 0.5}, initial conditions x4 in [-6, 6], x3 in [-15, 15].
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IC5 = [0.001, 0.0, 0.0, 0.0, 0.0]
MEM = [8.0, 35.0, 15.0, 0.01, -1.6676, 0.0676, 0.3682, 0.5]
RIC = [-6.0, 6.0, -15.0, 15.0]
MUT = [4, 3]
# parameter sweep: a0 in [7, 9], a6 in [0.3, 0.45]
RPAR = [7.0, 9.0, 0.3, 0.45]
MPAR = [0, 6]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def memristive(bh):
    bh.set_model(bh.MODEL_MEMRISTIVE)
    yield
    bh.set_model(bh.MODEL_ROSSLER)
    bh.set_sweep(bh.SWEEP_PARAMETERS)


def test_series_bit_exact(orc, bh):
    n, tmax, h, trans = 8, 5.0, 0.005, 10.0
    N = int(tmax / h)
    for sweep in (bh.SWEEP_INITIAL_CONDITIONS, bh.SWEEP_PARAMETERS):
        rng, mut = (RIC, MUT) if sweep else (RPAR, MPAR)
        bh.set_sweep(sweep)
        say("memristive series, sweep", sweep)
        data = bh.DeviceArray((n * n * N,), np.float64).zero_()
        flags = bh.DeviceArray((n * n,), np.int32).zero_()
        bh.series_device(data.ptr, flags.ptr, 2, tmax, n, h, IC5, rng, mut, 4, 0.0, trans, MEM,
                         1)
        bh.synchronize()
        got = data.to_host().reshape(n * n, N)
        orc.set_sweep_ic(sweep == bh.SWEEP_INITIAL_CONDITIONS)
        try:
            for cell in range(0, n * n, 7):
                ok, s = orc.bif_cell(cell, n, rng, mut, IC5, MEM, h, int(trans / h), N, 1, 4,
                                     model=orc.MODEL_MEMRISTIVE)
                np.testing.assert_array_equal(got[cell], s, err_msg=f"cell {cell}")
        finally:
            orc.set_sweep_ic(False)


def test_dbscan_maps_bit_exact(orc, bh):
    n = 12
    for sweep in (bh.SWEEP_INITIAL_CONDITIONS, bh.SWEEP_PARAMETERS):
        rng, mut = (RIC, MUT) if sweep else (RPAR, MPAR)
        bh.set_sweep(sweep)
        say("memristive dbscan map, sweep", sweep)
        m = bh.dbscan_map(20.0, n, 0.005, IC5, rng, mut, 4, 1e7, 100000, 20.0, MEM, 1, 0.02)
        orc.set_sweep_ic(sweep == bh.SWEEP_INITIAL_CONDITIONS)
        try:
            want = orc.dbscan_cells(20.0, n, 0.005, IC5, rng, mut, 4, 1e7, 100000, 20.0, MEM, 1,
                                    0.02, 0, n * n, model=orc.MODEL_MEMRISTIVE)
        finally:
            orc.set_sweep_ic(False)
        np.testing.assert_array_equal(m, want)
        say("distinct counts", sorted(set(want.tolist()))[:12])


def test_lle_and_ls_within_ulps(orc, bh):
    n = 8
    say("memristive LLE / LS maps")
    lle = bh.lle_map(2, 10.0, 0.4, n, 0.005, 1e-6, IC5, RPAR, MPAR, 1e7, 5.0, MEM)
    want = orc.lle_cells(2, 10.0, 0.4, n, 0.005, 1e-6, IC5, RPAR, MPAR, 1e7, 5.0, MEM,
                         orc.PERT_REFERENCE, 0, 0, n * n, model=orc.MODEL_MEMRISTIVE)
    fin = np.isfinite(want)
    np.testing.assert_array_equal(np.isfinite(lle), fin)
    assert np.abs(lle[fin] - want[fin]).max() <= 1e-12
    ls = bh.ls_map(2, 4.0, 0.4, n, 0.005, 1e-6, IC5, RPAR, MPAR, 1e7, 5.0, MEM)
    wls = orc.ls_cells(2, 4.0, 0.4, n, 0.005, 1e-6, IC5, RPAR, MPAR, 1e7, 5.0, MEM,
                       orc.PERT_REFERENCE, 0, 0, n * n, model=orc.MODEL_MEMRISTIVE)
    assert ls.shape == (n * n, 5)
    fin = np.isfinite(wls)
    np.testing.assert_array_equal(np.isfinite(ls), fin)
    err = np.abs(ls[fin] - wls[fin]).max()
    say("LS max err", float(err))
    assert err <= 1e-12


def test_unsupported_paths_refuse_the_model(bh):
    """The fused map, the tile path and FP32 bulk are three-variable paths."""
    d_pk = bh.DeviceArray((16,), np.int32)
    d_ll = bh.DeviceArray((16,), np.float64)
    with pytest.raises(bh.BifhipError):
        bh.lle_peak_map_device(d_pk.ptr, d_ll.ptr, 4, 0.005, IC5, RPAR, MPAR, MEM, 10.0, 5.0, 0,
                               1000, 0.01, 10.0, 0.5, 5.0, 1e-6)
    bh.set_precision(bh.PREC_FP32)
    try:
        with pytest.raises(bh.BifhipError):
            bh.lle_map(2, 2.0, 0.4, 4, 0.005, 1e-6, IC5, RPAR, MPAR, 1e7, 1.0, MEM)
    finally:
        bh.set_precision(bh.PREC_FP64)
    with pytest.raises(bh.BifhipError):
        bh.dbscan_map(2.0, 4, 0.005, IC5[:3], RPAR, MPAR, 0, 0.0, 100, 1.0, MEM, 1, 0.02)


def test_ostrovsky_classifier_matches_oracle(orc, bh, tmp_path):
    """calculationOfPeriodicityByOstrovsky on the reference's own setting
    (shortened): every cell's value, a count, its negation, +-1, 0 or
    +-0.5, equals the oracle's exactly; the file has the 2D layout."""
    from test_abi import _ref_map_text
    n = 12
    args = (30.0, n, 0.005, 5, IC5, RIC, MUT, 4, 1e7, 10000000, 20.0, MEM, 8, 1, 0.02, 2.0)
    say("Ostrovsky map", n)
    got = bh.ostrovsky_map(30.0, n, 0.005, IC5, RIC, MUT, 4, 1e7, 10000000, 20.0, MEM, 1, 0.02,
                           2.0)
    want = orc.ostrovsky_cells(30.0, n, 0.005, IC5, RIC, MUT, 4, 1e7, 10000000, 20.0, MEM, 1,
                               0.02, 2.0, 0, n * n, model=orc.MODEL_MEMRISTIVE)
    np.testing.assert_array_equal(got, want)
    say("values", sorted(set(want.tolist()))[:12])
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    bh.calculationOfPeriodicityByOstrovsky(*args)
    assert out.read_text() == _ref_map_text(n, RIC, want, 6, fmt=lambda v: f"{v:g}")


def test_ostrovsky_rossler(orc, bh):
    """The classifier on the Rossler system (initial conditions x0, z0
    swept), with divergence checks on."""
    bh.set_model(bh.MODEL_ROSSLER)
    n = 10
    ic = [0.1, 0.1, 0.1]
    rng = [-2.0, 2.0, 0.0, 1.0]
    got = bh.ostrovsky_map(100.0, n, 0.01, ic, rng, [0, 2], 0, 1e3, 100000, 100.0,
                           [0.2, 0.2, 5.7], 1, 0.01, 1.0)
    want = orc.ostrovsky_cells(100.0, n, 0.01, ic, rng, [0, 2], 0, 1e3, 100000, 100.0,
                               [0.2, 0.2, 5.7], 1, 0.01, 1.0, 0, n * n)
    np.testing.assert_array_equal(got, want)
