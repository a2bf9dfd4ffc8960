"""Pins the CPU oracle (CPU only, no GPU).

Each oracle function is checked against something other than a retyping
of its own formula: the reference's known-answer values (golden file),
closed forms of the linear test model, invariants of Gram-Schmidt and of
the Lyapunov spectrum, numpy's correctly rounded arithmetic, and brute
force on small inputs.
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
IC = [0.1, 0.1, 0.1]


# ------------------------------------------------------------ golden KATs

def test_peak_finder_reference_kat(orc):
    for case in GOLD["peak_finder"]:
        P, pk, iv = orc.peak_finder(case["series"], case["h"])
        assert P == case["count"]
        assert list(pk) == case["peaks"]
        assert list(iv) == case["intervals"]


def test_dbscan_reference_kats(orc):
    for case in GOLD["dbscan_eps001_intervals0"]:
        a = case["amps"]
        assert orc.dbscan(a, [0.0] * len(a), 0.01) == case["result"], a


def test_hypotf_reference_kat(orc):
    (c,) = GOLD["hypotf"]
    d = orc.distance(0.0, 0.0, c["x"], c["y"])
    assert float(np.float32(d)).hex() == float.fromhex(c["float_result_hex"]).hex()
    assert (d < 0.01) == c["neighbour_at_eps_0.01"]


@pytest.mark.parametrize("case", GOLD["rossler_dbscan_cells"]["maxAmountOfPeaks_1e5"],
                         ids=lambda c: f"c={c['c']}")
def test_rossler_dbscan_reference_counts(orc, case):
    c = case["c"]
    # 2D grid whose cell 0 is (a, c) = (0.2, c)
    rng2 = [0.2, 0.2, c, c]
    r = orc.dbscan_cells(500, 2, 0.01, IC, rng2, [0, 2], 0, 0.0, 100000, 1000, [0.2, 0.2, c], 1,
                         0.01, 0, 1)
    assert r[0] == case["result"]
    r32 = orc.dbscan_cells(500, 2, 0.01, IC, rng2, [0, 2], 0, 0.0, 32, 1000, [0.2, 0.2, c], 1,
                           0.01, 0, 1)
    assert r32[0] == 0
    cnt, _, _ = orc.peaks_cells(1, 500, 2, 0.01, IC, [c, c], [2], 0, 0.0, 1000, [0.2, 0.2, c], 1,
                                0, 1, 256)
    lo, hi = GOLD["rossler_peak_counts"]["approximate_range"]
    assert lo - 2 <= cnt[0] <= hi + 2


# ------------------------------------------------------ float semantics

def test_distance_is_double_evaluated_hypotf(orc):
    """glibc hypotf == (float)sqrt((double)x*x + (double)y*y) -- the identity
    the GPU neighbour test relies on.  numpy's sqrt is correctly rounded."""
    rng = np.random.default_rng(0)
    x1 = rng.normal(size=20000) * rng.choice([1e-3, 1e-2, 1, 1e3], 20000)
    y1 = rng.normal(size=20000) * 1e-2
    x2 = x1 + rng.normal(size=20000) * 1e-2
    y2 = y1 + rng.choice([0.0, 0.01, 0.02], 20000)
    fx = (x2 - x1).astype(np.float32).astype(np.float64)
    fy = (y2 - y1).astype(np.float32).astype(np.float64)
    want = np.sqrt(fx * fx + fy * fy).astype(np.float32).astype(np.float64)
    got = np.array([orc.distance(*v) for v in zip(x1, y1, x2, y2)])
    np.testing.assert_array_equal(got, want)


def test_value_by_idx_closed_form(orc):
    """getValueByIdx == start + i*((end-start)/(n-1)) with i = idx/n^k mod n
    for grids up to 4096^2 (float division exact there)."""
    for n, s, f in [(512, 0.05, 0.45), (512, 2.0, 10.0), (1000, 2.0, 7.0), (4096, 2.0, 10.0)]:
        step = (f - s) / (n - 1)
        for idx in list(range(0, n * n, max(1, n * n // 4099))) + [n * n - 1]:
            for k in (0, 1):
                i = (idx // n ** k) % n
                want = s + float(i) * step
                assert orc.value_by_idx(idx, n, s, f, k) == want, (n, idx, k)


# ------------------------------------------------------------ loop / step

def test_loop_records_before_step_and_stops_on_divergence(orc):
    """Sample 0 is the state after the transient; the divergence check stops
    the record loop and reports 0 (Library/cudaLibrary.cu:55-74)."""
    ok, s = orc.bif_cell(0, 2, [5.7, 5.7], [2], IC, [0.2, 0.2, 5.7], 0.01, 0, 10)
    assert ok == 1 and s[0] == 0.1
    x = orc.step(IC, [0.2, 0.2, 5.7], 0.01)
    assert s[1] == x[0]
    ok, s = orc.bif_cell(0, 2, [5.7, 5.7], [2], IC, [0.2, 0.2, 5.7], 0.01, 100, 50,
                         max_value=1e-3)
    assert ok == 0
    # NaN never trips the check (comparison is false)
    ok, _ = orc.bif_cell(0, 2, [5.7, 5.7], [2], [float("nan")] * 3, [0.2, 0.2, 5.7], 0.01, 0,
                         10, max_value=1.0)
    assert ok == 1


def test_rossler_step_is_sequential_euler(orc):
    """One step against an exact rational evaluation of the sequential
    scheme (x1 and x2 use the new x0), rounded once at the end: the oracle's
    per-operation rounding keeps it within a few ulps, while the
    simultaneous-Euler variant misses by far more."""
    from fractions import Fraction as F
    x = [0.7, -1.3, 0.4]
    a, b, c, h = 0.2, 0.2, 5.7, 0.01
    X = [F(v) for v in x]
    A, B, Cc, H = F(a), F(b), F(c), F(h)
    n0 = X[0] + H * (-X[1] - X[2])
    n1 = X[1] + H * (n0 + A * X[1])
    n2 = X[2] + H * (B + X[2] * (n0 - Cc))
    got = orc.step(x, [a, b, c], h)
    for g, w in zip(got, (n0, n1, n2)):
        assert abs(F(g) - w) <= F(4) * F(math.ulp(float(w)))
    s1 = X[1] + H * (X[0] + A * X[1])  # simultaneous variant
    assert abs(F(got[1]) - s1) > F(1000) * F(math.ulp(float(s1)))


def test_fp32_step_is_sequential_euler(orc):
    """The FP32 form (float state, fused FP32 step) against the same exact
    rational sequential step from the same float state: within a few float
    ulps, while the simultaneous variant (x1 from the old x0) or a flipped
    sign misses by far more."""
    from fractions import Fraction as F
    x = [0.75, -1.25, 0.375]                 # exact in float
    a, b, c, h = 0.2, 0.2, 5.7, 0.01
    X = [F(v) for v in x]
    A, B, Cc, H = F(a), F(b), F(c), F(h)
    n0 = X[0] + H * (-X[1] - X[2])
    n1 = X[1] + H * (n0 + A * X[1])
    n2 = X[2] + H * (B + X[2] * (n0 - Cc))
    got = []
    for wv in range(3):
        ok, s = orc.bif_cell(0, 2, [c, c], [2], x, [a, b, c], h, 0, 2, wv=wv, fp32=True)
        assert ok == 1 and s[0] == x[wv]
        got.append(s[1])
    for g, w in zip(got, (n0, n1, n2)):
        ulp = F(float(np.spacing(np.float32(abs(float(w))))))
        assert abs(F(g) - w) <= 4 * ulp, (g, float(w))
    s1 = X[1] + H * (X[0] + A * X[1])        # simultaneous variant
    assert abs(F(got[1]) - s1) > 100 * F(float(np.spacing(np.float32(abs(float(s1))))))


def test_fp32_linear_model_power_law(orc):
    """Linear model in FP32: x_n = x_0 (1+h*a)^n up to one float rounding of
    the factor and one per step; divergence sets the -1 count as in FP64."""
    h, a, n = 0.01, 0.3, 1000
    ok, s = orc.bif_cell(0, 2, [a, a], [0], [1.0, 1.0, 1.0], [a, a, a], h, 0, n + 1,
                         model=orc.MODEL_LINEAR, fp32=True)
    assert ok == 1
    k = np.arange(n + 1)
    want = np.exp(k * math.log1p(h * a))
    rel = np.abs(s - want) / want
    assert rel.max() < 3 * n * 2.0 ** -24, rel.max()
    assert rel[-1] > 1e-9                    # the state is float, not double
    for fp32 in (False, True):
        cnt, _, _ = orc.peaks_cells(1, 50, 2, h, [1.0, 1.0, 1.0], [50.0, 50.0], [0], 0, 1000.0,
                                    0, [50.0, 50.0, 50.0], 1, 0, 1, 16, model=orc.MODEL_LINEAR,
                                    fp32=fp32)
        assert cnt[0] == -1


def test_fp32_period1_rossler_matches_fp64(orc):
    """c = 2.5 (period 1): the FP32 trajectory settles on the same limit
    cycle -- the same DBSCAN count as the reference's (1) and the same
    peak amplitudes to float accuracy."""
    c = 2.5
    rng2 = [0.2, 0.2, c, c]
    args = (500, 2, 0.01, IC, rng2, [0, 2], 0, 0.0, 100000, 1000, [0.2, 0.2, c], 1, 0.01, 0, 1)
    assert orc.dbscan_cells(*args)[0] == 1
    assert orc.dbscan_cells(*args, fp32=True)[0] == 1
    c64, p64, _ = orc.peaks_cells(1, 500, 2, 0.01, IC, [c, c], [2], 0, 0.0, 1000, [0.2, 0.2, c],
                                  1, 0, 1, 256)
    c32, p32, _ = orc.peaks_cells(1, 500, 2, 0.01, IC, [c, c], [2], 0, 0.0, 1000, [0.2, 0.2, c],
                                  1, 0, 1, 256, fp32=True)
    assert abs(int(c32[0]) - int(c64[0])) <= 1
    m = min(c32[0], c64[0])
    assert np.abs(p32[0, :m].mean() - p64[0, :m].mean()) < 1e-4
    assert not np.array_equal(p32[0, :m], p64[0, :m])


# ------------------------------------------------------------ peak finder

def test_peak_finder_plateau_free_is_textbook(orc):
    """Without equal neighbours the state machine reduces to strict local
    maxima with the first one dropped; intervals are index gaps * h."""
    rng = np.random.default_rng(1)
    for n in (3, 5, 17, 200, 2001):
        d = rng.normal(size=n).cumsum()
        i = np.nonzero((d[1:-1] > d[:-2]) & (d[1:-1] > d[2:]))[0] + 1
        P, pk, iv = orc.peak_finder(d, 0.01)
        if len(i) <= 1:
            assert P == 0
            continue
        assert P == len(i) - 1
        np.testing.assert_array_equal(pk, d[i[1:]])
        np.testing.assert_array_equal(iv, np.diff(i).astype(np.float64) * 0.01)


def test_peak_finder_edge_cases(orc):
    assert orc.peak_finder([])[0] == 0
    assert orc.peak_finder([1.0])[0] == 0
    assert orc.peak_finder([0, 1, 0, 1, 0])[0] == 1
    # a rising plateau exit skips the next candidate (the reference quirk)
    P, pk, _ = orc.peak_finder([0, 9, 0, 5, 5, 6, 1, 0])
    assert P == 0


# ------------------------------------------------------------------- KDE

def test_kde_small_cases(orc):
    assert orc.kde([], 10, 10, 0, 1, 0.1) == 0
    assert orc.kde([0.5], 10, 10, 0, 1, 0.1) == 1
    assert orc.kde([0.5, 0.6], 10, 10, 0, 1, 0.1) == 1
    assert orc.kde([0.1] * 20, 5, 10, 0, 1, 0.1) == 5          # P > max returns max
    assert orc.kde([0.5, 0.5, 0.5], 10, 10, 0, 1, 0.05) == 1    # one Gaussian bump
    assert orc.kde([0.2, 0.2, 0.8, 0.8], 10, 10, 0, 1, 0.05) == 2
    assert orc.kde([0.1, 0.1, 0.5, 0.5, 0.9, 0.9], 10, 10, 0, 1, 0.03) == 3


# ---------------------------------------------------------- Gram-Schmidt

def test_gram_schmidt_invariants(orc):
    rng = np.random.default_rng(2)
    for n in (2, 3, 5, 8):
        for _ in range(20):
            a = rng.normal(size=(n, n))
            b, den = orc.gram_schmidt(a)
            np.testing.assert_allclose(b @ b.T, np.eye(n), atol=1e-10)
            assert abs(abs(np.linalg.det(a)) - np.prod(den)) <= 1e-9 * np.prod(den)
            q, r = np.linalg.qr(a.T)
            np.testing.assert_allclose(np.abs(b), np.abs(q.T), atol=1e-9)
            np.testing.assert_allclose(den, np.abs(np.diag(r)), rtol=1e-10)
    # zero vector -> zero row and zero norm
    b, den = orc.gram_schmidt(np.array([[1.0, 0, 0], [0, 0, 0], [0, 1.0, 0]]))
    assert den[1] == 0 and not b[1].any()


# ---------------------------------------------------------- Lyapunov

LAM = [0.3, -0.5, -1.0]


def _ref_lle_dir(idx):
    """0.5*sinf of the wrapped int idx*(i*idx+1)+1, normalised (float32 sin;
    a last-ulp difference from glibc's sinf moves the LLE by < 1e-9)."""
    z = []
    for i in range(3):
        arg = (idx * (i * idx + 1) + 1) & 0xFFFFFFFF
        arg = arg - (1 << 32) if arg >= (1 << 31) else arg
        z.append(0.5 * float(np.sin(np.float32(arg))))
    z = np.array(z)
    return z / np.linalg.norm(z)


def test_lle_linear_model_closed_form(orc):
    """x_i += h*a_i*x_i: with the perturbation along x the LLE is exactly
    log(1+h*a_0)/h; with the reference directions z it is the finite-time
    value log|M^N z|/T, M = diag(1+h*a).  The zero state keeps x exact so the
    perturbation evolves linearly."""
    h, T, NT = 0.01, 200, 0.5
    out = orc.lle_cells(1, T, NT, 4, h, 1e-8, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20, LAM,
                        orc.PERT_AXIS_X, 0, 0, 4, model=orc.MODEL_LINEAR)
    np.testing.assert_allclose(out, math.log(1 + h * 0.3) / h, rtol=1e-9)
    out = orc.lle_cells(1, T, NT, 4, h, 1e-8, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20, LAM,
                        orc.PERT_REFERENCE, 0, 0, 4, model=orc.MODEL_LINEAR)
    N = int(T / NT) * int(NT / h)
    lg = np.log1p(h * np.array(LAM)) * N
    for idx in range(4):
        z = _ref_lle_dir(idx)
        m = lg.max()
        want = (m + 0.5 * np.log(np.sum((z * np.exp(lg - m)) ** 2))) / T
        assert abs(out[idx] - want) < 1e-8, (idx, out[idx], want)


def _ic_lle_dir(idx):
    """LLEKernelICCUDA's direction: 0.5*sinf(0.2171828(i+1) + idx +
    (0.2171828 + i*idx)), normalised (float32 sin as in _ref_lle_dir)."""
    z = np.array([0.5 * float(np.sin(np.float32(0.2171828 * (i + 1) + idx + (0.2171828 + i * idx))))
                  for i in range(3)])
    return z / np.linalg.norm(z)


def test_lle_ic_sweep_directions_closed_form(orc):
    """A sweep over initial conditions seeds the perturbation with
    LLEKernelICCUDA's formula: on the linear model at the zero state (the
    swept initial value is 0 in every cell) the LLE is log|M^N z|/T for that
    z, and differs from the parameter-sweep direction's value."""
    h, T, NT = 0.01, 200, 0.5
    orc.set_sweep_ic(True)
    try:
        out = orc.lle_cells(1, T, NT, 4, h, 1e-8, [0, 0, 0], [0.0, 0.0], [0], 0.0, 20, LAM,
                            orc.PERT_REFERENCE, 0, 0, 4, model=orc.MODEL_LINEAR)
    finally:
        orc.set_sweep_ic(False)
    N = int(T / NT) * int(NT / h)
    lg = np.log1p(h * np.array(LAM)) * N
    m = lg.max()
    for idx in range(4):
        want = (m + 0.5 * np.log(np.sum((_ic_lle_dir(idx) * np.exp(lg - m)) ** 2))) / T
        other = (m + 0.5 * np.log(np.sum((_ref_lle_dir(idx) * np.exp(lg - m)) ** 2))) / T
        assert abs(out[idx] - want) < 1e-8, (idx, out[idx], want)
        assert abs(out[idx] - other) > 1e-6


def test_ic_sweep_substitutes_initial_state(orc):
    """Sweeping initial conditions writes the lattice values into the state
    (calculateDiscreteModelICCUDA): each cell's series equals a plain run
    from the initial conditions with those entries replaced by the closed-form
    lattice values, parameters untouched."""
    n, h, skip, N = 6, 0.01, 300, 400
    rng = [-1.0, 1.0, 0.0, 2.0]
    vals = [0.2, 0.2, 5.7]
    orc.set_sweep_ic(True)
    try:
        got = [orc.bif_cell(c, n, rng, [0, 2], IC, vals, h, skip, N)[1] for c in range(n * n)]
    finally:
        orc.set_sweep_ic(False)
    for c in range(n * n):
        i, j = c % n, c // n
        x = [rng[0] + i * ((rng[1] - rng[0]) / (n - 1)), IC[1],
             rng[2] + j * ((rng[3] - rng[2]) / (n - 1))]
        # a one-point parameter sweep [a, a] leaves the parameters as they are
        _, want = orc.bif_cell(0, n, [vals[0], vals[0]], [0], x, vals, h, skip, N)
        np.testing.assert_array_equal(got[c], want, err_msg=str(c))
    assert not np.array_equal(got[0], got[n * n - 1])


def test_ls_linear_model_closed_form(orc):
    """Axis directions stay aligned under a diagonal map: the spectrum is
    exactly log(1+h*a_i)/h; reference directions converge to the same set."""
    h = 0.01
    want = [math.log(1 + h * v) / h for v in LAM]
    out = orc.ls_cells(1, 200, 0.5, 3, h, 1e-6, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20, LAM,
                       orc.PERT_AXIS_X, 0, 0, 3, model=orc.MODEL_LINEAR)
    for row in out:
        np.testing.assert_allclose(row, want, rtol=1e-9)
    out = orc.ls_cells(1, 200, 0.5, 3, h, 1e-6, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20, LAM,
                       orc.PERT_REFERENCE, 0, 0, 3, model=orc.MODEL_LINEAR)
    for row in out:
        np.testing.assert_allclose(row, want, atol=2e-2)


def test_ls_sum_equals_volume_contraction(orc):
    """Sum of the spectrum == time average of log|det J|/h along the orbit.
    For the sequential Euler Rossler step det J = (1+h a)(1+h(x0'-c))."""
    a, b, c, h = 0.2, 0.2, 5.7, 0.01
    tmax, trans = 400, 100
    ls = orc.ls_cells(1, tmax, 0.5, 2, h, 1e-8, IC, [c, c], [2], 0.0, trans, [a, b, c],
                      orc.PERT_REFERENCE, 0, 0, 1)[0]
    ok, s = orc.bif_cell(0, 2, [c, c], [2], IC, [a, b, c], h, int(trans / h), int(tmax / h) + 1)
    x0_new = s[1:]  # x0 after each step of the measured window
    vol = np.mean(np.log(np.abs((1 + h * a) * (1 + h * (x0_new - c))))) / h
    assert abs(ls.sum() - vol) < 2e-3 * abs(vol)
    assert ls[0] > 0.03 and abs(ls[1]) < 0.02 and ls[2] < -1.0


def test_lle_rossler_sanity(orc):
    """Chaotic c=5.7: LLE ~ 0.07; period-1 c=2.5: LLE ~ 0 (neutral flow
    direction), within finite-time noise."""
    vals = [0.2, 0.2, 5.7]
    lle = orc.lle_cells(1, 500, 0.5, 2, 0.01, 1e-8, IC, [5.7, 5.7], [2], 0.0, 200, vals,
                        orc.PERT_REFERENCE, 0, 0, 1)[0]
    assert 0.04 < lle < 0.11
    vals = [0.2, 0.2, 2.5]
    lle = orc.lle_cells(1, 500, 0.5, 2, 0.01, 1e-8, IC, [2.5, 2.5], [2], 0.0, 200, vals,
                        orc.PERT_REFERENCE, 0, 0, 1)[0]
    assert abs(lle) < 0.01


def test_lle_direction_chunk_changes_direction_only(orc):
    """The reference seeds the direction from the chunk-local index; a
    different chunk must give a slightly different (finite-time) value."""
    args = (2, 50, 0.5, 8, 0.01, 1e-8, IC, [0.05, 0.45, 2, 10], [0, 2], 0.0, 10,
            [0.2, 0.2, 5.7], orc.PERT_REFERENCE)
    full = orc.lle_cells(*args, 0, 0, 64)
    chunked = orc.lle_cells(*args, 10, 0, 64)
    np.testing.assert_array_equal(full[:10], chunked[:10])
    assert (full[10:] != chunked[10:]).any()


def test_variational_ls_matches_finite_difference(orc):
    """The FP64 tangent-map spectrum (oracle of the tile path) agrees with the
    reference-style finite-difference LS on the same cells."""
    h, T, trans = 0.01, 400.0, 100.0
    for c in (2.5, 4.0, 5.7):
        fd = orc.ls_cells(1, T, 0.5, 2, h, 1e-8, IC, [c, c], [2], 0.0, trans, [0.2, 0.2, c],
                          orc.PERT_AXIS_X, 0, 0, 1)[0]
        var = orc.ls_var_cells([0.2], [0.2], [c], IC, h, int(trans / h), 50, [int(T / 0.5)])[0]
        np.testing.assert_allclose(var, fd, atol=2e-4)


# --------------------------------------------- five-variable memristive model

# parameter vector of the reference's active main.cpp call (main.cpp:127)
MEM = [8.0, 35.0, 15.0, 0.01, -1.6676, 0.0676, 0.3682, 0.5]


def _mem_rhs(x, a):
    """The ODE both passes of the composition step discretise."""
    return np.array([a[0] * (x[2] - (a[4] + a[5] + a[6] * abs(x[4])) * x[0]),
                     -x[2] + x[3],
                     a[1] * (x[1] - x[0] - x[2]),
                     -a[2] * x[1] - a[3] * x[3],
                     x[0]])


def test_memristive_step_converges_to_its_ode(orc):
    """For every split a7 (explicit share 1, 1/2, 0) repeated steps converge
    to the ODE's solution at first order: halving h halves the error.  A
    dropped term, a wrong sign or a swapped variable changes the limit and
    fails this."""
    from scipy.integrate import solve_ivp
    x0 = [0.5, -0.2, 0.3, 0.1, -0.4]
    T = 0.5
    for a7 in (1.0, 0.5, 0.0):
        a = MEM[:7] + [a7]
        ref = solve_ivp(lambda t, x: _mem_rhs(x, a), (0, T), x0, method="DOP853", rtol=1e-12,
                        atol=1e-14).y[:, -1]
        errs = []
        for h in (2e-3, 1e-3, 5e-4):
            x = np.array(x0)
            for _ in range(int(round(T / h))):
                x = orc.step(x, a, h, model=orc.MODEL_MEMRISTIVE)
            errs.append(float(np.abs(x - ref).max()))
        assert errs[0] < 0.1 * np.abs(ref).max(), (a7, errs)
        assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1], (a7, errs)


def test_memristive_step_rounds_the_state_to_float(orc):
    """With a7 = 1 the semi-implicit pass has zero length, so a step returns
    the explicit pass rounded to float (the reference's float y[5]); with
    a7 = 0 the explicit pass has zero length."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        x = rng.normal(size=5)
        y = orc.step(x, MEM[:7] + [1.0], 0.005, model=orc.MODEL_MEMRISTIVE)
        assert (y.astype(np.float32).astype(np.float64) == y).all()
        assert not (x.astype(np.float32).astype(np.float64) == x).all()
    # fixed point of the ODE: the origin
    z = orc.step(np.zeros(5), MEM, 0.005, model=orc.MODEL_MEMRISTIVE)
    assert (z == 0).all()


# -------------------------------------------------- Ostrovsky classifier

@pytest.mark.parametrize("case", GOLD["rossler_dbscan_cells"]["maxAmountOfPeaks_1e5"][:4],
                         ids=lambda c: f"c={c['c']}")
def test_ostrovsky_signs_over_reference_counts(orc, case):
    """The classifier's output is the DBSCAN count (pinned by the
    reference's values for these cells) with the reference's flag rules:
    every kept peak above the threshold -> +count, every one below ->
    -count, peaks on both sides -> 0."""
    c = case["c"]
    want = case["result"]
    args = (500.0, 2, 0.01, IC, [0.1, 0.1, 0.1, 0.1], [0, 2], 0, 0.0, 100000, 1000.0,
            [0.2, 0.2, c], 1, 0.01)
    assert orc.ostrovsky_cells(*args, -100.0, 0, 1)[0] == want
    assert orc.ostrovsky_cells(*args, 100.0, 0, 1)[0] == -want
    ok, s = orc.bif_cell(0, 2, [0.2, 0.2], [0], IC, [0.2, 0.2, c], 0.01, 100000, 50000)
    P, pk, _ = orc.peak_finder(s, 0.01)
    mid = 0.5 * (pk[1:P].min() + pk[1:P].max())
    if pk[1:P].max() - pk[1:P].min() > 1e-3:
        assert orc.ostrovsky_cells(*args, mid, 0, 1)[0] == 0


def test_ostrovsky_peakless_cells_follow_the_last_sample(orc):
    """A monotonically decaying linear trajectory has no peaks: flag 3 makes
    the output +1 when the last recorded sample is above the threshold, else
    -1, and flag 0 (no increasing pair of peaks, vacuously true) halves it.
    The initial value is swept (x0 in [0.5, 2]) and the last sample has the
    closed form x0 * (1 + h*a)^(skip + N - 1)."""
    h, a, skip, N = 0.01, -0.1, 100, 400
    lam = [a, a, a]
    n = 6
    x0 = np.array([orc.value_by_idx(i, n, 0.5, 2.0, 0) for i in range(n)])
    last = x0 * (1 + h * a) ** (skip + N - 1)
    thr = float(np.median(last))
    out = orc.ostrovsky_cells(N * h, n, h, [1.0, 0.0, 0.0], [0.5, 2.0, 0.0, 0.0], [0, 1], 0,
                              0.0, 100000, skip * h, lam, 1, 0.01, thr, 0, n,
                              model=orc.MODEL_LINEAR)
    np.testing.assert_array_equal(out, np.where(last > thr, 0.5, -0.5))


# ---------------------------------------------------- single-system series

def test_denis_series_linear_closed_form(orc):
    """Sample j = idx + i*T comes from trajectory idx after idx steps of h
    and i steps of hS: x0 (1+h a)^idx (1+hS a)^i on the linear model; the
    samples past T*iters stay 0."""
    h, hS, tMax, a = 0.01, 0.05, 1.03, -0.3
    T, iters, n = int(hS / h), int(tMax / hS), int(tMax / h)
    out = orc.denis(tMax, h, hS, [2.0, 0.0, 0.0], 0, [a, a, a], model=orc.MODEL_LINEAR)
    assert len(out) == n
    for j in range(n):
        idx, i = j % T, j // T
        if i < iters:
            want = 2.0 * (1 + h * a) ** idx * (1 + hS * a) ** i
            assert abs(out[j] - want) <= 1e-13 * abs(want), j
        else:
            assert out[j] == 0
