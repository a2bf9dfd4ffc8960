"""Pins of the oracle functions whose rules only the reference defines.

KDE (kde, Library/cudaLibrary.cu:356-422).  Three independent checks:
  * the values the count compares (orc_kde_sums) are the Gaussian KDE of
    the peaks at the k1-1 points w*(I2-I1)/(k1-1) + I1, w < k1-1 (so I2
    itself is never sampled), evaluated as the float exp of the
    float-rounded exponent;
  * the count is a property of that sequence of values alone (_runs_count):
    split it into runs of equal values; a run other than the first counts
    when it is above the run before it and either above the run after it
    or, as the last run, a single value.  This is what the reference's
    loop computes -- the w < 2 skip leaves a maximum at the first value
    uncounted, the strange-peak state counts a plateau once when it falls
    and not when it rises, and the tail rule counts a single last value
    above its neighbour -- stated without its state machine, and checked
    on a fuzz set that holds exact plateaus and exact zeros;
  * hand-built cases whose answers follow from the shape of the values:
    with kdeSampling 1 and three peaks only two points are sampled, so the
    answer is [S1 > S0]; identical peaks midway between two points of a
    power-of-two step give two bitwise-equal values (a plateau); peaks
    more than ~15 smoothing widths from a point add an exact 0.

Lyapunov spectrum reference directions (LSKernelCUDA,
Library/cudaLibrary.cu:1025-1040): one renormalisation of the diagonal
linear model from the zero state maps the orthonormalised directions Q
to M^n Q, so the norms are |diag R| of the QR factorisation of M^n Q^T,
with Z built from 0.5*sinf(0.2171828(i+1)(j+1) + idx + (0.2171828 +
i*j*idx)) here in numpy; the norms depend on Z, so a wrong direction
formula moves them.

tests/test_oracle_mutants.py plants a mistake in a copy of the oracle for
each rule and checks that one of these checks fails on it.
"""
import math

import numpy as np


# ----------------------------------------------------------------- KDE

def _runs_count(S):
    S = list(S)
    runs = []                        # [value, length]
    for v in S:
        if runs and runs[-1][0] == v:
            runs[-1][1] += 1
        else:
            runs.append([v, 1])
    count = 0
    for r in range(1, len(runs)):
        v, n = runs[r]
        if not v > runs[r - 1][0]:
            continue
        if r + 1 < len(runs):
            count += v > runs[r + 1][0]
        else:
            count += n == 1
    return count


def _kde_definition(peaks, sampling, i1, i2, H):
    p = np.asarray(peaks, dtype=np.float64)
    k1 = sampling * len(p)
    n = max(0, int(math.ceil(k1 - 1)))
    if n == 0:
        return np.zeros(0)
    k2 = (i2 - i1) / (k1 - 1)
    pts = np.arange(n) * k2 + i1
    t = (p[None, :] - pts[:, None]) / H
    arg = (-(t * t) / 2).astype(np.float32).astype(np.float64)
    return np.exp(arg).sum(axis=1)


def _fuzz_cases(seed=3, n=2500):
    """Peak sets on a power-of-two sample step: groups of identical peaks at
    sample points or midway between two, far enough apart that a group's
    far terms are exact zeros; some cases overlap groups or add random
    peaks."""
    rng = np.random.default_rng(seed)
    cases = []
    for c in range(n):
        P = int(rng.integers(3, 10))
        s = int(rng.integers(1, 4))
        k1 = s * P
        d = 2.0 ** -int(rng.integers(1, 5))
        i1 = float(rng.choice([0.0, -1.0, 0.5]))
        i2 = i1 + (k1 - 1) * d
        H = d * float(rng.choice([0.125, 0.25, 0.5, 1.0])) if c % 4 else d * rng.random() * 2
        peaks = []
        while len(peaks) < P:
            g = int(rng.integers(1, P - len(peaks) + 1))
            w = int(rng.integers(-2, k1 + 2))
            pos = i1 + (w + (0.5 if rng.random() < 0.5 else 0.0)) * d
            if rng.random() < 0.15:
                pos = i1 + rng.random() * (i2 - i1)
            peaks += [pos] * g
        rng.shuffle(peaks)
        cases.append((peaks, s, i1, i2, H))
    return cases


def check_kde_sums_are_the_kde(orc):
    rng = np.random.default_rng(1)
    for peaks, s, i1, i2, H in _fuzz_cases(seed=4, n=300) + [
            (list(rng.random(7)), 5, 0.0, 1.0, 0.05), ([0.2, 0.9, 0.4], 1, 0.1, 0.7, 0.3)]:
        got = orc.kde_sums(peaks, s, i1, i2, H)
        want = _kde_definition(peaks, s, i1, i2, H)
        assert len(got) == len(want) == s * len(peaks) - 1, (len(got), len(want))
        np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-44 * len(peaks))


def check_kde_count_is_the_run_rule(orc):
    plateaus = zeros = 0
    for peaks, s, i1, i2, H in _fuzz_cases():
        S = orc.kde_sums(peaks, s, i1, i2, H)
        got = orc.kde(peaks, 1000, s, i1, i2, H)
        assert got == _runs_count(S), (peaks, s, i1, i2, H, list(S), got)
        plateaus += any(S[k] == S[k + 1] and S[k] > 0 for k in range(len(S) - 1))
        zeros += any(v == 0 for v in S)
    # the fuzz set really exercises the equality rules
    assert plateaus > 300 and zeros > 600, (plateaus, zeros)


def check_kde_hand_cases(orc):
    # kdeSampling 1, P = 3: two points, I1 and the middle; answer [S1 > S0]
    assert orc.kde([0.0, 0.0, 1.0], 10, 1, 0.0, 1.0, 0.05) == 0
    assert orc.kde([0.5, 0.5, 0.5], 10, 1, 0.0, 1.0, 0.05) == 1
    # plateau that falls: points w/8 (w < 8); three peaks at 3.5/8 give
    # S3 == S4 bitwise; one maximum
    S = orc.kde_sums([0.4375] * 3, 3, 0.0, 1.0, 0.1)
    assert S[3] == S[4] and S[2] < S[3] and S[5] < S[4]
    assert orc.kde([0.4375] * 3, 10, 3, 0.0, 1.0, 0.1) == 1
    # plateau that rises into a higher single value: points w/4 (w < 7);
    # three peaks at 2.5/4 (a plateau of 1e-34 at w = 2, 3, exact zeros
    # elsewhere) and one at 4/4; only the later value is a maximum
    pk = [0.625, 0.625, 1.0, 0.625]
    S = orc.kde_sums(pk, 2, 0.0, 1.75, 0.01)
    assert S[0] == S[1] == 0 and S[2] == S[3] > 0 and S[4] > S[3] and S[5] == S[6] == 0
    assert orc.kde(pk, 10, 2, 0.0, 1.75, 0.01) == 1
    # every value an exact zero: no maximum
    S = orc.kde_sums([5.0, 5.0, 6.0], 2, 0.0, 1.0, 0.01)
    assert not S.any()
    assert orc.kde([5.0, 5.0, 6.0], 10, 2, 0.0, 1.0, 0.01) == 0
    # a maximum at the last point counts (tail rule), one at the first
    # point does not, one at the second point does
    assert orc.kde([1.75] * 3, 10, 3, 0.0, 2.0, 0.01) == 1
    assert orc.kde([0.0] * 3, 10, 3, 0.0, 2.0, 0.01) == 0
    assert orc.kde([0.25] * 3, 10, 3, 0.0, 2.0, 0.01) == 1


def test_kde_sums_are_the_kde(orc):
    check_kde_sums_are_the_kde(orc)


def test_kde_count_is_the_run_rule(orc):
    check_kde_count_is_the_run_rule(orc)


def test_kde_hand_cases(orc):
    check_kde_hand_cases(orc)


def test_runs_count_rule_on_small_sequences():
    """The run rule itself on sequences whose count is plain to see."""
    assert _runs_count([]) == 0
    assert _runs_count([1, 2]) == 1          # rising to a single last value
    assert _runs_count([2, 1]) == 0          # first value never counts
    assert _runs_count([1, 2, 2]) == 0       # plateau at the end
    assert _runs_count([1, 2, 2, 1]) == 1    # plateau that falls
    assert _runs_count([1, 2, 2, 3, 1]) == 1  # plateau that rises
    assert _runs_count([0, 3, 0, 3, 0, 2]) == 3


# --------------------------------------------- LS reference directions

LAM_S = [2.0, -3.0, -9.0]


def _ls_dirs(idx, n=3):
    z = np.zeros((n, n))
    for j in range(n):
        for i in range(n):
            arg = 0.2171828 * (i + 1) * (j + 1) + idx + (0.2171828 + i * j * idx)
            z[j, i] = float(np.sin(np.float32(arg))) * 0.5
        z[j] /= np.linalg.norm(z[j])
    return z


def check_ls_reference_directions(orc):
    h, NT, eps = 0.01, 0.5, 1e-6
    nsteps = int(NT / h)
    M = np.diag((1 + h * np.array(LAM_S)) ** nsteps)
    for idx in (0, 1, 2, 7, 123457, 199999):
        out = orc.ls_cells(1, NT, NT, 3, h, eps, [0, 0, 0], [2.0, 2.0], [0], 0.0, 0.0, LAM_S,
                           orc.PERT_REFERENCE, 0, idx, 1, model=orc.MODEL_LINEAR)[0]
        q, _ = np.linalg.qr(_ls_dirs(idx).T)
        _, r = np.linalg.qr(M @ q)
        want = np.log(np.abs(np.diag(r))) / NT
        np.testing.assert_allclose(out, want, atol=2e-6, err_msg=f"idx {idx}")


def test_ls_reference_directions_one_renormalisation(orc):
    check_ls_reference_directions(orc)
