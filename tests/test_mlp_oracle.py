"""Pins the CPU reference of the surrogate network (CPU only).

The analytic gradient of the exact (float64, no BF16 rounding) reference
must match central finite differences of its own loss on a sample of
parameters of every tensor; BF16 mode must stay close to exact mode.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
import mlp_ref as M  # noqa: E402


def _setup(seed=0, n=48, n_lat=32):
    rng = np.random.default_rng(seed)
    flat = np.zeros(M.N_PARAMS)
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        fan_in = shape[1] if len(shape) == 2 else 1
        flat[o:o + k] = rng.normal(size=k) * (np.sqrt(2.0 / fan_in) if len(shape) == 2 else 0.1)
        o += k
    fourier = rng.normal(size=64) * 3.0
    cells = rng.integers(0, n_lat * n_lat, size=n)
    cls = rng.integers(0, 16, size=n)
    lle = rng.normal(size=n)
    lle[::7] = np.nan
    return flat, fourier, cells, n_lat, cls, lle


def test_exact_gradient_matches_finite_differences():
    flat, fourier, cells, n_lat, cls, lle = _setup()
    _, _, g = M.forward_backward(flat, fourier, cells, n_lat, cls, lle, 0.7, bf=False)
    rng = np.random.default_rng(1)
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        for idx in rng.choice(k, size=min(6, k), replace=False):
            i = o + idx
            h = 1e-6
            fp, fm = flat.copy(), flat.copy()
            fp[i] += h
            fm[i] -= h
            fd = (M.loss_only(fp, fourier, cells, n_lat, cls, lle, 0.7) -
                  M.loss_only(fm, fourier, cells, n_lat, cls, lle, 0.7)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, g[i])
        o += k


def test_bf16_mode_close_to_exact():
    flat, fourier, cells, n_lat, cls, lle = _setup(2)
    l_ex, out_ex, g_ex = M.forward_backward(flat, fourier, cells, n_lat, cls, lle, bf=False)
    l_bf, out_bf, g_bf = M.forward_backward(flat, fourier, cells, n_lat, cls, lle, bf=True)
    assert np.allclose(l_ex, l_bf, rtol=5e-2)
    cos = g_ex @ g_bf / np.linalg.norm(g_ex) / np.linalg.norm(g_bf)
    assert cos > 0.99


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, 0.0])
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie rounds to even
    assert r[2] == 1.0078125                     # above the tie rounds up
    assert abs(r[3] + 3.0e-3) < 2e-5 and r[5] == 0.0


def test_adam_reference_closed_forms():
    """With a constant gradient g the bias-corrected moments are exactly g
    and g^2 at every step, so each step moves p by lr*g/(|g|+eps); with a
    gradient that changes, m_t = (1-b1) sum_k b1^(t-k) g_k and v_t likewise.
    Held entries stay zero."""
    rng = np.random.default_rng(3)
    n, lr, b1, b2, eps = 64, 1e-3, 0.9, 0.999, 1e-8
    g = rng.normal(size=n) * np.logspace(-12, 2, n)
    p0 = rng.normal(size=n)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    for t in range(1, 8):
        p, m, v = M.adam_step(p, m, v, g, t, lr, b1, b2, eps)
        np.testing.assert_allclose(p, p0 - t * lr * g / (np.abs(g) + eps), rtol=1e-12,
                                   atol=1e-15)
    gs = rng.normal(size=(6, n))
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    for t in range(1, 7):
        p_prev = p
        p, m, v = M.adam_step(p, m, v, gs[t - 1], t, lr, b1, b2, eps)
        mt = (1 - b1) * sum(b1 ** (t - k) * gs[k - 1] for k in range(1, t + 1))
        vt = (1 - b2) * sum(b2 ** (t - k) * gs[k - 1] ** 2 for k in range(1, t + 1))
        np.testing.assert_allclose(m, mt, rtol=1e-12)
        np.testing.assert_allclose(v, vt, rtol=1e-12)
        step = lr * (mt / (1 - b1 ** t)) / (np.sqrt(vt / (1 - b2 ** t)) + eps)
        np.testing.assert_allclose(p, p_prev - step, rtol=1e-12)
    keep = np.arange(n) % 5 != 0
    p, m, v = M.adam_step(p0, np.zeros(n), np.zeros(n), g, 1, keep=keep)
    assert not p[~keep].any() and not m[~keep].any() and not v[~keep].any()
