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
