"""CPU reference of the surrogate network (TEST INFRASTRUCTURE ONLY).

numpy float64 forward/backward of the network the library trains on the
GPU: 64 Fourier features of a lattice point -> 3 x 128 ReLU -> a 32-wide
head (column 0 LLE, columns 1..16 class logits).  With bf16=True every
tensor the GPU stores in BF16 (features, weights, hidden activations, head
gradient, hidden gradients) is rounded to BF16 at the same point; with
bf16=False it is exact float64 and its gradient is pinned by central
finite differences (tests/test_mlp_oracle.py).  Only tests/ load this.

Parameter layout (flat): W1[128x64] b1[128] W2[128x128] b2 W3[128x128] b3
Wh[32x128] bh[32], matrices row-major [out][in].
"""
import numpy as np

FEAT, HID, HEAD, CLASSES = 64, 128, 32, 16
SHAPES = [("W1", (HID, FEAT)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("Wh", (HEAD, HID)), ("bh", (HEAD,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = u.view(np.float32).astype(np.float64)
    return np.where(np.isfinite(f), r, f.astype(np.float64))


def features(cells, n_lat, fourier):
    cells = np.asarray(cells, dtype=np.int64)
    inv = 1.0 / (n_lat - 1)
    u = (cells % n_lat) * inv
    v = (cells // n_lat) * inv
    B = np.asarray(fourier, dtype=np.float64).reshape(FEAT // 2, 2)
    ph = 2 * np.pi * (np.outer(u, B[:, 0]) + np.outer(v, B[:, 1]))
    return np.concatenate([np.sin(ph), np.cos(ph)], axis=1)


CHAOS_COL = 1 + CLASSES


def forward_backward(flat, fourier, cells, n_lat, cls, lle, w_lle=1.0, bf=True, chaos_std=0.0):
    """Losses (mean CE, mean squared LLE error, mean chaos BCE), head outputs
    and the gradient of CE + w_lle * MSE + BCE with respect to the flat
    parameters.  The chaos target is (lle > chaos_std), NaN counting as 0."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(flat)
    W = {k: (r(P[k]) if k.startswith("W") else P[k]) for k in P}
    n = len(cells)
    x0 = r(features(cells, n_lat, fourier))
    h1 = r(np.maximum(x0 @ W["W1"].T + P["b1"], 0))
    h2 = r(np.maximum(h1 @ W["W2"].T + P["b2"], 0))
    h3 = r(np.maximum(h2 @ W["W3"].T + P["b3"], 0))
    out = h3 @ W["Wh"].T + P["bh"]
    logits = out[:, 1:1 + CLASSES]
    mx = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - mx)
    p = e / e.sum(axis=1, keepdims=True)
    cls = np.asarray(cls, dtype=np.int64)
    ce = -(logits[np.arange(n), cls] - mx[:, 0] - np.log(e.sum(axis=1)))
    lle = np.asarray(lle, dtype=np.float64)
    fin = np.isfinite(lle)
    err = np.where(fin, out[:, 0] - np.where(fin, lle, 0), 0)
    z = out[:, CHAOS_COL]
    tc = np.where(fin, lle > chaos_std, False).astype(np.float64)
    pc = 1.0 / (1.0 + np.exp(-z))
    bce = np.maximum(z, 0) - z * tc + np.log1p(np.exp(-np.abs(z)))
    dout = np.zeros((n, HEAD))
    dout[:, 1:1 + CLASSES] = (p - np.eye(CLASSES)[cls]) / n
    dout[:, 0] = 2 * w_lle * err / n
    dout[:, CHAOS_COL] = (pc - tc) / n
    dout = r(dout)
    g = {}
    g["Wh"] = dout.T @ h3
    g["bh"] = dout.sum(axis=0)
    d3 = r((dout @ W["Wh"]) * (h3 > 0))
    g["W3"] = d3.T @ h2
    g["b3"] = d3.sum(axis=0)
    d2 = r((d3 @ W["W3"]) * (h2 > 0))
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(axis=0)
    d1 = r((d2 @ W["W2"]) * (h1 > 0))
    g["W1"] = d1.T @ x0
    g["b1"] = d1.sum(axis=0)
    losses = np.array([ce.mean(), (err ** 2).sum() / n, bce.mean()])
    return losses, out, pack(g)


def loss_only(flat, fourier, cells, n_lat, cls, lle, w_lle=1.0, chaos_std=0.0):
    losses, _, _ = forward_backward(flat, fourier, cells, n_lat, cls, lle, w_lle, bf=False,
                                    chaos_std=chaos_std)
    return losses[0] + w_lle * losses[1] + losses[2]


def adam_step(p, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, keep=None):
    """One Adam step (Kingma & Ba, 2015, Algorithm 1) in float64 at step t
    (1-based): m <- b1 m + (1-b1) g; v <- b2 v + (1-b2) g^2;
    p <- p - lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).  Entries with
    keep == 0 are held at zero with m and v unchanged (the padded head rows
    of the network).  Returns new (p, m, v)."""
    p = np.asarray(p, dtype=np.float64).copy()
    m = np.asarray(m, dtype=np.float64).copy()
    v = np.asarray(v, dtype=np.float64).copy()
    g = np.asarray(grad, dtype=np.float64)
    live = np.ones(p.shape, dtype=bool) if keep is None else np.asarray(keep, dtype=bool)
    m[live] = b1 * m[live] + (1 - b1) * g[live]
    v[live] = b2 * v[live] + (1 - b2) * g[live] ** 2
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p[live] = p[live] - lr * mh[live] / (np.sqrt(vh[live]) + eps)
    p[~live] = 0.0
    return p, m, v
