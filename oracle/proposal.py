"""CPU restatement of the learned-proposal network (test infrastructure
only; the product path never imports this).

The network has no counterpart in the reference.  This restatement fixes
what the GPU kernels compute: positional-encoding features, three dense
layers (ReLU between), per-bin binary cross entropy with a positive weight,
its gradient, and Adam.  `bf16=True` rounds at the points where the GPU
rounds to bfloat16 (features, weights, hidden activations, backward
signals); `bf16=False` is the exact FP64 function the finite-difference
test differentiates.  Pinned in tests/test_oracle.py by central finite
differences of the loss and by the bfloat16 rounding rule (round to
nearest even, as the record conversion).
"""
import numpy as np

IN, WIDTH, BINS = 128, 64, 64
W1, W2, W3 = IN * WIDTH, WIDTH * WIDTH, WIDTH * BINS
WEIGHTS = W1 + W2 + W3
PARAMS = WEIGHTS + 3 * WIDTH


def bf16(x):
    """Round float64/float32 values to the nearest bfloat16 (ties to even),
    returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(o, d, tn, dims, f32=False):
    """[n, 128] features: entry point o + max(tn, 0) d mapped to [-1, 1]^3
    (grid dims W, H, D) and direction d, each [x, sin(2^i pi x), cos(2^i pi x)
    for i < 10] per component, ordered as the kernel lays them out.
    f32=True repeats the kernel's float32 operations one by one (the sine
    argument reaches 512 pi, where its float32 rounding shows)."""
    ft = np.float32 if f32 else np.float64
    o, d = np.asarray(o, ft), np.asarray(d, ft)
    t0 = np.maximum(np.asarray(tn, ft), ft(0.0))
    p = o + t0[:, None] * d
    scale = np.array([ft(2.0) / ft(dims[0] - 1), ft(2.0) / ft(dims[1] - 1),
                      ft(2.0) / ft(dims[2] - 1)], ft)
    v = np.concatenate([p * scale - ft(1.0), d], axis=1)  # [n, 6]
    pi = ft(3.14159265358979) if f32 else np.pi
    out = np.zeros((len(o), IN), np.float64)
    for f in range(126):
        g, q = divmod(f, 63)
        if q < 3:
            out[:, f] = v[:, 3 * g + q]
        else:
            i, rem = divmod(q - 3, 6)
            c = rem % 3
            x = (v[:, 3 * g + c] * ft(2 ** i) * pi).astype(np.float64)
            out[:, f] = np.sin(x) if rem < 3 else np.cos(x)
    return out


def split(params):
    p = np.asarray(params, np.float64)
    w1 = p[:W1].reshape(IN, WIDTH)
    w2 = p[W1:W1 + W2].reshape(WIDTH, WIDTH)
    w3 = p[W1 + W2:WEIGHTS].reshape(WIDTH, BINS)
    b1, b2, b3 = p[WEIGHTS:WEIGHTS + 64], p[WEIGHTS + 64:WEIGHTS + 128], p[WEIGHTS + 128:]
    return w1, w2, w3, b1, b2, b3


def forward(params, X, bf16_round=True):
    """Logits [n, 64] and the hidden activations."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    w1, w2, w3, b1, b2, b3 = split(params)
    w1, w2, w3 = rnd(w1), rnd(w2), rnd(w3)
    X = rnd(X)
    h1 = rnd(np.maximum(X @ w1 + b1, 0.0))
    h2 = rnd(np.maximum(h1 @ w2 + b2, 0.0))
    z = h2 @ w3 + b3
    return z, (X, h1, h2)


def loss_grad(params, X, targets, valid=None, pos_weight=1.0, bf16_round=True):
    """(weighted BCE sum, mean-loss gradient [PARAMS]) of one batch."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    n = X.shape[0]
    t = np.asarray(targets, np.uint64)
    y = ((t[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
    use = np.ones(n, bool) if valid is None else np.asarray(valid).astype(bool)
    n_valid = int(use.sum()) if valid is not None else n
    w1, w2, w3, b1, b2, b3 = split(params)
    w1r, w2r, w3r = rnd(w1), rnd(w2), rnd(w3)
    z, (Xr, h1, h2) = forward(params, X, bf16_round)
    p = 1.0 / (1.0 + np.exp(-z))
    wgt = np.where(y > 0, pos_weight, 1.0) * use[:, None]
    loss = float((wgt * (np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z))))).sum())
    scale = 1.0 / (n_valid * 64)
    dz = rnd(scale * wgt * (p - y))
    g = np.zeros(PARAMS)
    gw3 = h2.T @ dz
    gb3 = dz.sum(0)
    dh2 = rnd((dz @ w3r.T) * (h2 > 0))
    gw2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = rnd((dh2 @ w2r.T) * (h1 > 0))
    gw1 = Xr.T @ dh1
    gb1 = dh1.sum(0)
    g[:W1] = gw1.ravel()
    g[W1:W1 + W2] = gw2.ravel()
    g[W1 + W2:WEIGHTS] = gw3.ravel()
    g[WEIGHTS:WEIGHTS + 64] = gb1
    g[WEIGHTS + 64:WEIGHTS + 128] = gb2
    g[WEIGHTS + 128:] = gb3
    return loss, g


def adam(master, g, m, v, lr, t):
    """In place: FP32 moments (0.9, 0.999), bias correction, eps 1e-8."""
    gf = np.asarray(g, np.float32)
    m[:] = np.float32(0.9) * m + np.float32(0.1) * gf
    v[:] = np.float32(0.999) * v + np.float32(0.001) * gf * gf
    c1, c2 = 1.0 - 0.9 ** t, 1.0 - 0.999 ** t
    step = lr * (m.astype(np.float64) / c1) / (np.sqrt(v.astype(np.float64) / c2) + 1e-8)
    master[:] = (master.astype(np.float64) - step).astype(np.float32)
