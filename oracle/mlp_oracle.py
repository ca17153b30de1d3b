"""TEST INFRASTRUCTURE ONLY -- the float64 CPU oracle of the precision router.

Checker for the HIP router kernels (features, forward pass, gradient, Adam),
never part of the product path.  The router has no reference counterpart; this
file pins it with mathematics instead: its own analytic gradient is checked
against central finite differences (tests/test_oracle.py), and the HIP kernels
are checked against it.

Layout of the parameter vector (as include/tsdx_hip.h, TSDX_HIP_ROUTER_PARAMS):
feature means (12), scales (12), W1 (64x12), b1 (64), W2 (64x64), b2 (64),
W3 (3x64), b3 (3), row-major [out][in]; ReLU hidden layers; loss = mean
softmax cross-entropy.
"""
import numpy as np

F, H, T = 12, 64, 3
OFF_W1 = 24
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + T * H
PARAMS = OFF_B3 + T


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[0:12], p[12:24], p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(T, H), p[OFF_B3:PARAMS])


def forward(p, feat):
    mu, sd, W1, b1, W2, b2, W3, b3 = unpack(p)
    x = (np.asarray(feat, np.float64) - mu) / sd
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ W3.T + b3
    return x, h1, h2, z


def loss(p, feat, label):
    z = forward(p, feat)[3]
    zm = z.max(axis=1, keepdims=True)
    lse = zm[:, 0] + np.log(np.exp(z - zm).sum(axis=1))
    return float(np.mean(lse - z[np.arange(len(z)), label]))


def grad(p, feat, label):
    """Analytic gradient of loss(); the first 24 entries are zero."""
    mu, sd, W1, b1, W2, b2, W3, b3 = unpack(p)
    x, h1, h2, z = forward(p, feat)
    B = len(z)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    dz = e / e.sum(axis=1, keepdims=True)
    dz[np.arange(B), label] -= 1.0
    dz /= B
    g = np.zeros(PARAMS)
    g[OFF_W3:OFF_B3] = (dz.T @ h2).ravel()
    g[OFF_B3:PARAMS] = dz.sum(axis=0)
    dh2 = (dz @ W3) * (h2 > 0)
    g[OFF_W2:OFF_B2] = (dh2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    g[OFF_W1:OFF_B1] = (dh1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = dh1.sum(axis=0)
    return g


def finite_difference(p, feat, label, idx, h=1e-6):
    """Central differences of loss() at the parameter indices idx."""
    out = []
    for i in idx:
        q = np.array(p, np.float64)
        q[i] += h
        lp = loss(q, feat, label)
        q[i] -= 2 * h
        lm = loss(q, feat, label)
        out.append((lp - lm) / (2 * h))
    return np.array(out)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    p, m, v = np.array(p, np.float64), np.array(m, np.float64), np.array(v, np.float64)
    s = slice(OFF_W1, PARAMS)
    m[s] = b1 * m[s] + (1 - b1) * g[s]
    v[s] = b2 * v[s] + (1 - b2) * g[s] ** 2
    mh = m[s] / (1 - b1 ** step)
    vh = v[s] / (1 - b2 ** step)
    p[s] -= lr * mh / (np.sqrt(vh) + eps)
    return p, m, v


def features(A):
    """The 12 router features of float64 A (batch, n, n) from its lower triangle."""
    A = np.asarray(A, np.float64)
    out = []
    for a in A:
        n = a.shape[0]
        s = np.tril(a) + np.tril(a, -1).T
        d = np.diag(s).copy()
        r = np.abs(s).sum(axis=1) - np.abs(d)
        fro = np.sqrt((s * s).sum())

        def lg(x):
            return np.log10(x) if x > 1e-30 else -30.0
        f = np.zeros(12)
        f[0] = n / 64.0
        f[1] = lg(d.min())
        f[2] = lg(np.abs(d).max())
        f[3] = lg(fro)
        f[4] = lg((np.abs(d) + r).max())
        f[5] = max((d - r).min() / fro, -30.0) if fro > 0 else 0.0
        f[6] = 30.0 if not (d > 0).all() else lg((r / d).max())
        f[7] = lg(np.abs(d).max() / d.min()) if d.min() > 0 else 30.0
        f[8] = lg(fro / d.mean()) if d.sum() > 0 else 30.0
        f[9] = np.mean([lg(x) for x in d]) - lg(np.abs(d).max())
        f[10] = 1.0 if not (d > 0).all() else 0.0
        f[11] = float((d > r).mean())
        out.append(f)
    return np.array(out)


def round_bf16(x):
    """float32 values rounded to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def infer_bf16(p, feat):
    """Logits of the matrix-core inference path (tsdx_hip_router_infer): the
    feature normalisation folded into layer 1 in float32 (W1' = W1 / scale,
    b1' = b1 - sum_q W1'[:, q] mean_q in the order of q), weights, inputs and
    hidden activations rounded to bfloat16, sums in float64 here (the matrix
    cores accumulate bfloat16 products exactly in float32 in an order of their
    own, so GPU logits agree to float32 rounding of the sums)."""
    p32 = np.asarray(p, np.float32)
    mu, sd = p32[0:12], p32[12:24]
    W1 = p32[OFF_W1:OFF_B1].reshape(H, F)
    W1f = (W1 / sd[None, :]).astype(np.float32)
    b1f = p32[OFF_B1:OFF_W2].copy()
    for q in range(F):
        b1f = (b1f - (W1f[:, q] * mu[q]).astype(np.float32)).astype(np.float32)
    W2 = p32[OFF_W2:OFF_B2].reshape(H, H)
    W3 = p32[OFF_W3:OFF_B3].reshape(T, H)
    b2, b3 = p32[OFF_B2:OFF_W3], p32[OFF_B3:PARAMS]
    x = round_bf16(np.asarray(feat, np.float32)).astype(np.float64)
    a1 = x @ round_bf16(W1f).astype(np.float64).T + b1f
    h1 = round_bf16(np.maximum(a1, 0.0).astype(np.float32)).astype(np.float64)
    a2 = h1 @ round_bf16(W2).astype(np.float64).T + b2
    h2 = round_bf16(np.maximum(a2, 0.0).astype(np.float32)).astype(np.float64)
    return h2 @ round_bf16(W3).astype(np.float64).T + b3


def tiers(z):
    """First largest logit; a NaN logit gives tier 2."""
    z = np.asarray(z)
    t = np.argmax(z, axis=1)
    t[np.isnan(z).any(axis=1)] = 2
    return t.astype(np.int32)
