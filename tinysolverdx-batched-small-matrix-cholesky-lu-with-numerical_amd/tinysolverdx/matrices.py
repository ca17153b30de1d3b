"""Seeded test-matrix generators with a prescribed 2-norm condition number.

Same method as the reference's python/tinysolverdx/matrices.py:5-21: Haar
orthogonal factors from the QR of a Gaussian matrix with the signs of R's
diagonal folded in, and a geometric spectrum between 1 and 1/kappa.  The
*_kappas variants take one condition number per matrix (used by the
mixed-precision workloads, whose kappa is drawn per matrix).
"""
import numpy as np


def orthogonal(rng, batch, n):
    """Haar-distributed orthogonal matrices (batch, n, n)."""
    z = rng.standard_normal((batch, n, n))
    q, r = np.linalg.qr(z)
    d = np.sign(np.diagonal(r, axis1=1, axis2=2))
    d[d == 0] = 1.0
    return q * d[:, None, :]


def _spectrum(kappas, n):
    kappas = np.asarray(kappas, dtype=np.float64).reshape(-1, 1)
    t = np.linspace(0.0, 1.0, n)[None, :]
    return kappas ** (-t)  # 1 ... 1/kappa, geometric


def spd(rng, batch, n, kappa):
    """SPD matrices Q diag(lambda) Q^T with eigenvalues geometric in [1/kappa, 1], symmetrized."""
    return spd_kappas(rng, np.full(batch, float(kappa)), n)


def spd_kappas(rng, kappas, n):
    lam = _spectrum(kappas, n)
    q = orthogonal(rng, len(lam), n)
    a = np.einsum("bij,bj,bkj->bik", q, lam, q)
    return (a + a.transpose(0, 2, 1)) / 2


def general(rng, batch, n, kappa):
    """General matrices U diag(sigma) V^T with singular values geometric in [1/kappa, 1]."""
    return general_kappas(rng, np.full(batch, float(kappa)), n)


def general_kappas(rng, kappas, n):
    s = _spectrum(kappas, n)
    u = orthogonal(rng, len(s), n)
    v = orthogonal(rng, len(s), n)
    return np.einsum("bij,bj,bkj->bik", u, s, v)


# ----------------------------------------------------------- on the device
def haar_device(z):
    """Orthogonal factor of the QR of z (batch, n, n), R's diagonal made
    positive, by classical Gram-Schmidt with one re-orthogonalisation pass
    (CGS2), batched with torch on z's device."""
    import torch
    m, n, _ = z.shape
    q = torch.empty_like(z)
    for j in range(n):
        v = z[:, :, j]
        if j:
            qj = q[:, :, :j]
            for _ in range(2):
                c = (qj * v[:, :, None]).sum(dim=1)
                v = v - torch.matmul(qj, c[:, :, None])[:, :, 0]
        q[:, :, j] = v / torch.linalg.vector_norm(v, dim=1, keepdim=True)
    return q


def spd_device(n, batch, kmin, kmax, seed, device="cuda", chunk=16384):
    """float64 SPD matrices on the device: Q diag(lambda) Q^T symmetrised,
    lambda_0 = 1, lambda_{n-1} = 1/kappa, the rest log-uniform in between,
    kappa log-uniform on [kmin, kmax] per matrix.  Returns (A, kappa)."""
    import math
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    A = torch.empty((batch, n, n), dtype=torch.float64, device=device)
    lo, hi = math.log10(kmin), math.log10(kmax)
    kappa = 10.0 ** (lo + (hi - lo) * torch.rand(batch, generator=g, device=device, dtype=torch.float64))
    for s in range(0, batch, chunk):
        e = min(batch, s + chunk)
        m = e - s
        q = haar_device(torch.randn((m, n, n), generator=g, device=device, dtype=torch.float64))
        t = torch.rand((m, n), generator=g, device=device, dtype=torch.float64)
        t[:, 0] = 0.0
        t[:, n - 1] = 1.0
        lam = kappa[s:e, None] ** (-t)
        a = torch.matmul(q * lam[:, None, :], q.transpose(1, 2).contiguous())
        A[s:e] = 0.5 * (a + a.transpose(1, 2))
    return A, kappa
