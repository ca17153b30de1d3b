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
