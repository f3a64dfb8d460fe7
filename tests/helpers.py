"""Shared helpers of the test suite (problems, oracle runs, error measures)."""
import numpy as np

from clique_amd import problems as P


def backward_error(indptr, indices, vals, x, b):
    """eta_j = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) per column."""
    A = P.to_scipy(indptr, indices, vals)
    X = np.asarray(x).reshape(A.shape[0], -1)
    B = np.asarray(b).reshape(A.shape[0], -1)
    R = B - A @ X
    normA = abs(A).sum(axis=1).max()
    return np.abs(R).max(axis=0) / (normA * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))


def rel_diff(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def problem(kind, n, **kw):
    if kind == "lap":
        return P.laplacian7(n)
    if kind == "box":
        nx, ny, nz = n
        return P.stencil7(nx, ny, nz, kw.get("diag", 6.05), -1.0)
    if kind == "helm":
        return P.helmholtz7(n)
    if kind == "pml":
        return P.helmholtz_pml(n, pml=kw.get("pml", 4))
    if kind == "lap2d":
        return P.laplacian5(n[0], n[1], kw.get("s", 0.1))
    if kind == "grf":
        return P.diffusion_grf(n, seed=kw.get("seed", 0))
    raise ValueError(kind)


def hermitian_matrix(n, seed=0):
    """A Hermitian positive definite 7-point operator with complex couplings."""
    ip, ind, v, dims = P.stencil7(n, n, n, 7.0, -1.0)
    rng = np.random.default_rng(seed)
    N = ip.size - 1
    rows = np.repeat(np.arange(N), np.diff(ip))
    vals = v.astype(np.complex128)
    phase = np.exp(1j * rng.uniform(0, 2 * np.pi, size=len(vals)))
    # make the coupling of (i,j) the conjugate of (j,i)
    key = {}
    for e in range(len(vals)):
        i, j = rows[e], ind[e]
        if i < j:
            key[(i, j)] = phase[e]
    for e in range(len(vals)):
        i, j = rows[e], ind[e]
        if i < j:
            vals[e] = v[e] * key[(i, j)]
        elif i > j:
            vals[e] = v[e] * np.conj(key[(j, i)])
    return ip, ind, vals, dims
