"""Shared helpers of the test suite (problems, oracle runs, error measures)."""
import numpy as np
import scipy.sparse as sp

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


def indefinite_box(dims, seed=0, cplx=False, herm=False):
    """A 7-point operator with a weak diagonal of random sign: indefinite, so
    Bunch-Kaufman takes swaps and 2x2 pivots."""
    ip, ind, v, dims = P.stencil7(*dims, 1.0, -1.0)
    rng = np.random.default_rng(seed)
    N = ip.size - 1
    rows = np.repeat(np.arange(N), np.diff(ip))
    v = v.astype(np.complex128) if cplx else v.copy()
    diag = rows == ind
    v[diag] = rng.uniform(-0.6, 0.6, N) + (1j * rng.uniform(-0.3, 0.3, N) if cplx and not herm else 0)
    if herm:
        # complex couplings, Hermitian: (j,i) holds the conjugate of (i,j)
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi, ind.size))
        A = sp.csr_matrix((v * np.where(diag, 1, ph), ind, ip), shape=(N, N))
        A = sp.triu(A, 1) + sp.triu(A, 1).conj().T + sp.diags(A.diagonal().real)
        A = A.tocsr()
        A.sort_indices()
        return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, dims
    return ip, ind, v, dims


def indefinite_random(dims, seed=0, cplx=False, herm=False):
    """A 7-point pattern with random couplings of magnitude in [0.5, 1.5]
    (random phases when complex) and a random diagonal in [-0.6, 0.6]:
    indefinite, and without the exact magnitude ties of a constant stencil,
    so the Bunch-Kaufman pivot choices do not hinge on rounding."""
    ip, ind, v, dims = P.stencil7(*dims, 1.0, -1.0)
    rng = np.random.default_rng(seed)
    N = ip.size - 1
    A = sp.csr_matrix((np.ones(ind.size), ind, ip), shape=(N, N))
    U = sp.triu(A, 1).tocoo()
    mag = rng.uniform(0.5, 1.5, U.nnz)
    if cplx:
        vals = mag * np.exp(1j * rng.uniform(0, 2 * np.pi, U.nnz))
    else:
        vals = mag * np.where(rng.uniform(size=U.nnz) < 0.5, -1.0, 1.0)
    U = sp.csr_matrix((vals, (U.row, U.col)), shape=(N, N))
    d = rng.uniform(-0.6, 0.6, N)
    if cplx and not herm:
        d = d + 1j * rng.uniform(-0.3, 0.3, N)
    L = U.conj().T if herm else U.T
    M = (U + L + sp.diags(d)).tocsr()
    M.sort_indices()
    return M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data, dims
