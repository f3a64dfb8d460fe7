"""Synthetic sparse systems of the benchmark configurations (host, numpy).

Every generator returns ``(indptr, indices, vals, dims)``: a CSR matrix in the
natural grid ordering ``i = x + y*nx + z*nx*ny`` with sorted, de-duplicated
rows (the state ``DistSparseMatrix::StopAssembly`` leaves, reference
include/clique/core/dist_sparse_matrix/impl.hpp:209-257), and the grid
dimensions needed by natural nested dissection.

Stencils follow the reference drivers: the 7-point Laplacian of
tests/Solve.cpp:59-81 and the PML-stretched Helmholtz operator of
examples/HelmholtzPML3D.cpp:95-148.
"""
from __future__ import annotations

import numpy as np


def _assemble(N, rows, cols, vals):
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    key = rows.astype(np.int64) * N + cols
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, start) if len(vals) else vals
    r = (uniq // N).astype(np.int64)
    c = (uniq % N).astype(np.int32)
    indptr = np.zeros(N + 1, np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr, c, summed


def _grid(nx, ny, nz):
    i = np.arange(nx * ny * nz, dtype=np.int64)
    return i, i % nx, (i // nx) % ny, i // (nx * ny)


def stencil7(nx, ny, nz, diag, off=-1.0):
    """Constant-coefficient 7-point (5-point if nz == 1) stencil, Dirichlet."""
    N = nx * ny * nz
    i, x, y, z = _grid(nx, ny, nz)
    rows = [i]
    cols = [i]
    vals = [np.full(N, diag, dtype=np.result_type(diag, off))]
    for cond, step in ((x > 0, -1), (x < nx - 1, 1), (y > 0, -nx), (y < ny - 1, nx),
                       (z > 0, -nx * ny), (z < nz - 1, nx * ny)):
        idx = i[cond]
        rows.append(idx)
        cols.append(idx + step)
        vals.append(np.full(len(idx), off, dtype=vals[0].dtype))
    indptr, ind, v = _assemble(N, np.concatenate(rows), np.concatenate(cols),
                               np.concatenate(vals))
    return indptr, ind, v, (nx, ny, nz)


def laplacian7(n, shift=6.05):
    """Config 1: 3D 7-point operator with diagonal `shift`, off-diagonal -1."""
    return stencil7(n, n, n, float(shift), -1.0)


def laplacian5(nx, ny, s):
    """Config 3 member: 2D 5-point, diagonal 4+s, off-diagonal -1."""
    return stencil7(nx, ny, 1, 4.0 + float(s), -1.0)


def wave_speed(x, y, z):
    return 1.0 + 0.3 * np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / 0.02)


def helmholtz7(n, omega=2 * np.pi * 6, damping=0.01):
    """Config 2: complex-symmetric Helmholtz on n^3, h = 1/(n+1), node
    coordinate (i+1)h; diagonal 6/h^2 - (omega/c)^2 (1 + damping i)."""
    N = n ** 3
    h = 1.0 / (n + 1)
    i, x, y, z = _grid(n, n, n)
    c = wave_speed((x + 1) * h, (y + 1) * h, (z + 1) * h)
    diag = 6.0 / h ** 2 - (omega / c) ** 2 * (1 + 1j * damping)
    rows, cols, vals = [i], [i], [diag.astype(np.complex128)]
    offv = -1.0 / h ** 2
    for cond, step in ((x > 0, -1), (x < n - 1, 1), (y > 0, -n), (y < n - 1, n),
                       (z > 0, -n * n), (z < n - 1, n * n)):
        idx = i[cond]
        rows.append(idx)
        cols.append(idx + step)
        vals.append(np.full(len(idx), offv, np.complex128))
    indptr, ind, v = _assemble(N, np.concatenate(rows), np.concatenate(cols),
                               np.concatenate(vals))
    return indptr, ind, v, (n, n, n)


def helmholtz_pml(n, omega=2 * np.pi * 10, pml=10, strength=20.0):
    """Config 5: PML-stretched complex Helmholtz on n^3 in the form of
    examples/HelmholtzPML3D.cpp:102-148, with the stretching
    s(d) = 1 + strength*i*(d/L)^2 over a `pml`-cell layer on every face
    (d = depth into the layer, L = layer width) and c(x) as in config 2."""
    N = n ** 3
    h = 1.0 / (n + 1)
    width = pml * h

    def s_inv(j):
        # the profile value the reference's sInv() returns at grid index j
        # (HelmholtzPML3D.cpp:26-35; j may be -1 or n just outside the grid)
        j = np.asarray(j, dtype=np.float64)
        d = np.zeros_like(j)
        lo = j < pml - 1
        hi = j > n - pml
        d[lo] = (pml - 1 - j[lo]) * h
        d[hi] = (j[hi] - (n - pml)) * h
        return 1.0 + 1j * strength * (d / width) ** 2

    i, x, y, z = _grid(n, n, n)
    s1 = [s_inv(x - 1), s_inv(x), s_inv(x + 1)]
    s2 = [s_inv(y - 1), s_inv(y), s_inv(y + 1)]
    s3 = [s_inv(z - 1), s_inv(z), s_inv(z + 1)]
    h2 = h * h

    def terms(sa, top):
        tl, tm, tr = top / sa[0], top / sa[1], top / sa[2]
        return (tl + tm) / (2 * h2), (tm + tr) / (2 * h2)

    xl, xr = terms(s1, s2[1] * s3[1])
    yl, yr = terms(s2, s1[1] * s3[1])
    zl, zr = terms(s3, s1[1] * s2[1])
    c = wave_speed((x + 1) * h, (y + 1) * h, (z + 1) * h)
    k2 = (omega / c) ** 2
    main = xl + xr + yl + yr + zl + zr - k2 * s1[1] * s2[1] * s3[1]
    rows, cols, vals = [i], [i], [main]
    for cond, step, t in ((x > 0, -1, xl), (x < n - 1, 1, xr), (y > 0, -n, yl),
                          (y < n - 1, n, yr), (z > 0, -n * n, zl), (z < n - 1, n * n, zr)):
        idx = i[cond]
        rows.append(idx)
        cols.append(idx + step)
        vals.append(-t[cond])
    indptr, ind, v = _assemble(N, np.concatenate(rows), np.concatenate(cols),
                               np.concatenate(vals).astype(np.complex128))
    return indptr, ind, v, (n, n, n)


def gaussian_random_field(n, length=0.1, sigma=1.0, seed=0):
    """Squared-exponential Gaussian random field on an n^3 grid of spacing
    1/(n+1) via circulant embedding on the doubled grid (FFT)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n + 1)
    m = 2 * n
    k = np.minimum(np.arange(m), m - np.arange(m)) * h
    r2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    cov = sigma ** 2 * np.exp(-r2 / (2 * length ** 2))
    lam = np.real(np.fft.fftn(cov))
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal((m, m, m)) + 1j * rng.standard_normal((m, m, m))
    field = np.fft.ifftn(np.sqrt(lam) * noise) * np.sqrt(m ** 3)
    return np.real(field[:n, :n, :n])


def diffusion_grf(n, seed=0, length=0.1, sigma=1.0):
    """Config 4: -div(kappa grad u) on n^3, 7-point, face kappa = exp(g) with
    g a Gaussian random field (value at a face = mean of its two nodes)."""
    g = gaussian_random_field(n, length, sigma, seed)  # indexed [x, y, z]
    kap = np.exp(g)
    N = n ** 3
    i, x, y, z = _grid(n, n, n)
    kv = kap[x, y, z]
    rows, cols, vals = [], [], []
    diag = np.zeros(N)
    for axis, step, coord, size in ((0, 1, x, n), (1, n, y, n), (2, n * n, z, n)):
        # face between i and i+step
        has_up = coord < size - 1
        iu = i[has_up]
        xu, yu, zu = x[has_up], y[has_up], z[has_up]
        nb = (xu + (axis == 0), yu + (axis == 1), zu + (axis == 2))
        kf = 0.5 * (kv[has_up] + kap[nb])
        rows += [iu, iu + step]
        cols += [iu + step, iu]
        vals += [-kf, -kf]
        np.add.at(diag, iu, kf)
        np.add.at(diag, iu + step, kf)
        # Dirichlet boundary faces
        lo = coord == 0
        hi = coord == size - 1
        np.add.at(diag, i[lo], kv[lo])
        np.add.at(diag, i[hi], kv[hi])
    rows.append(i)
    cols.append(i)
    vals.append(diag)
    indptr, ind, v = _assemble(N, np.concatenate(rows), np.concatenate(cols),
                               np.concatenate(vals))
    return indptr, ind, v, (n, n, n)


def to_scipy(indptr, indices, vals, N=None):
    import scipy.sparse as sp
    N = len(indptr) - 1 if N is None else N
    return sp.csr_matrix((vals, indices, indptr), shape=(N, N))


def random_rhs(N, k, seed=0, complex_=False):
    rng = np.random.default_rng(seed)
    if complex_:
        return np.asfortranarray(rng.standard_normal((N, k)) + 1j * rng.standard_normal((N, k)))
    return np.asfortranarray(rng.standard_normal((N, k)))
