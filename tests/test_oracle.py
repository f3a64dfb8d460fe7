"""Pins the CPU oracle (CPU only).

No golden vectors exist in the reference (its tests print norms without
thresholds, tests/Solve.cpp:306-322, and its dense kernels live in the
absent Elemental submodule). The oracle is therefore pinned by facts that a
plausible mistake (dropped term, wrong sign/index, transposed operand) breaks:
closed forms of natural nested dissection, brute-force symbolic and numeric
LDL^T of the permuted matrix, Laplacian eigenpairs, the 1D pivot recurrence,
and scipy's sparse direct solver.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from clique_amd import problems as P
from helpers import (backward_error, hermitian_matrix, indefinite_box, indefinite_random, problem,
                     rel_diff)


def tree_of(kind, n, cut, **kw):
    ip, ind, v, dims = problem(kind, n, **kw)
    return ip, ind, v, dims, oracle.Tree(*dims, ip, ind, cutoff=cut)


def fronts(t):
    return [t.node(s) for s in range(t.num)]


# ---------------------------------------------------------------------------
# symbolic: invariants and closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dims,cut", [((6, 6, 6), 8), ((7, 5, 3), 4), ((30, 30, 30), 128),
                                      ((17, 1, 1), 2), ((12, 9, 1), 10), ((2, 1, 1), 1)])
def test_symbolic_invariants(dims, cut):
    ip, ind, v, d, t = tree_of("box", dims, cut)
    N = ip.size - 1
    nodes = fronts(t)
    mp = t.map()
    assert sorted(mp.tolist()) == list(range(N))          # map is a permutation
    assert sum(nd["size"] for nd in nodes) == N
    off = 0
    for s, nd in enumerate(nodes):
        assert nd["off"] == off                            # contiguous nodal ranges
        off += nd["size"]
        for c in (nd["child0"], nd["child1"]):
            if c >= 0:
                assert c < s and nodes[c]["parent"] == s   # postorder
        low = t.array(s, oracle.OR_LOWER)
        assert np.all(np.diff(low) > 0)
        if low.size:
            assert low[0] >= nd["off"] + nd["size"]
        inds = t.array(s, oracle.OR_INDS)
        assert np.array_equal(mp[inds], np.arange(nd["off"], nd["off"] + nd["size"]))
        full = np.concatenate([np.arange(nd["off"], nd["off"] + nd["size"]), low])
        for c, which in ((nd["child0"], oracle.OR_LEFT_REL), (nd["child1"], oracle.OR_RIGHT_REL)):
            if c < 0:
                continue
            rel = t.array(s, which)
            clow = t.array(c, oracle.OR_LOWER)
            assert np.all(np.diff(rel) > 0) and rel.size == clow.size
            assert np.array_equal(full[rel], clow)
        orel = t.array(s, oracle.OR_ORIG_LOWER_REL)
        assert np.array_equal(full[orel], t.array(s, oracle.OR_ORIG_LOWER))
    assert nodes[-1]["parent"] == -1 and nodes[-1]["lower_size"] == 0


def test_natural_dissection_closed_forms():
    # root separator = plane across the longest dimension (ties x, y, z),
    # left box gets (d-1)/2 planes; config 5's root front is 120*120
    for (nx, ny, nz), root in (((30, 30, 30), 900), ((10, 7, 7), 49), ((5, 9, 4), 20),
                               ((120, 120, 120), 14400)):
        ip, ind, v, dims = P.stencil7(nx, ny, nz, 6.0)
        t = oracle.Tree(nx, ny, nz, ip, ind, cutoff=128)
        nd = t.node(t.num - 1)
        assert nd["size"] == root
        inds = t.array(t.num - 1, oracle.OR_INDS)
        x, y, z = inds % nx, (inds // nx) % ny, inds // (nx * ny)
        if nx >= ny and nx >= nz:
            assert np.all(x == (nx - 1) // 2)
        elif ny >= nz:
            assert np.all(y == (ny - 1) // 2)
        else:
            assert np.all(z == (nz - 1) // 2)


def test_benchmark_tree_sizes():
    """Front counts and FP64 factor work of the 30^3 and 64^3 benchmark trees
    (363 fronts and 4.365 GFLOP; 4,095 fronts and 1,696.9 GFLOP complex)."""
    for n, fronts_, gflop in ((30, 363, 4.365), (64, 4095, 1696.9 / 4)):
        ip, ind, v, dims = P.stencil7(n, n, n, 6.0)
        t = oracle.Tree(*dims, ip, ind, cutoff=128)
        assert t.num == fronts_
        assert abs(t.work(False)[0] / 1e9 - gflop) / gflop < 1e-3


def symbolic_cholesky(A):
    """Boolean pattern of L for the dense elimination of A (no cancellation)."""
    Pat = (np.abs(A) > 0)
    N = Pat.shape[0]
    L = np.tril(Pat)
    for j in range(N):
        rows = np.nonzero(L[j + 1:, j])[0] + j + 1
        if rows.size:
            L[np.ix_(rows, rows)] |= np.tril(np.ones((rows.size, rows.size), bool))
    return L


@pytest.mark.parametrize("dims,cut", [((6, 5, 4), 6), ((9, 8, 1), 5), ((5, 5, 5), 16)])
def test_lower_struct_equals_symbolic_cholesky(dims, cut):
    ip, ind, v, d, t = tree_of("box", dims, cut)
    A = P.to_scipy(ip, ind, v).toarray()
    mp = t.map()
    Ap = np.zeros_like(A)
    Ap[np.ix_(mp, mp)] = A
    Lpat = symbolic_cholesky(Ap)
    for s in range(t.num):
        nd = t.node(s)
        cols = range(nd["off"], nd["off"] + nd["size"])
        below = set()
        for c in cols:
            below |= set(np.nonzero(Lpat[:, c])[0].tolist())
        below -= set(cols)
        assert sorted(below) == t.array(s, oracle.OR_LOWER).tolist()


# ---------------------------------------------------------------------------
# numeric: brute force and closed forms
# ---------------------------------------------------------------------------
def dense_ldl(A, conj=False):
    A = A.astype(np.complex128 if np.iscomplexobj(A) else np.float64).copy()
    N = A.shape[0]
    L = np.eye(N, dtype=A.dtype)
    d = np.zeros(N, dtype=A.dtype)
    for j in range(N):
        d[j] = A[j, j]
        l = A[j + 1:, j] / d[j]
        L[j + 1:, j] = l
        A[j + 1:, j + 1:] -= np.outer(l, (np.conj(l) if conj else l)) * d[j]
    return L, d


@pytest.mark.parametrize("kind,n,cut,cplx,conj", [
    ("box", (6, 5, 4), 6, False, False),
    ("box", (9, 8, 1), 5, False, False),
    ("helm", 5, 8, True, False),
    ("herm", 4, 8, True, True),
])
@pytest.mark.parametrize("nb", [3, 96])
def test_fronts_equal_dense_ldl(kind, n, cut, cplx, conj, nb):
    if kind == "herm":
        ip, ind, v, dims = hermitian_matrix(n)
    else:
        ip, ind, v, dims = problem(kind, n)
    t = oracle.Tree(*dims, ip, ind, cutoff=cut)
    f = oracle.Factor(t, is_complex=cplx, conjugate=conj)
    f.initialize(ip, ind, v)
    f.ldl(nb=nb)
    A = P.to_scipy(ip, ind, v).toarray()
    mp = t.map()
    Ap = np.zeros_like(A)
    Ap[np.ix_(mp, mp)] = A
    Ld, dd = dense_ldl(Ap, conj=conj)
    for s in range(t.num):
        nd = t.node(s)
        rows = np.concatenate([np.arange(nd["off"], nd["off"] + nd["size"]),
                               t.array(s, oracle.OR_LOWER)]).astype(int)
        cols = np.arange(nd["off"], nd["off"] + nd["size"])
        Lf, df = f.front(s)
        assert np.abs(Lf - Ld[np.ix_(rows, cols)]).max(initial=0) <= 1e-12
        assert np.abs(df - dd[cols]).max(initial=0) <= 1e-12 * np.abs(dd).max()


def test_chain_pivots_closed_form():
    """1D chain (n x 1 x 1) ordered naturally in each leaf: d_1 = a, d_k = a - 1/d_{k-1}."""
    a, n = 2.5, 40
    ip, ind, v, dims = P.stencil7(n, 1, 1, a, -1.0)
    t = oracle.Tree(*dims, ip, ind, cutoff=n)        # a single leaf: natural order
    f = oracle.Factor(t)
    f.initialize(ip, ind, v)
    f.ldl(nb=7)
    L, d = f.front(0)
    ref = np.zeros(n)
    ref[0] = a
    for k in range(1, n):
        ref[k] = a - 1.0 / ref[k - 1]
    assert np.allclose(d, ref, rtol=1e-14, atol=0)
    assert np.allclose(np.diag(L, -1), -1.0 / ref[:-1], rtol=1e-14, atol=0)


def test_laplacian_eigenvector_solve():
    """A^{-1} v = v / lambda for the Dirichlet eigenvectors of the 7-point operator."""
    n, shift = 12, 6.05
    ip, ind, v, dims = P.laplacian7(n, shift)
    t = oracle.Tree(*dims, ip, ind, cutoff=32)
    f = oracle.Factor(t)
    f.initialize(ip, ind, v)
    f.ldl(nb=96)
    i = np.arange(n ** 3)
    x, y, z = i % n, (i // n) % n, i // (n * n)
    B, lam = [], []
    for p, q, r in ((1, 1, 1), (2, 3, 1), (5, 4, 7), (n, n, n)):
        vec = (np.sin(p * np.pi * (x + 1) / (n + 1)) * np.sin(q * np.pi * (y + 1) / (n + 1))
               * np.sin(r * np.pi * (z + 1) / (n + 1)))
        B.append(vec)
        lam.append(shift - 2 * (np.cos(p * np.pi / (n + 1)) + np.cos(q * np.pi / (n + 1))
                                + np.cos(r * np.pi / (n + 1))))
    B = np.array(B).T
    X = f.solve(B)
    assert np.abs(X - B / np.array(lam)).max() <= 1e-12 * np.abs(B / np.array(lam)).max()


@pytest.mark.parametrize("kind,n,cut", [("lap", 10, 16), ("helm", 8, 16), ("pml", 10, 32),
                                        ("lap2d", (20, 13), 16), ("grf", 8, 16)])
@pytest.mark.parametrize("block", [False, True])
def test_solve_matches_scipy(kind, n, cut, block):
    ip, ind, v, dims = problem(kind, n)
    cplx = np.iscomplexobj(v)
    t = oracle.Tree(*dims, ip, ind, cutoff=cut)
    f = oracle.Factor(t, is_complex=cplx)
    f.initialize(ip, ind, v)
    f.ldl(nb=16, block=block)
    N = ip.size - 1
    b = P.random_rhs(N, 3, seed=1, complex_=cplx)
    x = f.solve(b)
    xr = spla.spsolve(P.to_scipy(ip, ind, v).tocsc(), b)
    assert rel_diff(x, xr) <= 1e-10
    assert backward_error(ip, ind, v, x, b).max() <= 1e-14


def test_block_ldl_front_is_inverse():
    """FrontBlockLDL leaves ATL = (L11 D L11^T)^{-1} and restores ABL."""
    ip, ind, v, dims = problem("lap", 7)
    t = oracle.Tree(*dims, ip, ind, cutoff=16)
    f1 = oracle.Factor(t)
    f1.initialize(ip, ind, v)
    f1.ldl(nb=5)
    f2 = oracle.Factor(t)
    f2.initialize(ip, ind, v)
    f2.ldl(nb=5, block=True)
    f0 = oracle.Factor(t)
    f0.initialize(ip, ind, v)            # assembled only: ABL of a leaf is A21
    for s in range(t.num):
        n = t.node(s)["size"]
        L1, d1 = f1.front(s)
        L2, _ = f2.front(s)
        T = np.tril(L1[:n, :n])
        A11 = T @ np.diag(d1) @ T.T
        assert np.allclose(L2[:n, :n] @ A11, np.eye(n), atol=1e-12)
        if t.node(s)["child0"] < 0:
            L0, _ = f0.front(s)
            assert np.array_equal(L2[n:], L0[n:])


def test_multiply_matches_scipy():
    ip, ind, v, dims = problem("pml", 9)
    N = ip.size - 1
    x = P.random_rhs(N, 4, seed=3, complex_=True)
    y0 = P.random_rhs(N, 4, seed=4, complex_=True)
    y = oracle.multiply(ip, ind, v, x, alpha=0.5, beta=-2.0, y=y0)
    ref = 0.5 * (P.to_scipy(ip, ind, v) @ x) - 2.0 * y0
    assert rel_diff(y, ref) <= 1e-15


def test_planted_sign_error_is_caught():
    """A plausible mistake planted in the oracle's output (a flipped L entry)
    must fail the dense brute-force check used above."""
    ip, ind, v, dims = problem("box", (5, 4, 3))
    t = oracle.Tree(*dims, ip, ind, cutoff=6)
    f = oracle.Factor(t)
    f.initialize(ip, ind, v)
    f.ldl(nb=4)
    A = P.to_scipy(ip, ind, v).toarray()
    mp = t.map()
    Ap = np.zeros_like(A)
    Ap[np.ix_(mp, mp)] = A
    Ld, _ = dense_ldl(Ap)
    s = t.num - 2
    nd = t.node(s)
    L, _ = f.front(s)
    L = L.copy()
    L[nd["size"], 0] = -L[nd["size"], 0] if L[nd["size"], 0] != 0 else 1.0
    rows = np.concatenate([np.arange(nd["off"], nd["off"] + nd["size"]),
                           t.array(s, oracle.OR_LOWER)]).astype(int)
    cols = np.arange(nd["off"], nd["off"] + nd["size"])
    assert np.abs(L - Ld[np.ix_(rows, cols)]).max() > 1e-3


@pytest.mark.parametrize("kind,n,cut,cplx", [("lap", 30, 128, False), ("pml", 16, 64, True)])
def test_dropped_panel_update_exceeds_fp32_front_bound(kind, n, cut, cplx):
    """The GPU FP32 factor is checked front by front against the oracle
    within F32_FRONT_BOUND (tests/test_gpu_configs.py). A factor that leaves
    out one panel's trailing update (planted in the oracle: the root's first
    panel, ldl/local_front.hpp:65-66) must miss that bound by far, so the
    bound would catch such a mistake."""
    from test_gpu_configs import F32_FRONT_BOUND
    ip, ind, v, dims = problem(kind, n)
    t = oracle.Tree(*dims, ip, ind, cutoff=cut)
    good = oracle.Factor(t, is_complex=cplx)
    good.initialize(ip, ind, v)
    good.ldl(nb=96)
    bad = oracle.Factor(t, is_complex=cplx)
    bad.initialize(ip, ind, v)
    bad.plant(t.num - 1, 0)
    try:
        bad.ldl(nb=96)
        s = t.num - 1
        Lg, dg = good.front(s)
        Lb, db = bad.front(s)
        worst = max(np.abs(Lb - Lg).max() / np.abs(Lg).max(), np.abs(db - dg).max() / np.abs(dg).max())
    except ZeroDivisionError:
        worst = np.inf   # the mistake even produced a zero pivot
    assert worst > 1e3 * F32_FRONT_BOUND, worst


@pytest.mark.parametrize("kind,n,cut,cplx,conj", [("lap", 12, 32, False, False),
                                                  ("pml", 12, 32, True, False),
                                                  ("herm", 6, 16, True, True)])
def test_blas_baseline_matches_oracle(kind, n, cut, cplx, conj):
    """The CPU-baseline path (the same LocalLDL and solves with the dense
    calls made through a CPU BLAS) reproduces the pinned oracle's fronts and
    solutions to rounding."""
    from helpers import hermitian_matrix
    oracle.blas_load(4)
    ip, ind, v, dims = hermitian_matrix(n) if kind == "herm" else problem(kind, n)
    t = oracle.Tree(*dims, ip, ind, cutoff=cut)
    ref = oracle.Factor(t, is_complex=cplx, conjugate=conj)
    ref.initialize(ip, ind, v)
    ref.ldl(nb=8)
    f = oracle.Factor(t, is_complex=cplx, conjugate=conj)
    f.initialize(ip, ind, v)
    f.ldl_blas(nb=8)
    for s in range(t.num):
        La, da = ref.front(s)
        Lb, db = f.front(s)
        assert np.abs(La - Lb).max(initial=0) <= 1e-12 * max(np.abs(La).max(initial=0), 1)
        assert np.abs(da - db).max(initial=0) <= 1e-12 * max(np.abs(da).max(initial=0), 1)
    N = ip.size - 1
    rng = np.random.default_rng(9)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    xa = ref.solve_nodal(b)
    xb = f.solve_nodal_blas(b)
    assert rel_diff(xb, xa) <= 1e-12


def _dense_nodal(ip, ind, v, t):
    A = P.to_scipy(ip, ind, v).toarray()
    mp = t.map()
    Ap = np.zeros_like(A)
    Ap[np.ix_(mp, mp)] = A
    return Ap


def _ldl_parts(L, d, sub, conj):
    n = L.shape[1]
    Lu = np.tril(L[:n, :n], -1) + np.eye(n)
    D = np.diag(d[:n]).astype(L.dtype)
    for k in range(n - 1):
        if sub[k] != 0:
            D[k + 1, k] = sub[k]
            D[k, k + 1] = np.conj(sub[k]) if conj else sub[k]
    return Lu, D


def test_bunch_kaufman_matches_lapack_sytrf_real():
    """One front holding a whole real symmetric indefinite matrix: the
    oracle's Bunch-Kaufman (the "A" rule) gives LAPACK dsytrf's pivots, D and
    L (scipy.linalg.ldl), an independent implementation of the same rule."""
    import scipy.linalg as sla
    ip, ind, v, dims = indefinite_box((6, 5, 4), seed=3)
    t = oracle.Tree(*dims, ip, ind, cutoff=1000)
    assert t.num == 1
    f = oracle.Factor(t)
    f.initialize(ip, ind, v)
    f.ldl(pivoted=True)
    L, d = f.front(0)
    sub, piv = f.pivots(0)
    Ap = _dense_nodal(ip, ind, v, t)
    lu, dd, perm = sla.ldl(Ap, lower=True, hermitian=False)
    assert np.array_equal(piv, perm)
    assert (sub != 0).sum() > 3           # 2x2 pivots were taken
    assert not np.array_equal(piv, np.arange(piv.size))
    Lu, D = _ldl_parts(L, d, sub, False)
    assert np.abs(D - dd).max() <= 1e-12 * np.abs(dd).max()
    assert np.abs(Lu - lu[perm]).max() <= 1e-12 * np.abs(lu).max()


@pytest.mark.parametrize("cplx,herm", [(False, False), (True, False), (True, True)])
def test_bunch_kaufman_reconstructs_the_front(cplx, herm):
    """P ATL P^T = L D op(L)^T for real, complex-symmetric and Hermitian
    indefinite matrices held in one front."""
    ip, ind, v, dims = indefinite_box((5, 5, 5), seed=4, cplx=cplx, herm=herm)
    t = oracle.Tree(*dims, ip, ind, cutoff=1000)
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(pivoted=True)
    L, d = f.front(0)
    sub, piv = f.pivots(0)
    Ap = _dense_nodal(ip, ind, v, t)
    Lu, D = _ldl_parts(L, d, sub, herm)
    R = Lu @ D @ (Lu.conj().T if herm else Lu.T)
    assert np.abs(R - Ap[np.ix_(piv, piv)]).max() <= 1e-12 * np.abs(Ap).max()
    assert (sub != 0).sum() > 0


@pytest.mark.parametrize("cplx,herm", [(False, False), (True, False), (True, True)])
def test_pivoted_solve_matches_scipy(cplx, herm):
    """LDL_INTRAPIV over a whole tree (pivoting inside every front's ATL,
    PermuteCols / Trsm / QuasiDiagonalSolve / Trrk for ABL and ABR) and the
    pivoted solves (PermuteRows, QuasiDiagonalSolve, InversePermuteRows)
    against SuperLU."""
    import scipy.sparse.linalg as sla
    ip, ind, v, dims = indefinite_box((9, 8, 7), seed=5, cplx=cplx, herm=herm)
    t = oracle.Tree(*dims, ip, ind, cutoff=24)
    assert t.num > 7
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(pivoted=True)
    N = ip.size - 1
    rng = np.random.default_rng(6)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    x = f.solve(b)
    xr = sla.spsolve(P.to_scipy(ip, ind, v).tocsc(), b)
    assert rel_diff(x, xr) <= 1e-10
    assert backward_error(ip, ind, v, x, b).max() <= 1e-13
    assert sum((f.pivots(s)[0] != 0).sum() for s in range(t.num)) > 0


def test_bunch_kaufman_takes_identity_pivots_when_diagonally_dominant():
    """On the strictly diagonally dominant 30^3 Laplacian (config 1) every
    pivot is 1x1 without interchange, so the pivoted factor is the
    unpivoted one."""
    ip, ind, v, dims = problem("lap", 12)
    t = oracle.Tree(*dims, ip, ind, cutoff=32)
    a = oracle.Factor(t)
    a.initialize(ip, ind, v)
    a.ldl(pivoted=True)
    b = oracle.Factor(t)
    b.initialize(ip, ind, v)
    b.ldl(nb=96)
    for s in range(t.num):
        sub, piv = a.pivots(s)
        assert np.array_equal(piv, np.arange(piv.size)) and not sub.any()
        La, da = a.front(s)
        Lb, db = b.front(s)
        assert np.abs(La - Lb).max(initial=0) <= 1e-13 * max(np.abs(Lb).max(initial=0), 1)
        assert np.abs(da - db).max(initial=0) <= 1e-13 * np.abs(db).max(initial=1)


@pytest.mark.parametrize("cplx,herm", [(False, False), (True, False), (True, True)])
def test_block_ldl_with_pivoting_inverts_the_front(cplx, herm):
    """FrontBlockLDL with intra-front pivoting (ldl/local_front_block.hpp:
    35-57) on one front holding the whole indefinite matrix: ATL becomes the
    inverse of the front's ATL in its own ordering (the permutation undone)."""
    ip, ind, v, dims = indefinite_random((5, 4, 4), seed=21, cplx=cplx, herm=herm)
    t = oracle.Tree(*dims, ip, ind, cutoff=1000)
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(block=True, pivoted=True)
    L, _ = f.front(0)
    Ap = _dense_nodal(ip, ind, v, t)
    n = Ap.shape[0]
    assert np.abs(L[:n, :n] @ Ap - np.eye(n)).max() <= 1e-10
    assert (f.pivots(0)[1] != np.arange(n)).any()


@pytest.mark.parametrize("cplx,herm", [(False, False), (True, True)])
def test_block_ldl_with_pivoting_solve_matches_scipy(cplx, herm):
    """BLOCK_LDL_INTRAPIV over a tree: the block solves (GEMMs with the
    inverted ATL and the original ABL) against SuperLU. Explicit inverses of
    indefinite fronts cost accuracy: backward error measured up to 2.9e-12,
    bound 1e-10 (the unpivoted block fronts' bound on the GPU)."""
    import scipy.sparse.linalg as sla
    ip, ind, v, dims = indefinite_random((9, 8, 7), seed=22, cplx=cplx, herm=herm)
    t = oracle.Tree(*dims, ip, ind, cutoff=24)
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(block=True, pivoted=True)
    N = ip.size - 1
    rng = np.random.default_rng(23)
    b = rng.standard_normal((N, 2)) + (1j * rng.standard_normal((N, 2)) if cplx else 0)
    x = f.solve(b)
    xr = sla.spsolve(P.to_scipy(ip, ind, v).tocsc(), b)
    assert rel_diff(x, xr) <= 1e-9
    assert backward_error(ip, ind, v, x, b).max() <= 1e-10
