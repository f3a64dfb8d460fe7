"""GPU parity of the factorization and solves against the CPU oracle.

The oracle restates Clique's FP64 LocalLDL/FrontLDL (nb = 96, the blocksize
of tests/Solve.cpp:39) and its solves; the GPU path uses 64-wide diagonal
blocks and one delayed Schur update per front, so FP64 results agree to
rounding, not bitwise. Tolerances are stated per test.
"""
import numpy as np
import pytest

import clique_amd as cq
import oracle
from helpers import backward_error, hermitian_matrix, problem, rel_diff

torch = pytest.importorskip("torch")


def gpu_tree(ip, ind, v, dims, cutoff, is_complex=False, conjugate=False, precision="f64",
             front_type="ldl"):
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cutoff)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=is_complex, conjugate=conjugate)
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision, front_type=front_type)
    return sym, ft


def oracle_tree(ip, ind, v, dims, cutoff, is_complex=False, conjugate=False, block=False):
    t = oracle.Tree(*dims, ip, ind, cutoff=cutoff)
    f = oracle.Factor(t, is_complex=is_complex, conjugate=conjugate)
    f.initialize(ip, ind, v)
    f.ldl(nb=96, block=block)
    return t, f


def to_dev(a):
    return cq.colmajor(torch.tensor(np.asarray(a), device="cuda:0"))


# FP32 factor + FP64 refinement: every workload of this file converges to
# 1e-10 in at most one step on MI355X; the cap allows two more.
REFINE_CAP = 3


CASES = [
    ("box", (2, 1, 1), 1),
    ("box", (7, 5, 3), 4),
    ("box", (17, 1, 1), 2),
    ("box", (12, 9, 1), 10),
    ("lap", 6, 8),
    ("lap", 11, 16),
    ("lap", 30, 128),      # config 1 (N = 27,000, 363 fronts)
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut", CASES)
def test_fronts_match_oracle_fp64(device, kind, n, cut):
    ip, ind, v, dims = problem(kind, n)
    sym, ft = gpu_tree(ip, ind, v, dims, cut)
    t, f = oracle_tree(ip, ind, v, dims, cut)
    worst = 0.0
    for s in range(sym.num_nodes):
        Lg, dg = ft.front(s)
        Lo, do = f.front(s)
        scale = max(np.abs(Lo).max(initial=0), 1.0)
        worst = max(worst, np.abs(Lg - Lo).max(initial=0) / scale)
        if do.size:
            assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    assert worst <= 1e-12, worst


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut", CASES)
@pytest.mark.parametrize("k", [1, 5])
def test_solve_matches_oracle_fp64(device, kind, n, cut, k):
    ip, ind, v, dims = problem(kind, n)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut)
    t, f = oracle_tree(ip, ind, v, dims, cut)
    b = np.random.default_rng(1).standard_normal((N, k))
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    xo = f.solve(b)
    assert rel_diff(xg, xo) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("n,cut", [(8, 16), (16, 64)])
def test_complex_symmetric_helmholtz_fp64(device, n, cut):
    ip, ind, v, dims = problem("helm", n)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut, is_complex=True)
    t, f = oracle_tree(ip, ind, v, dims, cut, is_complex=True)
    for s in range(sym.num_nodes):
        Lg, dg = ft.front(s)
        Lo, do = f.front(s)
        assert np.abs(Lg - Lo).max(initial=0) <= 1e-11 * max(np.abs(Lo).max(initial=0), 1)
    rng = np.random.default_rng(2)
    b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("precision,front_type,tol", [("f64", "ldl", 1e-13), ("f32", "ldl", 1e-5),
                                                      ("f64", "block_ldl", 1e-10)])
def test_pml40_factor_backward_error(device, precision, front_type, tol):
    # fronts up to 2,360 x 760, whose TRSMs (C = A Linv^T in place) span many
    # row tiles: the in-place product must keep a row block's columns in one
    # tile (two 32-wide tile columns raced here, backward error 1e-5 in FP64);
    # the solves' update-row products (K = u up to 1,600) are split in K
    ip, ind, v, dims = problem("pml", 40, pml=5)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 128, is_complex=True, precision=precision,
                       front_type=front_type)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    x = to_dev(b)
    ft.solve(x)
    eta = backward_error(ip, ind, v, x.cpu().numpy(), b)
    assert eta.max() <= tol, eta


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_helmholtz64_split_k_solves(device, precision):
    # the 64^3 Helmholtz operator (N = 262,144): fronts up to 4,096 wide whose
    # update-row products (K = n or u >= 1,024, few tiles) are split in K in
    # both sweeps; FP64 solves to rounding, FP32 refines to 1e-10
    ip, ind, v, dims = problem("helm", 64)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 128, is_complex=True, precision=precision)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
    if precision == "f64":
        x = to_dev(b)
        ft.solve(x)
        eta = backward_error(ip, ind, v, x.cpu().numpy(), b)
        assert eta.max() <= 1e-13, eta
    else:
        x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=REFINE_CAP)
        eta = backward_error(ip, ind, v, x.cpu().numpy(), b)
        assert info["converged"] and eta.max() <= 1e-10, (info, eta)


@pytest.mark.gpu
def test_hermitian_ldlh_fp64(device):
    ip, ind, v, dims = hermitian_matrix(6)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 16, is_complex=True, conjugate=True)
    t, f = oracle_tree(ip, ind, v, dims, 16, is_complex=True, conjugate=True)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut,cplx", [("lap", 30, 128, False), ("helm", 16, 64, True),
                                             ("pml", 16, 64, True)])
def test_fp32_factor_with_fp64_refinement(device, kind, n, cut, cplx):
    ip, ind, v, dims = problem(kind, n)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut, is_complex=cplx, precision="f32")
    rng = np.random.default_rng(4)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    bd = to_dev(b)
    x, info = ft.refine(bd, tol=1e-10, max_iters=REFINE_CAP)
    xg = x.cpu().numpy()
    eta = backward_error(ip, ind, v, xg, b)
    assert info["converged"] and eta.max() <= 1e-10, (info, eta)
    assert abs(info["backward_error"] - eta.max()) <= 1e-2 * eta.max() + 1e-16


@pytest.mark.gpu
def test_refinement_skips_converged_columns(device):
    """Columns that reach the tolerance drop out of the correction solves:
    with 24 zero right-hand sides (converged at the first residual) among 40,
    every correction solve runs on the 16 others, gathered into a narrower
    solve. Each column must still reach the tolerance, the zero columns stay
    zero, and the others agree with a refinement of those 16 alone."""
    ip, ind, v, dims = problem("pml", 16)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 64, is_complex=True, precision="f32")
    rng = np.random.default_rng(11)
    live = np.sort(rng.choice(40, 16, replace=False))
    b = np.zeros((N, 40), dtype=np.complex128)
    b[:, live] = rng.standard_normal((N, 16)) + 1j * rng.standard_normal((N, 16))
    x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=REFINE_CAP)
    xg = x.cpu().numpy()
    eta = backward_error(ip, ind, v, xg[:, live], b[:, live])
    assert info["converged"] and eta.max() <= 1e-10, (info, eta)
    assert info["iterations"] >= 1
    dead = np.setdiff1d(np.arange(40), live)
    assert np.all(xg[:, dead] == 0)
    x16, info16 = ft.refine(to_dev(b[:, live]), tol=1e-10, max_iters=REFINE_CAP)
    assert info16["iterations"] == info["iterations"]
    assert rel_diff(xg[:, live], x16.cpu().numpy()) <= 1e-8


@pytest.mark.gpu
def test_multiply_matches_oracle(device):
    ip, ind, v, dims = problem("helm", 10)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 32, is_complex=True)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    y0 = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    yd = to_dev(y0)
    ft.multiply(to_dev(x), yd, alpha=-1.0, beta=2.0)
    yo = oracle.multiply(ip, ind, v, x, alpha=-1.0, beta=2.0, y=y0)
    assert rel_diff(yd.cpu().numpy(), yo) <= 1e-14


@pytest.mark.gpu
def test_forest_of_independent_problems(device):
    """Concatenated symbolics factor as one forest; each block matches its own solve."""
    import scipy.sparse as sp
    shapes = [(9, 7, 1), (5, 5, 5), (12, 3, 1)]
    syms, mats, rhs = [], [], []
    for i, sh in enumerate(shapes):
        ip, ind, v, dims = problem("box", sh, diag=4.3 + i)
        syms.append(cq.NaturalNestedDissection(*dims, ip, ind, cutoff=8))
        mats.append(sp.csr_matrix((v, ind, ip)))
    A = sp.block_diag(mats, format="csr")
    A.sort_indices()
    sym = cq.Symbolic.concat(syms)
    ft = cq.DistSymmFrontTree(sym, A.indptr, A.indices)
    ft.factor(torch.tensor(A.data, device="cuda:0"))
    b = np.random.default_rng(6).standard_normal((A.shape[0], 2))
    x = to_dev(b)
    ft.solve(x)
    eta = backward_error(A.indptr, A.indices, A.data, x.cpu().numpy(), b)
    assert eta.max() <= 1e-14


@pytest.mark.gpu
def test_zero_pivot_is_reported(device):
    ip, ind, v, dims = problem("box", (4, 4, 1), diag=4.0)
    v = v.copy()
    N = ip.size - 1
    # a zero row/column makes the first pivot of its front exactly zero
    rows = np.repeat(np.arange(N), np.diff(ip))
    v[(rows == 0) | (ind == 0)] = 0.0
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=4)
    ft = cq.DistSymmFrontTree(sym, ip, ind)
    with pytest.raises(cq.CliqgError) as e:
        ft.factor(torch.tensor(v, device="cuda:0"))
    assert e.value.code == 2


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batched_2d_problems_one_forest(device, precision):
    """Config-3 shape at small scale: independent 2D 5-point problems of
    uneven sizes factored as one forest; per-problem backward error."""
    import scipy.sparse as sp
    rng = np.random.default_rng(7)
    sizes = rng.integers(8, 41, size=(12, 2))
    shifts = rng.uniform(0.01, 0.5, size=12)
    syms, mats = [], []
    for (nx, ny), s in zip(sizes, shifts):
        ip, ind, v, dims = problem("lap2d", (int(nx), int(ny)), s=float(s))
        syms.append(cq.NaturalNestedDissection(*dims, ip, ind, cutoff=16))
        mats.append(sp.csr_matrix((v, ind, ip)))
    A = sp.block_diag(mats, format="csr")
    A.sort_indices()
    sym = cq.Symbolic.concat(syms)
    ft = cq.DistSymmFrontTree(sym, A.indptr, A.indices)
    ft.factor(torch.tensor(A.data, device="cuda:0"), precision=precision)
    N = A.shape[0]
    b = rng.standard_normal((N, 16))
    x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=REFINE_CAP)
    xg = x.cpu().numpy()
    o = 0
    for m in mats:
        n = m.shape[0]
        eta = backward_error(m.indptr, m.indices, m.data, xg[o:o + n], b[o:o + n])
        assert eta.max() <= 1e-10
        o += n


BLOCK_CASES = [
    ("box", (7, 5, 3), 4, False),
    ("box", (17, 1, 1), 2, False),
    ("lap", 11, 16, False),
    ("lap", 20, 128, False),    # root n = 400: several 64-wide blocks, a partial last one
    ("helm", 8, 16, True),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut,cplx", BLOCK_CASES)
def test_block_ldl_fronts_and_solve_match_oracle(device, kind, n, cut, cplx):
    """BLOCK_LDL (ldl/local_front_block.hpp:25-73): ATL holds the full
    symmetric inverse of L11 D L11^T, ABL the assembled A21; the block solves
    (lower_solve/local_front_block.hpp) are products with those blocks."""
    ip, ind, v, dims = problem(kind, n)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut, is_complex=cplx, front_type="block_ldl")
    t, f = oracle_tree(ip, ind, v, dims, cut, is_complex=cplx, block=True)
    worst = 0.0
    for s in range(sym.num_nodes):
        Lg, dg = ft.front(s)
        Lo, _ = f.front(s)
        assert dg.size == 0
        scale = max(np.abs(Lo).max(initial=0), 1.0)
        worst = max(worst, np.abs(Lg - Lo).max(initial=0) / scale)
        nn = Lg.shape[1]
        atl = Lg[:nn]
        assert np.array_equal(atl, atl.T)    # MakeSymmetric: both triangles written
    assert worst <= 1e-11, worst
    rng = np.random.default_rng(11)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-12


@pytest.mark.gpu
def test_block_ldl_hermitian(device):
    ip, ind, v, dims = hermitian_matrix(6)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, 16, is_complex=True, conjugate=True,
                       front_type="block_ldl")
    t, f = oracle_tree(ip, ind, v, dims, 16, is_complex=True, conjugate=True, block=True)
    for s in range(sym.num_nodes):
        Lg, _ = ft.front(s)
        Lo, _ = f.front(s)
        nn = Lg.shape[1]
        assert np.abs(Lg - Lo).max(initial=0) <= 1e-11 * max(np.abs(Lo).max(initial=0), 1)
        atl = Lg[:nn]
        assert np.array_equal(atl, atl.conj().T)   # Hermitian inverse, real diagonal
    rng = np.random.default_rng(12)
    b = rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut,cplx", [("lap", 30, 128, False), ("pml", 16, 64, True)])
def test_block_ldl_fp32_with_refinement(device, kind, n, cut, cplx):
    ip, ind, v, dims = problem(kind, n)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut, is_complex=cplx, precision="f32",
                       front_type="block_ldl")
    rng = np.random.default_rng(13)
    b = rng.standard_normal((N, 4)) + (1j * rng.standard_normal((N, 4)) if cplx else 0)
    x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=REFINE_CAP)
    eta = backward_error(ip, ind, v, x.cpu().numpy(), b)
    assert info["converged"] and eta.max() <= 1e-10, (info, eta)


@pytest.mark.gpu
def test_refactor_switches_between_ldl_and_block(device):
    """One plan refactored as LDL, BLOCK_LDL, then LDL again gives the same solution."""
    ip, ind, v, dims = problem("lap", 12)
    N = ip.size - 1
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=32)
    ft = cq.DistSymmFrontTree(sym, ip, ind)
    b = np.random.default_rng(14).standard_normal((N, 2))
    xs = []
    for ftype in ("ldl", "block_ldl", "ldl"):
        ft.factor(torch.tensor(v, device="cuda:0"), front_type=ftype)
        x = to_dev(b)
        ft.solve(x)
        xs.append(x.cpu().numpy())
    assert rel_diff(xs[1], xs[0]) <= 1e-12
    assert np.array_equal(xs[2], xs[0])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_graph_replays_match_direct_launches(device, precision):
    """The factor and solve launch sequences are recorded once as HIP graphs
    and replayed; every replay (with host work and synchronizations between
    them, and a refactorization with new values) must equal, bit for bit,
    the same plan type launching each kernel directly."""
    ip, ind, v, dims = problem("pml", 14)
    N = ip.size - 1
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=32)
    g = cq.DistSymmFrontTree(sym, ip, ind, is_complex=True)
    d = cq.DistSymmFrontTree(sym, ip, ind, is_complex=True)
    g.set_graphs(True)
    d.set_graphs(False)
    rng = np.random.default_rng(21)
    for scale in (1.0, 1.5):
        vals = torch.tensor(v * scale, device="cuda:0")
        g.factor(vals, precision=precision)
        d.factor(vals, precision=precision)
        for s in (0, sym.num_nodes // 2, sym.num_nodes - 1):
            assert np.array_equal(g.front(s)[0], d.front(s)[0])
        for rep in range(3):
            b = rng.standard_normal((N, 5)) + 1j * rng.standard_normal((N, 5))
            xg, xd = to_dev(b), to_dev(b)
            g.solve(xg)
            d.solve(xd)
            assert np.array_equal(xg.cpu().numpy(), xd.cpu().numpy()), (scale, rep)


@pytest.mark.gpu
@pytest.mark.parametrize("n,cut", [(10, 16), (16, 64)])
def test_grf_diffusion_matches_oracle(device, n, cut):
    """Config-4 family (7-point diffusion with a log-normal coefficient):
    strongly varying but SPD; FP64 fronts and solutions against the oracle,
    and FP32 factor + FP64 refinement to 1e-10."""
    ip, ind, v, dims = problem("grf", n, seed=5)
    N = ip.size - 1
    sym, ft = gpu_tree(ip, ind, v, dims, cut)
    t, f = oracle_tree(ip, ind, v, dims, cut)
    for s in range(sym.num_nodes):
        Lg, dg = ft.front(s)
        Lo, do = f.front(s)
        assert np.abs(Lg - Lo).max(initial=0) <= 1e-12 * max(np.abs(Lo).max(initial=0), 1.0)
        assert np.abs(dg - do).max(initial=0) <= 1e-12 * np.abs(do).max(initial=1.0)
    b = np.random.default_rng(31).standard_normal((N, 8))
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-11
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-14
    sym32, ft32 = gpu_tree(ip, ind, v, dims, cut, precision="f32")
    xr, info = ft32.refine(to_dev(b), tol=1e-10, max_iters=REFINE_CAP)
    assert info["converged"]
    assert backward_error(ip, ind, v, xr.cpu().numpy(), b).max() <= 1e-10


_SWITCH_SCRIPT = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import clique_amd as cq
from helpers import backward_error, problem
ip, ind, v, dims = problem("pml", 40, pml=5)
N = ip.size - 1
sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=128)
ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=True)
ft.factor(torch.tensor(v, device="cuda:0"), precision="f64")
rng = np.random.default_rng(7)
b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
x = cq.colmajor(torch.tensor(b, device="cuda:0"))
ft.solve(x)
print("%.3e" % backward_error(ip, ind, v, x.cpu().numpy(), b).max())
"""


@pytest.mark.gpu
def test_tuning_switches_keep_results(device):
    # the non-default schedule paths (no small tiles, no split K, no
    # lookahead, 128-wide outer panels, 64-wide solve blocks) solve to the
    # same FP64 rounding level; the switches are read once per process
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, CLIQG_SMALL_TILES="0", CLIQG_SPLIT_TILES="0", CLIQG_LOOKAHEAD="0",
               CLIQG_OUTER_PANEL="128", CLIQG_SOLVE_BLOCK="64")
    r = subprocess.run([sys.executable, "-c", _SWITCH_SCRIPT, root], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert float(r.stdout.strip().splitlines()[-1]) <= 1e-13, r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_staged_fronts_replace_assembled_values(device, precision):
    """Fronts written from the host before a factorization (the reference's
    tests/Refactor.cpp writes localFronts[s].frontL and factors again) are
    factored in place of A's values: staging every front of 2A (the oracle's
    assembled fronts, doubled) gives the solution of 2A x = b, x/2 (scaling
    by two is exact, so to rounding of the division only); the next
    factorization without staging is A's again, bit for bit."""
    ip, ind, v, dims = problem("helm", 10)
    N = ip.size - 1
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=16)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=True)
    vals = torch.tensor(v, device="cuda:0")
    t = oracle.Tree(*dims, ip, ind, cutoff=16)
    f = oracle.Factor(t, is_complex=True)
    f.initialize(ip, ind, v)        # assembled, not factored
    b = np.random.default_rng(41).standard_normal((N, 2)) + 0j

    def solve():
        y = to_dev(b)
        ft.solve(y)
        return y.cpu().numpy()

    ft.factor(vals, precision=precision)
    y1 = solve()
    for s in range(sym.num_nodes):
        ft.set_front(s, 2.0 * f.front(s)[0])
    ft.factor(vals, precision=precision)
    y2 = solve()
    ft.factor(vals, precision=precision)
    y3 = solve()
    assert rel_diff(2.0 * y2, y1) <= 1e-12
    assert np.array_equal(y3, y1)
