"""GPU parity of intra-front Bunch-Kaufman pivoting (LDL_INTRAPIV) against
the oracle.

The reference's FrontLDLIntraPiv (ldl/local_front.hpp:73-94) factors each
front's ATL with El::LDL(..., BUNCH_KAUFMAN_A), permutes ABL's columns,
solves and scales them with the quasi-diagonal D and updates ABR; the solves
permute each front's rows before the forward sweep and after the backward
sweep (lower_solve/local_front.hpp:53-62, 84-93) and divide by the
quasi-diagonal D (diagonal_solve.hpp:58-60). The oracle restates that
algorithm unblocked (oracle/oracle_num.inc, pinned against LAPACK's sytrf);
the GPU factors ATL in panels with delayed updates. On matrices without
exact magnitude ties (tests/helpers.py indefinite_random) both make the same
pivot choices, so pivots are compared exactly and values to rounding.

Bounds: pivots and the 2x2 structure equal; L, D and D's subdiagonal within
3e-10 of the largest entry of the front (measured up to 6.9e-11 on MI355X:
Bunch-Kaufman's element growth on these indefinite matrices amplifies the
rounding differences of the two summation orders); solutions within 1e-10 of the
oracle's; FP64 backward error <= 1e-13; FP32 factor + FP64 refinement
reaches 1e-10.
"""
import ctypes

import numpy as np
import pytest

import clique_amd as cq
import oracle
from helpers import backward_error, indefinite_random, problem, rel_diff

torch = pytest.importorskip("torch")


def to_dev(a):
    return cq.colmajor(torch.tensor(np.asarray(a), device="cuda:0"))


def gpu_piv(ip, ind, v, dims, cutoff, cplx=False, herm=False, precision="f64"):
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cutoff)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx, conjugate=herm)
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision, front_type="ldl_intrapiv")
    return sym, ft


def oracle_piv(ip, ind, v, dims, cutoff, cplx=False, herm=False):
    t = oracle.Tree(*dims, ip, ind, cutoff=cutoff)
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(pivoted=True)
    return t, f


# (grid, cutoff, complex, Hermitian): a single front of order 168 and 720
# (many panels of 32 columns, 2x2 pivots across panel ends), and trees with
# fronts of many shapes.
PIV_CASES = [
    ((7, 6, 4), 1000, False, False),
    ((7, 6, 4), 1000, True, False),
    ((7, 6, 4), 1000, True, True),
    ((10, 9, 8), 1000, False, False),
    ((9, 8, 7), 24, False, False),
    ((9, 8, 7), 24, True, False),
    ((9, 8, 7), 24, True, True),
    ((14, 13, 12), 64, True, False),
]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,cut,cplx,herm", PIV_CASES)
def test_pivoted_fronts_match_oracle(device, dims, cut, cplx, herm):
    ip, ind, v, dims = indefinite_random(dims, seed=11, cplx=cplx, herm=herm)
    sym, ft = gpu_piv(ip, ind, v, dims, cut, cplx, herm)
    t, f = oracle_piv(ip, ind, v, dims, cut, cplx, herm)
    assert sym.num_nodes == t.num
    twos = swaps = 0
    worst = 0.0
    for s in range(t.num):
        sub_r, piv_r = f.pivots(s)
        sub_g, piv_g = ft.pivots(s)
        assert np.array_equal(piv_g, piv_r), f"front {s}: pivots differ"
        assert np.array_equal(sub_g != 0, sub_r != 0), f"front {s}: 2x2 pivots differ"
        twos += int((sub_r != 0).sum())
        swaps += int((piv_r != np.arange(piv_r.size)).sum())
        Lr, dr = f.front(s)
        Lg, dg = ft.front(s)
        scale = max(np.abs(Lr).max(initial=0), np.abs(dr).max(initial=0), 1.0)
        worst = max(worst, np.abs(Lg - Lr).max(initial=0) / scale,
                    np.abs(dg - dr).max(initial=0) / scale,
                    np.abs(sub_g - sub_r).max(initial=0) / scale)
    assert twos > 0 and swaps > 0
    print(f"fronts {t.num}, 2x2 pivots {twos}, interchanges {swaps}, worst {worst:.2e}")
    assert worst <= 3e-10, worst


@pytest.mark.gpu
@pytest.mark.parametrize("dims,cut,cplx,herm", [c for c in PIV_CASES if c[1] < 1000])
@pytest.mark.parametrize("k", [1, 6])
def test_pivoted_solve_matches_oracle(device, dims, cut, cplx, herm, k):
    ip, ind, v, dims = indefinite_random(dims, seed=12, cplx=cplx, herm=herm)
    sym, ft = gpu_piv(ip, ind, v, dims, cut, cplx, herm)
    t, f = oracle_piv(ip, ind, v, dims, cut, cplx, herm)
    N = ip.size - 1
    rng = np.random.default_rng(13)
    b = rng.standard_normal((N, k)) + (1j * rng.standard_normal((N, k)) if cplx else 0)
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    xr = f.solve(b)
    assert rel_diff(xg, xr) <= 1e-10
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-13


@pytest.mark.gpu
def test_pivoted_equals_unpivoted_when_diagonally_dominant(device):
    """On the strictly diagonally dominant Laplacian every Bunch-Kaufman pivot
    is 1x1 without interchange: the pivoted GPU factor is the unpivoted one."""
    ip, ind, v, dims = problem("lap", 20)
    sym, a = gpu_piv(ip, ind, v, dims, 64)
    b = cq.DistSymmFrontTree(sym, ip, ind)
    b.factor(torch.tensor(v, device="cuda:0"), precision="f64")
    for s in range(sym.num_nodes):
        sub, piv = a.pivots(s)
        assert np.array_equal(piv, np.arange(piv.size)) and not sub.any()
        La, da = a.front(s)
        Lb, db = b.front(s)
        assert np.abs(La - Lb).max(initial=0) <= 1e-12 * max(np.abs(Lb).max(initial=0), 1)
        assert np.abs(da - db).max(initial=0) <= 1e-12 * np.abs(db).max(initial=1)


@pytest.mark.gpu
@pytest.mark.parametrize("cplx,herm", [(False, False), (True, False), (True, True)])
def test_pivoted_fp32_factor_with_fp64_refinement(device, cplx, herm):
    """FP32 pivoted factor + FP64 refinement on an indefinite problem the
    unpivoted LDL^T handles poorly (a weak, sign-changing diagonal)."""
    ip, ind, v, dims = indefinite_random((16, 15, 14), seed=14, cplx=cplx, herm=herm)
    sym, ft = gpu_piv(ip, ind, v, dims, 64, cplx, herm, precision="f32")
    N = ip.size - 1
    rng = np.random.default_rng(15)
    b = rng.standard_normal((N, 4)) + (1j * rng.standard_normal((N, 4)) if cplx else 0)
    x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=12)
    assert info["converged"]
    eta = backward_error(ip, ind, v, x.cpu().numpy(), b)
    assert eta.max() <= 1e-10


@pytest.mark.gpu
def test_pivoted_solve_paths_agree(device):
    """The allocation-free solve, graph replay and nodal solve run the
    pivoted schedule (row permutations, quasi-diagonal solve) and agree
    with the plain solve bit for bit."""
    ip, ind, v, dims = indefinite_random((11, 10, 9), seed=16, cplx=True)
    sym, ft = gpu_piv(ip, ind, v, dims, 32, True, False)
    N = ip.size - 1
    rng = np.random.default_rng(17)
    b = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    x1 = to_dev(b)
    ft.solve(x1)
    x2 = to_dev(b)
    t, ld, k = cq._check_rhs(x2, ft.N, ft.is_complex)
    assert cq.lib().cliqg_plan_solve_noalloc(ft._h, ctypes.c_void_p(t.data_ptr()), ld, k) == 0
    assert torch.equal(x1, x2)
    ft.set_graphs(True)
    for _ in range(3):
        x3 = to_dev(b)
        ft.solve(x3)
        torch.cuda.synchronize()
        assert torch.equal(x1, x3)
    ft.set_graphs(False)
    mp = sym.map()
    bn = np.empty_like(b)
    bn[mp] = b
    xn = to_dev(bn)
    ft.solve(xn, nodal=True)
    assert np.array_equal(xn.cpu().numpy()[mp], x1.cpu().numpy())


@pytest.mark.gpu
def test_pivoted_zero_pivot_is_reported(device):
    """A front whose ATL is singular (a zero column) raises SINGULAR."""
    ip, ind, v, dims = problem("box", (5, 1, 1))
    v = np.zeros_like(v)
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=8)
    ft = cq.DistSymmFrontTree(sym, ip, ind)
    with pytest.raises(cq.CliqgError) as e:
        ft.factor(torch.tensor(v, device="cuda:0"), front_type="ldl_intrapiv")
    assert e.value.code == 2


@pytest.mark.gpu
@pytest.mark.parametrize("dims,cut,cplx,herm",
                         [((7, 6, 4), 1000, True, False), ((9, 8, 7), 24, False, False),
                          ((9, 8, 7), 24, True, True), ((14, 13, 12), 64, True, False)])
def test_block_pivoted_fronts_and_solve_match_oracle(device, dims, cut, cplx, herm):
    """BLOCK_LDL_INTRAPIV (FrontBlockLDL with intraPiv,
    ldl/local_front_block.hpp:35-57): each front's ATL becomes the inverse of
    the front's ATL in its own (unpermuted) ordering, ABL the assembled A21.
    Fronts against the oracle's mode 3 within 3e-10 of the largest entry (the
    inverse is formed from the pivoted factor, whose rounding Bunch-Kaufman's
    growth amplifies, as for LDL_INTRAPIV); solves within 1e-10 of the
    oracle's; FP64 backward error <= 1e-10 (a solve through an explicit
    inverse of an indefinite ATL is not backward stable at rounding level:
    the oracle's own block-pivoted solves reach 2.9e-12, the GPU's 1.9e-12 on
    the Hermitian case); FP32 factor + FP64 refinement to 1e-10."""
    ip, ind, v, dims = indefinite_random(dims, seed=21, cplx=cplx, herm=herm)
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cut)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx, conjugate=herm)
    ft.factor(torch.tensor(v, device="cuda:0"), precision="f64", front_type="block_ldl_intrapiv")
    t = oracle.Tree(*dims, ip, ind, cutoff=cut)
    f = oracle.Factor(t, is_complex=cplx, conjugate=herm)
    f.initialize(ip, ind, v)
    f.ldl(block=True, pivoted=True)
    worst = 0.0
    for s in range(t.num):
        Lr, _ = f.front(s)
        Lg, _ = ft.front(s)
        scale = max(np.abs(Lr).max(initial=0), 1.0)
        worst = max(worst, np.abs(Lg - Lr).max(initial=0) / scale)
    print(f"fronts {t.num}, worst {worst:.2e}")
    assert worst <= 3e-10, worst
    N = ip.size - 1
    rng = np.random.default_rng(22)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    x = to_dev(b)
    ft.solve(x)
    xg = x.cpu().numpy()
    assert rel_diff(xg, f.solve(b)) <= 1e-10
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-10
    ft32 = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx, conjugate=herm)
    ft32.factor(torch.tensor(v, device="cuda:0"), precision="f32",
                front_type="block_ldl_intrapiv")
    x, info = ft32.refine(to_dev(b), tol=1e-10, max_iters=12)
    assert info["converged"]
    assert backward_error(ip, ind, v, x.cpu().numpy(), b).max() <= 1e-10
