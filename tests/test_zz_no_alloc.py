"""The allocation-free solve (cliqg_plan_solve_noalloc) against the
allocating one (cliqg_plan_solve), through the C ABI.

Both run the same kernels on the same launch schedule, so the results must be
bit-identical. Calls in a row share the library's scratch array: the second
call of each pair has a smaller input than the first (fewer right-hand sides,
then a smaller plan), so stale contents of the array would show.
"""
import ctypes

import numpy as np
import pytest

import clique_amd as cq
from clique_amd import _lib
from helpers import problem

torch = pytest.importorskip("torch")


def _noalloc(ft, x):
    t, ld, k = cq._check_rhs(x, ft.N, ft.is_complex)
    return cq.lib().cliqg_plan_solve_noalloc(ft._h, ctypes.c_void_p(t.data_ptr()), ld, k)


def _plan(kind, n, cut, cplx, precision):
    ip, ind, v, dims = problem(kind, n)
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cut)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision)
    return ft


def _rhs(N, k, cplx, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((N, k)) + (1j * rng.standard_normal((N, k)) if cplx else 0)
    return b


@pytest.mark.gpu
def test_noalloc_solve_matches_solve(device):
    big = _plan("pml", 14, 32, True, "f32")
    small = _plan("lap", 9, 16, False, "f64")
    for ft, k, seed in ((big, 6, 1), (big, 3, 2), (small, 2, 3), (big, 6, 4)):
        b = _rhs(ft.N, k, ft.is_complex, seed)
        ref = cq.colmajor(torch.tensor(b, device=device))
        ft.solve(ref)
        x = cq.colmajor(torch.tensor(b, device=device))
        st = _noalloc(ft, x)
        assert st == 0, cq.lib().cliqg_last_error()
        assert np.array_equal(x.cpu().numpy(), ref.cpu().numpy()), (ft.N, k)


@pytest.mark.gpu
def test_noalloc_solve_rejects_a_workspace_beyond_its_scratch(device):
    ft = _plan("lap", 40, 128, False, "f64")
    k = 300   # two nodal buffers of 64,000 x 300 doubles alone exceed 128 MiB
    x = cq.colmajor(torch.zeros((ft.N, k), dtype=torch.float64, device=device))
    st = _noalloc(ft, x)
    assert st == 1
    assert b"scratch" in cq.lib().cliqg_last_error()
    # the plan is unharmed: an allocating solve still works
    b = _rhs(ft.N, 2, False, 5)
    y = cq.colmajor(torch.tensor(b, device=device))
    ft.solve(y)
    assert np.isfinite(y.cpu().numpy()).all()
