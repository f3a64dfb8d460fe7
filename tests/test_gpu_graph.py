"""Factor and solve on general graphs ordered by NestedDissection (the
reference's nested_dissection.hpp with the library's own bisection in place
of METIS), against SuperLU (scipy) on the same matrix: FP64 LDL and the
pivoted LDL to backward error 1e-13 and within 1e-10 of SuperLU's solution;
FP32 + FP64 refinement to 1e-10."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import clique_amd as cq
from helpers import backward_error, rel_diff
from test_host import irregular_graph

torch = pytest.importorskip("torch")


def graph_matrix(seed, cplx):
    """A symmetric matrix on the irregular graph: random couplings, a
    diagonal that makes it indefinite when shift < 0."""
    ip, ind = irregular_graph(seed)
    N = ip.size - 1
    rng = np.random.default_rng(seed + 100)
    A = sp.csr_matrix((np.ones(ind.size), ind, ip), shape=(N, N))
    U = sp.triu(A, 1).tocoo()
    vals = -rng.uniform(0.5, 1.5, U.nnz)
    if cplx:
        vals = vals * np.exp(1j * rng.uniform(0, 0.3, U.nnz))
    U = sp.csr_matrix((vals, (U.row, U.col)), shape=(N, N))
    rowsum = np.asarray(abs(U + U.T).sum(axis=1)).ravel()
    M = (U + U.T + sp.diags(rowsum * 1.05 - (0.7 if cplx else 0.0))).tocsr()
    M.sort_indices()
    assert np.array_equal(M.indptr, ip) and np.array_equal(M.indices, ind)
    return ip, ind, M.data, M


@pytest.mark.gpu
@pytest.mark.parametrize("seed,cplx,front_type,cut", [(0, False, "ldl", 32),
                                                      (1, True, "ldl", 64),
                                                      (2, True, "ldl_intrapiv", 32),
                                                      (3, False, "ldl_intrapiv", 128)])
def test_graph_dissection_solve_matches_superlu(device, seed, cplx, front_type, cut):
    ip, ind, v, M = graph_matrix(seed, cplx)
    N = ip.size - 1
    sym = cq.NestedDissection(ip, ind, cutoff=cut)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
    ft.factor(torch.tensor(v, device="cuda:0"), precision="f64", front_type=front_type)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
    x = cq.colmajor(torch.tensor(b, device="cuda:0"))
    ft.solve(x)
    xg = x.cpu().numpy()
    xr = spla.spsolve(M.tocsc(), b)
    assert backward_error(ip, ind, v, xg, b).max() <= 1e-13
    assert rel_diff(xg, xr) <= 1e-10
    ft.factor(torch.tensor(v, device="cuda:0"), precision="f32", front_type=front_type)
    x32, info = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-10, max_iters=5)
    assert info["converged"]
    assert backward_error(ip, ind, v, x32.cpu().numpy(), b).max() <= 1e-10
