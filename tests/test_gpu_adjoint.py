"""Adjoint of the solve (GPU): against the two-solve oracle gradient, central
finite differences, and through torch.autograd.

Conventions (stated in include/cliqg.h): stored entries of A are independent
parameters, dL/dA[e] = -lambda(i) x(j) with lambda = A^{-T} dL/dx. A
perturbation that keeps A symmetric moves the two stored entries (i,j) and
(j,i) together, so its derivative is dA[(i,j)] + dA[(j,i)].
"""
import numpy as np
import pytest

import clique_amd as cq
import oracle
from clique_amd import problems as P
from helpers import problem, rel_diff

torch = pytest.importorskip("torch")


# refinement to backward error 1e-13 from an FP32 factor takes at most two
# steps on these problems; the cap allows two more
REFINE_CAP = 4


def plan_for(ip, ind, v, dims, cut=32, cplx=False):
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cut)
    return cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)


def solve_np(ft, v, b, precision):
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision)
    x, _ = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-13, max_iters=REFINE_CAP)
    return x.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_adjoint_matches_two_solve_reference(device, precision):
    ip, ind, v, dims = problem("grf", 10, seed=3)
    N = ip.size - 1
    ft = plan_for(ip, ind, v, dims)
    rng = np.random.default_rng(0)
    b = rng.standard_normal((N, 3))
    g = rng.standard_normal((N, 3))
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision)
    x, _ = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-13, max_iters=REFINE_CAP)
    lam, dA = ft.adjoint(cq.colmajor(torch.tensor(g, device="cuda:0")), x)
    t = oracle.Tree(*dims, ip, ind, cutoff=32)
    f = oracle.Factor(t)
    f.initialize(ip, ind, v)
    f.ldl(nb=96)
    xr, lr = f.solve(b), f.solve(g)
    rows = np.repeat(np.arange(N), np.diff(ip))
    dAr = -(lr[rows] * xr[ind]).sum(axis=1)
    assert rel_diff(lam.cpu().numpy(), lr) <= 1e-11
    assert rel_diff(dA.cpu().numpy(), dAr) <= 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_adjoint_against_central_differences(device, precision):
    """Config-4-style loss 0.5 ||P x - d||^2 on a GRF diffusion operator."""
    ip, ind, v, dims = problem("grf", 9, seed=1)
    N = ip.size - 1
    k = 2
    rng = np.random.default_rng(1)
    b = rng.standard_normal((N, k))
    samp = rng.choice(N, size=max(N // 100, 4), replace=False)
    d = rng.standard_normal((samp.size, k))
    ft = plan_for(ip, ind, v, dims)

    def loss(vv, bb):
        x = solve_np(ft, vv, bb, precision)
        return 0.5 * np.sum((x[samp] - d) ** 2), x

    L0, x = loss(v, b)
    g = np.zeros((N, k))
    g[samp] = x[samp] - d
    ft.factor(torch.tensor(v, device="cuda:0"), precision=precision)
    lam, dA = ft.adjoint(cq.colmajor(torch.tensor(g, device="cuda:0")),
                         cq.colmajor(torch.tensor(x, device="cuda:0")))
    lam, dA = lam.cpu().numpy(), dA.cpu().numpy()
    rows = np.repeat(np.arange(N), np.diff(ip))
    h = 1e-4
    # b directions
    for trial in range(3):
        vdir = rng.standard_normal((N, k))
        fd = (loss(v, b + h * vdir)[0] - loss(v, b - h * vdir)[0]) / (2 * h)
        an = np.sum(lam * vdir)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (fd, an)
    # symmetric A perturbations: an off-diagonal pair and a diagonal entry
    offd = np.nonzero(rows < ind)[0]
    for e in list(rng.choice(offd, 3, replace=False)) + [int(np.nonzero(rows == ind)[0][5])]:
        i, j = rows[e], ind[e]
        pair = np.nonzero((rows == j) & (ind == i))[0][0]
        E = np.zeros_like(v)
        E[e] = 1.0
        E[pair] = 1.0
        scale = abs(v[e])
        fd = (loss(v + h * scale * E, b)[0] - loss(v - h * scale * E, b)[0]) / (2 * h * scale)
        an = dA[e] + (dA[pair] if pair != e else 0.0)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (e, fd, an)


@pytest.mark.gpu
def test_autograd_function_real(device):
    from clique_amd.autograd import sparse_solve
    ip, ind, v, dims = problem("grf", 8, seed=2)
    N = ip.size - 1
    ft = plan_for(ip, ind, v, dims)
    vals = torch.tensor(v, device="cuda:0", requires_grad=True)
    b = torch.randn(N, 2, dtype=torch.float64, device="cuda:0", requires_grad=True)
    x = sparse_solve(vals, b, ft, "f32")
    loss = 0.5 * (x ** 2).sum()
    gv, gb = torch.autograd.grad(loss, (vals, b))
    # analytic: lambda = A^{-1} x, dL/dA = -lambda x^T on the pattern
    A = torch.sparse_csr_tensor(torch.tensor(ip), torch.tensor(ind, dtype=torch.int64), torch.tensor(v),
                                size=(N, N)).to_dense().to("cuda:0")
    xr = torch.linalg.solve(A, b.detach())
    lr = torch.linalg.solve(A.T.contiguous(), xr)
    rows = torch.tensor(np.repeat(np.arange(N), np.diff(ip)), device="cuda:0")
    cols = torch.tensor(ind, dtype=torch.long, device="cuda:0")
    gvr = -(lr[rows] * xr[cols]).sum(dim=1)
    assert torch.linalg.norm(gb - lr) <= 1e-10 * torch.linalg.norm(lr)
    assert torch.linalg.norm(gv - gvr) <= 1e-10 * torch.linalg.norm(gvr)


@pytest.mark.gpu
def test_autograd_function_complex_wirtinger(device):
    """PyTorch convention: for L = ||x||^2 with x = A^{-1} b (A complex
    symmetric), grad_b = A^{-H} (2x)/... checked against torch.linalg.solve."""
    from clique_amd.autograd import sparse_solve
    ip, ind, v, dims = problem("helm", 6)
    N = ip.size - 1
    ft = plan_for(ip, ind, v, dims, cplx=True)
    vals = torch.tensor(v, device="cuda:0", requires_grad=True)
    b = torch.randn(N, 2, dtype=torch.complex128, device="cuda:0", requires_grad=True)
    x = sparse_solve(vals, b, ft, "f64")
    loss = (x.abs() ** 2).sum()
    gv, gb = torch.autograd.grad(loss, (vals, b))
    A = torch.sparse_csr_tensor(torch.tensor(ip), torch.tensor(ind, dtype=torch.int64), torch.tensor(v),
                                size=(N, N)).to_dense().to("cuda:0").detach().requires_grad_(True)
    b2 = b.detach().clone().requires_grad_(True)
    x2 = torch.linalg.solve(A, b2)
    gA2, gb2 = torch.autograd.grad((x2.abs() ** 2).sum(), (A, b2))
    rows = torch.tensor(np.repeat(np.arange(N), np.diff(ip)), device="cuda:0")
    cols = torch.tensor(ind, dtype=torch.long, device="cuda:0")
    assert torch.linalg.norm(gb - gb2) <= 1e-9 * torch.linalg.norm(gb2)
    assert torch.linalg.norm(gv - gA2[rows, cols]) <= 1e-9 * torch.linalg.norm(gA2[rows, cols])
