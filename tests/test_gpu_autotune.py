"""The GEMM tile autotuner (cliqg_plan_autotune) and every tile shape it
chooses among.

Tile shapes change the blocking of the grouped GEMMs, not the algorithm, so
the factor and the solves keep their accuracy: FP64 fronts within 1e-12 of
the oracle and backward error <= 1e-13 (complex) after tuning; in every
forced shape, FP32 fronts as close to the oracle as with the default shapes
(within a factor of three) and FP64 refinement to 1e-10.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import clique_amd as cq
import oracle
from helpers import backward_error, problem

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def to_dev(a):
    return cq.colmajor(torch.tensor(np.asarray(a), device="cuda:0"))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,cut,cplx,precision",
                         [("helm", 16, 64, True, "f64"), ("lap", 20, 128, False, "f64"),
                          ("pml", 24, 128, True, "f32")])
def test_autotune_keeps_results(device, kind, n, cut, cplx, precision):
    ip, ind, v, dims = problem(kind, n)
    N = ip.size - 1
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cut)
    ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
    vals = torch.tensor(v, device="cuda:0")
    nb, nc = ft.autotune(vals, precision=precision, rhs=4)
    assert nb > 0
    table = cq.autotune_table()
    assert len(table) >= nb
    for entry in table.values():
        assert entry["shape"] in entry["ms"] and entry["ms"]
        assert entry["ms"][entry["shape"]] == min(entry["ms"].values())
    rng = np.random.default_rng(3)
    b = rng.standard_normal((N, 4)) + (1j * rng.standard_normal((N, 4)) if cplx else 0)
    if precision == "f64":
        # the tuned plan is factored (autotune factors again with the tuned
        # shapes): fronts against the oracle, solves to rounding
        t = oracle.Tree(*dims, ip, ind, cutoff=cut)
        f = oracle.Factor(t, is_complex=cplx)
        f.initialize(ip, ind, v)
        f.ldl(nb=96)
        worst = 0.0
        for s in range(sym.num_nodes):
            Lg, dg = ft.front(s)
            Lo, do = f.front(s)
            worst = max(worst, np.abs(Lg - Lo).max(initial=0) / max(np.abs(Lo).max(initial=0), 1),
                        np.abs(dg - do).max(initial=0) / np.abs(do).max(initial=1))
        assert worst <= 1e-12, worst
        x = to_dev(b)
        ft.solve(x)
        assert backward_error(ip, ind, v, x.cpu().numpy(), b).max() <= 1e-13
    else:
        x, info = ft.refine(to_dev(b), tol=1e-10, max_iters=3)
        assert info["converged"]
        assert backward_error(ip, ind, v, x.cpu().numpy(), b).max() <= 1e-10
    # tuning again, and refactoring, reuse the table
    ft.factor(vals, precision=precision)


_FORCED = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import clique_amd as cq
import oracle
from helpers import backward_error, problem
kind, n, cplx, prec = sys.argv[2], int(sys.argv[3]), sys.argv[4] == "1", sys.argv[5]
ip, ind, v, dims = problem(kind, n)
N = ip.size - 1
sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=128)
ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
ft.set_profiling(True)
ft.factor(torch.tensor(v, device="cuda:0"), precision=prec)
prof = ft.profile()
ft.set_profiling(False)
t = oracle.Tree(*dims, ip, ind, cutoff=128)
f = oracle.Factor(t, is_complex=cplx)
f.initialize(ip, ind, v)
f.ldl(nb=96)
worst = 0.0
for s in range(sym.num_nodes):
    Lg, dg = ft.front(s)
    Lo, do = f.front(s)
    worst = max(worst, np.abs(Lg - Lo).max(initial=0) / max(np.abs(Lo).max(initial=0), 1),
                np.abs(dg - do).max(initial=0) / np.abs(do).max(initial=1))
rng = np.random.default_rng(5)
b = rng.standard_normal((N, 3)) + (1j * rng.standard_normal((N, 3)) if cplx else 0)
x, info = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-10, max_iters=8)
eta = backward_error(ip, ind, v, x.cpu().numpy(), b).max()
print(sorted(prof))
print("%d %.3e %.3e" % (info["iterations"], worst, eta))
"""


@pytest.mark.gpu
@pytest.mark.parametrize("shape,kind,n,cplx,prec,family",
                         [("6", "pml", 24, True, "f32", "gemm_update_w8"),
                          ("6", "helm", 20, True, "f32", "gemm_update_w8"),
                          ("4", "lap", 24, False, "f64", "gemm_update_large"),
                          ("4", "grf", 20, False, "f32", "gemm_update_large"),
                          ("5", "helm", 16, True, "f64", "gemm_update_small"),
                          ("7", "pml", 24, True, "f32", "gemm_update_k64"),
                          ("7", "lap", 24, False, "f64", "gemm_update_k64")])
def test_every_tile_shape_keeps_results(device, shape, kind, n, cplx, prec, family):
    """Each tile shape, forced on every launch that admits it
    (CLIQG_GEMM_TILE, read once per process, hence subprocesses), against
    the default shapes on the same problem: FP64 fronts within 1e-12 of the
    oracle either way; FP32 fronts within three times the default shapes'
    distance to the oracle (the blocking changes only the rounding), and
    FP64 refinement to 1e-10 in at most one step more than the default."""
    def run(env_extra):
        env = dict(os.environ, **env_extra)
        if not env_extra:
            env.pop("CLIQG_GEMM_TILE", None)
        r = subprocess.run([sys.executable, "-c", _FORCED, ROOT, kind, str(n),
                            "1" if cplx else "0", prec], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        lines = r.stdout.strip().splitlines()
        its, worst, eta = lines[-1].split()
        return lines[-2], int(its), float(worst), float(eta)

    fams0, its0, worst0, eta0 = run({})
    fams, its, worst, eta = run({"CLIQG_GEMM_TILE": shape})
    print(f"default: {its0} steps, fronts {worst0:.2e}; shape {shape}: {its} steps, "
          f"fronts {worst:.2e}")
    assert family in fams, fams
    if prec == "f64":
        assert worst0 <= 1e-12 and worst <= 1e-12, (worst0, worst)
    else:
        assert worst <= 3 * worst0, (worst0, worst)
    assert eta <= 1e-10 and its <= its0 + 1, (its0, its, eta)


_STREAM = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import clique_amd as cq
import oracle
from helpers import backward_error, hermitian_matrix, problem
kind, n, k = sys.argv[2], int(sys.argv[3]), int(sys.argv[4])
conj = kind == "herm"
if conj:
    ip, ind, v, dims = hermitian_matrix(n)
else:
    ip, ind, v, dims = problem(kind, n)
cplx = np.iscomplexobj(v)
N = ip.size - 1
sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=64)
ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx, conjugate=conj)
ft.factor(torch.tensor(v, device="cuda:0"), precision="f32")
rng = np.random.default_rng(7)
b = rng.standard_normal((N, k)) + (1j * rng.standard_normal((N, k)) if cplx else 0)
x = cq.colmajor(torch.tensor(b, device="cuda:0"))
ft.set_profiling(True)
ft.solve(x)
prof = ft.profile()
ft.set_profiling(False)
t = oracle.Tree(*dims, ip, ind, cutoff=64)
f = oracle.Factor(t, is_complex=cplx, conjugate=conj)
f.initialize(ip, ind, v)
f.ldl(nb=96)
xo = f.solve(b)
err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
xr, info = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-10, max_iters=8)
eta = backward_error(ip, ind, v, xr.cpu().numpy(), b).max()
print(sorted(prof))
print("%d %.3e %.3e" % (info["iterations"], err, eta))
"""


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,k", [("pml", 24, 64), ("pml", 24, 20), ("lap", 24, 64),
                                      ("herm", 14, 33)])
def test_streaming_solve_tiles_keep_results(device, kind, n, k):
    """The streaming solve tiles (TILE_STREAM: op(A) read from global memory
    straight into the MFMAs), forced on every solve product that admits them
    (plain and transposed A, conjugated for the Hermitian case, N = k <= 64
    right-hand sides), against the default shapes: one FP32 solve within
    three times the default shapes' distance to the oracle's FP64 solution,
    and FP64 refinement to 1e-10 in at most one step more."""
    def run(env_extra):
        env = dict(os.environ, **env_extra)
        if not env_extra:
            env.pop("CLIQG_GEMM_TILE", None)
        r = subprocess.run([sys.executable, "-c", _STREAM, ROOT, kind, str(n), str(k)], env=env,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        lines = r.stdout.strip().splitlines()
        its, err, eta = lines[-1].split()
        return lines[-2], int(its), float(err), float(eta)

    fams0, its0, err0, eta0 = run({})
    fams, its, err, eta = run({"CLIQG_GEMM_TILE": "8"})
    print(f"default: {its0} steps, solve error {err0:.2e}; streaming: {its} steps, "
          f"solve error {err:.2e}")
    assert "solve_gemm_stream" in fams, fams
    assert err <= 3 * err0, (err0, err)
    assert eta <= 1e-10 and its <= its0 + 1, (its0, its, eta)
