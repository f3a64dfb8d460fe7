"""The LDL solves' fused panel steps (launch_panel_step) against the two
grouped products per panel they replace.

The triangular sweeps over the inverted SB-wide diagonal blocks
(FrontLowerForwardSolve / FrontLowerBackwardSolve, lower_solve/
local_front.hpp:35-82) run, for levels with fronts wider than one panel, as
one fused launch per panel (the panel's product and the update of the band
rows of its super-panel) plus one product per super-panel for the rows past
it; CLIQG_SOLVE_CHAIN sets the panels per super-panel, 0 restores two
products per panel. Only the summation order changes: FP64 backward error
stays <= 1e-13 on both paths, and the solutions agree to 1e-10 (relative,
max norm; rounding differences amplified by the conditioning of the PML
operator, measured 1.1e-11 on the 40^3 case; a wrong update would differ
at O(1)), and FP32
factor + FP64 refinement reaches 1e-10 in the same number of steps (+-1).
Cases cover several super-panels per front (chain 2: 256 columns), one
partial super-panel (the default 8 on a 1,600-column root), 1 and 5
right-hand sides (one partial group of 16), the conjugated backward sweep
of a Hermitian pivoted factor and 64-wide panels.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
import clique_amd as cq
from helpers import backward_error, indefinite_random, problem
kind, n, prec, ftype, k, out = sys.argv[2], int(sys.argv[3]), sys.argv[4], sys.argv[5], int(sys.argv[6]), sys.argv[7]
herm = kind == "herm"
if herm:
    ip, ind, v, dims = indefinite_random((n, n - 1, n - 2), seed=31, cplx=True, herm=True)
    cplx = True
else:
    ip, ind, v, dims = problem(kind, n) if kind != "pml" else problem("pml", n, pml=5)
    cplx = kind in ("pml", "helm")
N = ip.size - 1
sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=64)
ft = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx, conjugate=herm)
ft.factor(torch.tensor(v, device="cuda:0"), precision=prec, front_type=ftype)
rng = np.random.default_rng(9)
b = rng.standard_normal((N, k)) + (1j * rng.standard_normal((N, k)) if cplx else 0)
if prec == "f64":
    x = cq.colmajor(torch.tensor(b, device="cuda:0"))
    ft.solve(x)
    its = 0
else:
    x, info = ft.refine(cq.colmajor(torch.tensor(b, device="cuda:0")), tol=1e-10, max_iters=10)
    assert info["converged"]
    its = info["iterations"]
xs = x.cpu().numpy()
np.save(out, xs)
print("%d %.3e" % (its, backward_error(ip, ind, v, xs, b).max()))
"""


def run(chain, kind, n, prec, ftype, k, tmp, extra=None):
    env = dict(os.environ, CLIQG_SOLVE_CHAIN=str(chain), **(extra or {}))
    out = os.path.join(tmp, f"x{chain}.npy")
    r = subprocess.run([sys.executable, "-c", _SCRIPT, ROOT, kind, str(n), prec, ftype, str(k), out],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    its, eta = r.stdout.strip().splitlines()[-1].split()
    return np.load(out), int(its), float(eta)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,n,prec,ftype,k,chains,extra", [
    ("pml", 40, "f64", "ldl", 5, (2, 8), None),
    ("lap", 24, "f64", "ldl", 1, (3,), None),
    ("herm", 12, "f64", "ldl_intrapiv", 5, (2,), None),
    ("pml", 32, "f32", "ldl", 4, (2, 8), None),
    ("helm", 24, "f64", "ldl", 17, (4,), {"CLIQG_SOLVE_BLOCK": "64"}),
])
def test_fused_panel_steps_match_two_products(device, kind, n, prec, ftype, k, chains, extra):
    with tempfile.TemporaryDirectory() as tmp:
        x0, its0, eta0 = run(0, kind, n, prec, ftype, k, tmp, extra)
        scale = np.abs(x0).max()
        for c in chains:
            x, its, eta = run(c, kind, n, prec, ftype, k, tmp, extra)
            d = np.abs(x - x0).max() / scale
            print(f"chain {c}: rel diff {d:.2e}, eta {eta:.2e} (two products {eta0:.2e}), "
                  f"steps {its} ({its0})")
            if prec == "f64":
                assert d <= 1e-10, d
                assert eta <= 1e-13 and eta0 <= 1e-13, (eta, eta0)
            else:
                assert eta <= 1e-10 and abs(its - its0) <= 1, (eta, its, its0)
