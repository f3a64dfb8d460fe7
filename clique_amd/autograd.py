"""torch.autograd binding of the sparse solve x = A^{-1} b (test binding).

Forward factors A (values on A's stored pattern) and solves, refining in
FP64 when the factor is FP32. Backward calls the engine's adjoint through
the C ABI: lambda = A^{-T} g from the same factor (A^{-H} g for a Hermitian
plan), dL/db = lambda and dL/dA[e] = -lambda[row_e] x[col_e] on the stored
pattern (conj(x) for a Hermitian plan); each stored entry is an independent
parameter. Complex tensors follow PyTorch's conjugate-Wirtinger convention:
grad_b = A^{-H} g and grad_A[e] = -grad_b[row_e] conj(x[col_e]).

  * complex-symmetric plan (A^T = A): the engine's adjoint of conj(g) gives
    conj(A)^{-1} g = A^{-H} g after one more conjugation, and the conjugated
    dA is -grad_b[row] conj(x[col]);
  * Hermitian plan (A^H = A): the engine already returns A^{-1} g = A^{-H} g
    and -lambda[row] conj(x[col]), so g and the results pass straight through.

If the plan was refactored (with other values) between forward and backward,
backward refactors it with the values saved at forward time before the
adjoint, so the gradient always belongs to the factorization of this call.
"""
import torch

from . import DistSymmFrontTree, colmajor


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, plan: DistSymmFrontTree, precision="f64", tol=1e-12):
        vd = vals.detach()
        plan.factor(vd, precision=precision)
        bb = colmajor(b.detach().reshape(plan.N, -1))
        x, _ = plan.refine(bb, tol=tol, max_iters=40)
        ctx.plan = plan
        ctx.precision = precision
        ctx.factor_id = plan.factor_count
        ctx.save_for_backward(x, vd)
        ctx.shape = b.shape
        return x.reshape(b.shape)

    @staticmethod
    def backward(ctx, gx):
        x, vd = ctx.saved_tensors
        plan = ctx.plan
        if plan.factor_count != ctx.factor_id:
            # another factorization replaced this call's factor: restore it
            plan.factor(vd, precision=ctx.precision)
        g = colmajor(gx.reshape(plan.N, -1).contiguous())
        if torch.is_complex(g) and not plan.conjugate:
            lam, dA = plan.adjoint(colmajor(g.conj().resolve_conj()), x)
            lam, dA = lam.conj().resolve_conj(), dA.conj().resolve_conj()
        else:
            lam, dA = plan.adjoint(g, x)
        return dA, lam.reshape(ctx.shape), None, None, None


def sparse_solve(vals, b, plan, precision="f64", tol=1e-12):
    """x = A(vals)^{-1} b, differentiable in vals and b."""
    return SparseSolve.apply(vals, b, plan, precision, tol)
