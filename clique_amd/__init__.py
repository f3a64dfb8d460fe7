"""clique_amd: Python test binding of libcliqg, the MI355X multifrontal LDL^T engine.

Everything here calls the C ABI declared in include/cliqg.h through ctypes;
PyTorch only supplies device memory and the stream. Names mirror the
reference's C++ surface (namespace cliq):

    NaturalNestedDissection   include/clique/symbolic/natural_nested_dissection.hpp:17-26
    DistSymmFrontTree         include/clique/numeric/dist_symm_front_tree/decl.hpp:206-249
    LDL                       include/clique/numeric/ldl.hpp:23-26
    Solve                     include/clique/numeric/solve.hpp:17-24
    Multiply                  include/clique/numeric/multiply.hpp:17-21
    SymmetricSolve            include/clique/numeric/solve.hpp:94-116

There is no CPU fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref

import numpy as np

from . import _lib
from ._lib import CliqgError, build, lib  # noqa: F401

# SymmFrontType values accepted by LDL (decl.hpp:17-47); 1D/2D coincide at P=1
LDL_1D = LDL_2D = "ldl"
BLOCK_LDL_1D = BLOCK_LDL_2D = "block_ldl"
LDL_INTRAPIV_1D = LDL_INTRAPIV_2D = "ldl_intrapiv"
BLOCK_LDL_INTRAPIV_1D = BLOCK_LDL_INTRAPIV_2D = "block_ldl_intrapiv"

_FRONT_TYPES = {"ldl": 0, "block_ldl": 1, "ldl_intrapiv": 2, "block_ldl_intrapiv": 3}
_PRECISIONS = {"f64": 0, "fp64": 0, "double": 0, "f32": 1, "fp32": 1, "single": 1}


_LIVE = weakref.WeakSet()


@atexit.register
def _release_all():
    """Destroy live plans and symbolics while the HIP runtime is still up."""
    for obj in list(_LIVE):
        try:
            obj.close()
        except Exception:
            pass


def _dev_ptr(t):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (HIP) tensor")
    return ctypes.c_void_p(t.data_ptr())


def _check_rhs(t, N, complex_):
    import torch
    want = torch.complex128 if complex_ else torch.float64
    if t.dtype != want:
        raise TypeError(f"right-hand side must be {want}, got {t.dtype}")
    if t.dim() == 1:
        t2 = t.view(-1, 1)
    else:
        t2 = t
    if t2.shape[0] != N:
        raise ValueError(f"right-hand side has {t2.shape[0]} rows, expected {N}")
    # column-major N x k == row-major k x N storage: require columns contiguous
    if t2.shape[1] > 1 and not (t2.stride(0) == 1):
        raise ValueError("right-hand sides must be column-major (use .t().contiguous().t())")
    ld = t2.stride(1) if t2.shape[1] > 1 else N
    return t2, int(ld), int(t2.shape[1])


def colmajor(t):
    """Return a column-major (Fortran-ordered) copy of a 2-D tensor."""
    return t.t().contiguous().t() if t.dim() == 2 else t.contiguous()


class Symbolic:
    """Separator tree, reordering map and DistSymmInfo (symbolic analysis)."""

    def __init__(self, handle, N):
        self._h = handle
        self.N = N
        n = ctypes.c_int()
        lib().cliqg_symbolic_num_nodes(self._h, ctypes.byref(n), None)
        self.num_nodes = n.value
        _LIVE.add(self)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            lib().cliqg_symbolic_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def node(self, s):
        info = (ctypes.c_int * 8)()
        _lib.check(lib().cliqg_symbolic_node(self._h, s, info))
        keys = ("size", "off", "parent", "child0", "child1", "lower_size", "orig_lower_size",
                "on_left")
        return dict(zip(keys, list(info)))

    def node_array(self, s, which):
        nd = self.node(s)
        n = {0: nd["size"], 1: nd["lower_size"], 2: nd["orig_lower_size"],
             3: nd["orig_lower_size"],
             4: self.node(nd["child0"])["lower_size"] if nd["child0"] >= 0 else 0,
             5: self.node(nd["child1"])["lower_size"] if nd["child1"] >= 0 else 0}[which]
        out = np.zeros(max(n, 1), np.int32)
        _lib.check(lib().cliqg_symbolic_node_array(
            self._h, s, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out[:n].copy()

    def map(self):
        m = np.zeros(max(self.N, 1), np.int32)
        _lib.check(lib().cliqg_symbolic_map(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return m[:self.N]

    def work(self, is_complex=False):
        f, s, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _lib.check(lib().cliqg_symbolic_work(self._h, int(is_complex), ctypes.byref(f),
                                             ctypes.byref(s), ctypes.byref(e)))
        return f.value, s.value, e.value

    @staticmethod
    def concat(syms):
        arr = (ctypes.c_void_p * len(syms))(*[s._h for s in syms])
        h = ctypes.c_void_p()
        _lib.check(lib().cliqg_symbolic_concat(arr, len(syms), ctypes.byref(h)))
        return Symbolic(h, sum(s.N for s in syms))


def NaturalNestedDissection(nx, ny, nz, indptr, indices, cutoff=128):
    """Natural nested dissection of an nx*ny*nz grid graph given as CSR."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ind = np.ascontiguousarray(indices, dtype=np.int32)
    N = ip.size - 1
    if N < 0 or ind.size < ip[-1]:
        raise ValueError("malformed CSR arrays")
    h = ctypes.c_void_p()
    _lib.check(lib().cliqg_symbolic_natural(
        int(nx), int(ny), int(nz), N, ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        ind.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(cutoff), ctypes.byref(h)))
    return Symbolic(h, N)


def NestedDissection(indptr, indices, cutoff=128):
    """Nested dissection of a general graph given as CSR (the reference's
    NestedDissection, symbolic/nested_dissection.hpp:491-533, with this
    library's own bisection in place of METIS)."""
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ind = np.ascontiguousarray(indices, dtype=np.int32)
    N = ip.size - 1
    if N <= 0 or ind.size < ip[-1]:
        raise ValueError("malformed CSR arrays")
    h = ctypes.c_void_p()
    _lib.check(lib().cliqg_symbolic_graph(
        N, ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
        ind.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(cutoff), ctypes.byref(h)))
    return Symbolic(h, N)


class DistSymmFrontTree:
    """Device front tree (a plan): A's pattern, scatter map and level schedule."""

    def __init__(self, sym: Symbolic, indptr, indices, is_complex=False, conjugate=False,
                 stream=None):
        import torch
        self.sym = sym
        self.N = sym.N
        self.is_complex = bool(is_complex)
        self.conjugate = bool(conjugate)
        self._indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self._indices = np.ascontiguousarray(indices, dtype=np.int32)
        self.nnz = int(self._indptr[-1] - self._indptr[0])
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream().cuda_stream
        h = ctypes.c_void_p()
        _lib.check(lib().cliqg_plan_create(
            sym._h, self._indptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            self._indices.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            1 if is_complex else 0, int(conjugate), ctypes.c_void_p(stream or 0), ctypes.byref(h)))
        self._h = h
        self.front_type = None
        self.precision = None
        self.factor_count = 0   # numeric factorizations so far (refactoring changes it)
        _LIVE.add(self)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._h = None
            lib().cliqg_plan_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- numeric -------------------------------------------------------------
    def factor(self, vals, precision="f64", front_type="ldl"):
        import torch
        want = torch.complex128 if self.is_complex else torch.float64
        if vals.dtype != want or not vals.is_cuda or vals.numel() != self.nnz:
            raise TypeError(f"values must be a CUDA {want} tensor of {self.nnz} entries")
        vals = vals.contiguous()
        _lib.check(lib().cliqg_plan_factor(self._h, _dev_ptr(vals), _PRECISIONS[precision],
                                           _FRONT_TYPES[front_type]))
        self.precision = precision
        self.front_type = front_type
        self._vals = vals
        self.factor_count += 1

    def autotune(self, vals, precision="f64", front_type="ldl", rhs=0):
        """Time every admissible GEMM tile shape of this factorization (and of
        a solve with `rhs` right-hand sides), keep the fastest per launch
        bucket in the process-wide table, and factor `vals` again with the
        tuned shapes. Returns (buckets measured, buckets changed)."""
        import torch
        want = torch.complex128 if self.is_complex else torch.float64
        if vals.dtype != want or not vals.is_cuda or vals.numel() != self.nnz:
            raise TypeError(f"values must be a CUDA {want} tensor of {self.nnz} entries")
        vals = vals.contiguous()
        nb, nc = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib().cliqg_plan_autotune(self._h, _dev_ptr(vals), _PRECISIONS[precision],
                                             _FRONT_TYPES[front_type], int(rhs), ctypes.byref(nb),
                                             ctypes.byref(nc)))
        self.precision = precision
        self.front_type = front_type
        self._vals = vals
        self.factor_count += 1
        return nb.value, nc.value

    def solve(self, x, nodal=False):
        t, ld, k = _check_rhs(x, self.N, self.is_complex)
        fn = lib().cliqg_plan_solve_nodal if nodal else lib().cliqg_plan_solve
        _lib.check(fn(self._h, _dev_ptr(t), ld, k))
        return x

    def multiply(self, x, y=None, alpha=1.0, beta=0.0):
        import torch
        t, ld, k = _check_rhs(x, self.N, self.is_complex)
        if y is None:
            y = torch.zeros_like(x)
        ty, ldy, _ = _check_rhs(y, self.N, self.is_complex)
        _lib.check(lib().cliqg_plan_multiply(self._h, float(alpha), _dev_ptr(t), ld, float(beta),
                                             _dev_ptr(ty), ldy, k))
        return y

    def refine(self, b, x=None, tol=1e-10, max_iters=20, raise_on_fail=True):
        import torch
        tb, ldb, k = _check_rhs(b, self.N, self.is_complex)
        if x is None:
            x = torch.empty_like(b)
        tx, ldx, _ = _check_rhs(x, self.N, self.is_complex)
        info = _lib.RefineInfo()
        st = lib().cliqg_plan_refine(self._h, _dev_ptr(tb), ldb, _dev_ptr(tx), ldx, k,
                                     float(tol), int(max_iters), ctypes.byref(info))
        if st != 0 and (raise_on_fail or st != 3):
            _lib.check(st)
        return x, {"iterations": info.iterations, "converged": bool(info.converged),
                   "backward_error": info.max_backward_error}

    def adjoint(self, g, x):
        """lambda = A^{-T} g and dL/dA on A's pattern (one value per stored nonzero)."""
        import torch
        tg, ldg, k = _check_rhs(g, self.N, self.is_complex)
        tx, ldx, _ = _check_rhs(x, self.N, self.is_complex)
        lam = colmajor(torch.empty_like(g))
        tl, ldl, _ = _check_rhs(lam, self.N, self.is_complex)
        dA = torch.empty(self.nnz, dtype=g.dtype, device=g.device)
        _lib.check(lib().cliqg_plan_adjoint(self._h, _dev_ptr(tg), ldg, _dev_ptr(tx), ldx, k,
                                            _dev_ptr(tl), ldl, _dev_ptr(dA)))
        return lam, dA

    def front(self, s):
        nd = self.sym.node(s)
        n, m = nd["size"], nd["size"] + nd["lower_size"]
        dt = np.complex128 if self.is_complex else np.float64
        L = np.zeros((m, n), dtype=dt, order="F")
        d = np.zeros(max(n, 1), dtype=dt)
        _lib.check(lib().cliqg_plan_get_front(self._h, s, L.ctypes.data_as(ctypes.c_void_p),
                                              d.ctypes.data_as(ctypes.c_void_p)))
        if self.front_type in ("block_ldl", "block_ldl_intrapiv"):
            # block fronts: ATL holds inv(L11 D L11^T), ABL the original A21, no diagonal
            return L, d[:0]
        return L, d[:n]

    def pivots(self, s):
        """D's subdiagonal and the pivot permutation of front s of a pivoted
        factorization (localFronts[s].subdiag and .piv, ldl/local.hpp:108-115):
        sub[i] = D(i+1, i) (nonzero where a 2x2 pivot starts), and
        (P ATL P^T)(i, j) = ATL(piv[i], piv[j])."""
        n = self.sym.node(s)["size"]
        dt = np.complex128 if self.is_complex else np.float64
        sub = np.zeros(max(n, 1), dtype=dt)
        piv = np.zeros(max(n, 1), dtype=np.int32)
        _lib.check(lib().cliqg_plan_get_front_pivots(
            self._h, s, sub.ctypes.data_as(ctypes.c_void_p),
            piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return sub[:n], piv[:n]

    def set_front(self, s, L):
        """Stage front s's unfactored values (an m x n array, FP64 field type)
        for the next factor(), which uses them in place of A's assembled
        values for that front (tests/Refactor.cpp:150-160 writes
        localFronts[s].frontL and factors again)."""
        nd = self.sym.node(s)
        n, m = nd["size"], nd["size"] + nd["lower_size"]
        dt = np.complex128 if self.is_complex else np.float64
        A = np.asfortranarray(np.asarray(L, dtype=dt))
        if A.shape != (m, n):
            raise ValueError(f"front {s} is {m} x {n}, got {A.shape}")
        _lib.check(lib().cliqg_plan_set_front(self._h, s, A.ctypes.data_as(ctypes.c_void_p)))

    def memory(self):
        b = ctypes.c_size_t()
        _lib.check(lib().cliqg_plan_memory(self._h, ctypes.byref(b)))
        return b.value

    def set_graphs(self, enable=True):
        """Replay the factor and solve launch sequences as HIP graphs (default off)."""
        _lib.check(lib().cliqg_plan_set_graphs(self._h, int(enable)))

    def set_profiling(self, enable=True):
        _lib.check(lib().cliqg_plan_set_profiling(self._h, int(enable)))

    def profile(self):
        """{kernel family: {launches, ms, flops, bytes}} since profiling was enabled."""
        n = ctypes.c_int()
        _lib.check(lib().cliqg_plan_get_profile(self._h, None, 0, ctypes.byref(n)))
        arr = (_lib.ProfileEntry * max(n.value, 1))()
        _lib.check(lib().cliqg_plan_get_profile(self._h, arr, n.value, ctypes.byref(n)))
        return {e.name.decode(): {"launches": e.launches, "ms": e.total_ms, "flops": e.flops,
                                  "bytes": e.bytes} for e in arr[:n.value]}

    def last_factor_ms(self):
        v = ctypes.c_double()
        _lib.check(lib().cliqg_plan_last_factor_ms(self._h, ctypes.byref(v)))
        return v.value


_SHAPES = {0: "64x64", 4: "128x128", 5: "32x32", 6: "128x128x8w", 7: "32x32k64", 8: "stream"}
_SHAPE_ORDER = ("32x32", "64x64", "128x128", "128x128x8w", "32x32k64", "stream")   # GEMM_VARIANTS


def autotune_table():
    """The GEMM tile autotuner's table: {bucket: {"shape": ..., "ms": {shape: ms}}}."""
    n = ctypes.c_size_t()
    _lib.check(lib().cliqg_autotune_table(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _lib.check(lib().cliqg_autotune_table(buf, n.value + 1, ctypes.byref(n)))
    out = {}
    for line in buf.value.decode().splitlines():
        key, v, *ms = line.split()
        out[key] = {"shape": _SHAPES.get(int(v), v),
                    "ms": {name: float(m) for name, m in zip(_SHAPE_ORDER, ms)
                           if float(m) >= 0}}
    return out


def LDL(info, L: DistSymmFrontTree, vals, front_type=LDL_2D, precision="f64"):
    """Factor the front tree (values given as a CUDA tensor in CSR order)."""
    L.factor(vals, precision=precision, front_type=front_type)


def Solve(info, L: DistSymmFrontTree, X, nodal=True):
    """Solve in place; X in nodal ordering (a DistNodalMultiVec) by default."""
    return L.solve(X, nodal=nodal)


def Multiply(alpha, L: DistSymmFrontTree, X, beta, Y):
    return L.multiply(X, Y, alpha=alpha, beta=beta)


def SymmetricSolve(nx, ny, nz, indptr, indices, vals, X, conjugate=False, cutoff=128,
                   precision="f64", front_type=LDL_INTRAPIV_1D):
    """Black-box solve: dissection, factorization and solve (original
    ordering). Like the reference (numeric/solve.hpp:95-117) it factors with
    intra-front Bunch-Kaufman pivoting (LDL_INTRAPIV_1D) unless told otherwise."""
    import torch
    sym = NaturalNestedDissection(nx, ny, nz, indptr, indices, cutoff)
    is_c = torch.is_complex(X)
    ft = DistSymmFrontTree(sym, indptr, indices, is_complex=is_c, conjugate=conjugate)
    ft.factor(vals, precision=precision, front_type=front_type)
    return ft.solve(X)
