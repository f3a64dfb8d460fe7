"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a plain-C restatement of Clique's world-size-1 multifrontal
LDL^T path (see oracle.h for the reference file:line of every function).
Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module; the product library never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OR_INDS, OR_LOWER, OR_ORIG_LOWER, OR_ORIG_LOWER_REL, OR_LEFT_REL, OR_RIGHT_REL = range(6)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", _HERE],
                              stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, ip, lp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), \
            ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_double)
        L.or_natural_nd.restype = vp
        L.or_natural_nd.argtypes = [ctypes.c_int] * 4 + [lp, ip, ctypes.c_int]
        L.or_tree_free.argtypes = [vp]
        L.or_tree_num_nodes.argtypes = [vp]
        L.or_tree_node.argtypes = [vp, ctypes.c_int, ip]
        L.or_tree_array.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip]
        L.or_tree_map.argtypes = [vp, ip]
        L.or_factor_create.restype = vp
        L.or_factor_create.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_factor_free.argtypes = [vp]
        L.or_initialize.argtypes = [vp, ctypes.c_int, lp, ip, dp]
        L.or_ldl.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_solve.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int]
        L.or_plant.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_get_pivots.argtypes = [vp, ctypes.c_int, dp, ip]
        L.or_blas_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.or_ldl_blas.argtypes = [vp, ctypes.c_int]
        L.or_solve_blas.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int]
        L.or_get_front.argtypes = [vp, ctypes.c_int, dp, dp]
        L.or_multiply.argtypes = [ctypes.c_int, lp, ip, dp, ctypes.c_int, ctypes.c_double,
                                  dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int,
                                  ctypes.c_int]
        L.or_factor_work.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        _lib = L
    return _lib


_BLAS = None


def _blas_candidates():
    import glob
    cands = [("/opt/conda/lib/libmkl_rt.so", "", "MKL (libmkl_rt)")]
    try:
        import scipy
        root = os.path.dirname(os.path.dirname(scipy.__file__))
        for p in sorted(glob.glob(os.path.join(root, "scipy.libs", "libscipy_openblas*.so"))):
            cands.append((p, "scipy_", "OpenBLAS (scipy bundle)"))
    except ImportError:
        pass
    return cands


def blas_load(threads: int) -> str:
    """Load a multithreaded CPU BLAS for the CPU-baseline path (MKL if
    present, else scipy's bundled OpenBLAS); returns its name. In a process
    that has imported torch, MKL is skipped: torch's own MKL and OpenMP
    symbols take the place of a second MKL's and its results are wrong
    (scipy's OpenBLAS has prefixed symbols and is unaffected)."""
    import sys
    global _BLAS
    if _BLAS is not None:
        return _BLAS
    for path, prefix, name in _blas_candidates():
        if prefix == "" and "torch" in sys.modules:
            continue
        if os.path.exists(path) and lib().or_blas_load(path.encode(), prefix.encode(),
                                                      int(threads)) == 0:
            _BLAS = f"{name}, {threads} threads"
            return _BLAS
    raise RuntimeError("no CPU BLAS could be loaded for the CPU baseline")


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _as_f64(vals):
    vals = np.ascontiguousarray(vals)
    if np.iscomplexobj(vals):
        return np.ascontiguousarray(vals.astype(np.complex128)).view(np.float64), True
    return np.ascontiguousarray(vals, dtype=np.float64), False


class Tree:
    """Natural nested dissection + symbolic analysis (oracle restatement)."""

    def __init__(self, nx, ny, nz, indptr, indices, cutoff=128):
        self.N = int(nx * ny * nz)
        self._indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self._indices = np.ascontiguousarray(indices, dtype=np.int32)
        h = lib().or_natural_nd(nx, ny, nz, self.N, _lp(self._indptr), _ip(self._indices), cutoff)
        if not h:
            raise ValueError("oracle nested dissection rejected its input")
        self._h = ctypes.c_void_p(h)
        self.num = lib().or_tree_num_nodes(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_tree_free(self._h)
            self._h = None

    def node(self, s):
        info = np.zeros(8, np.int32)
        lib().or_tree_node(self._h, s, _ip(info))
        keys = ("size", "off", "parent", "child0", "child1", "lower_size", "orig_lower_size",
                "on_left")
        return dict(zip(keys, (int(v) for v in info)))

    def array(self, s, which):
        n = lib().or_tree_array(self._h, s, which, ctypes.POINTER(ctypes.c_int)())
        out = np.zeros(max(n, 1), np.int32)
        lib().or_tree_array(self._h, s, which, _ip(out))
        return out[:n].copy()

    def map(self):
        m = np.zeros(self.N, np.int32)
        lib().or_tree_map(self._h, _ip(m))
        return m

    def work(self, is_complex=False):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().or_factor_work(self._h, int(is_complex), ctypes.byref(a), ctypes.byref(b),
                             ctypes.byref(c))
        return a.value, b.value, c.value


class Factor:
    """Multifrontal FP64 LDL^T of the oracle (double or complex double)."""

    def __init__(self, tree: Tree, is_complex=False, conjugate=False):
        self.tree = tree
        self.is_complex = bool(is_complex)
        h = lib().or_factor_create(tree._h, int(is_complex), int(conjugate))
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_factor_free(self._h)
            self._h = None

    def initialize(self, indptr, indices, vals):
        self._indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self._indices = np.ascontiguousarray(indices, dtype=np.int32)
        v, cplx = _as_f64(vals)
        if cplx != self.is_complex:
            raise TypeError("value type does not match the factor's field")
        self._vals = v
        r = lib().or_initialize(self._h, self.tree.N, _lp(self._indptr), _ip(self._indices),
                                _dp(v))
        if r != 0:
            raise RuntimeError(f"or_initialize failed ({r})")

    def ldl(self, nb=96, block=False, nthreads=0, pivoted=False):
        """LocalLDL: FrontLDL (default), FrontBlockLDL (block),
        FrontLDLIntraPiv with Bunch-Kaufman pivoting (pivoted) or
        FrontBlockLDL with intra-front pivoting (both)."""
        mode = (3 if block else 2) if pivoted else int(bool(block))
        r = lib().or_ldl(self._h, nb, mode, nthreads)
        if r != 0:
            raise ZeroDivisionError(f"oracle LDL hit a zero pivot ({r})")

    def pivots(self, s):
        """(subdiag, piv) of front s of a pivoted factor."""
        n = self.tree.node(s)["size"]
        dt = np.complex128 if self.is_complex else np.float64
        sub = np.zeros(max(n, 1), dtype=dt)
        piv = np.zeros(max(n, 1), dtype=np.int32)
        lib().or_get_pivots(self._h, s, _dp(sub.view(np.float64)), _ip(piv))
        return sub[:n], piv[:n]

    def ldl_blas(self, nb=96):
        """LocalLDL with FrontLDL's dense calls made through the CPU BLAS
        loaded by blas_load() (the CPU baseline of the benchmark)."""
        r = lib().or_ldl_blas(self._h, nb)
        if r != 0:
            raise RuntimeError(f"oracle BLAS LDL failed ({r})")

    def solve_nodal_blas(self, x):
        """solve_nodal() with the front solves made through the CPU BLAS."""
        return self._solve_nodal(x, lib().or_solve_blas)

    def plant(self, front, panel):
        """Test hook: the next ldl() skips the trailing update of one panel
        of one front (a deliberate mistake the tolerances must catch)."""
        lib().or_plant(self._h, int(front), int(panel))

    def solve_nodal(self, x):
        """Solve A_nodal y = x for x in nodal ordering (N or N x k)."""
        return self._solve_nodal(x, lib().or_solve)

    def _solve_nodal(self, x, fn):
        dt = np.complex128 if self.is_complex else np.float64
        X = np.array(x, dtype=dt, order="F", copy=True)
        squeeze = X.ndim == 1
        if squeeze:
            X = X.reshape(-1, 1, order="F")
        k = X.shape[1]
        buf = X.view(np.float64) if not self.is_complex else X.reshape(-1, order="F").view(np.float64)
        r = fn(self._h, _dp(buf), self.tree.N, k)
        if r != 0:
            raise RuntimeError(f"or_solve failed ({r})")
        if self.is_complex:
            X = buf.view(np.complex128).reshape(self.tree.N, k, order="F")
        return X[:, 0] if squeeze else X

    def solve(self, b):
        """Solve A x = b in the original ordering (Pull, Solve, Push)."""
        mp = self.tree.map()
        B = np.asarray(b)
        xn = np.empty_like(B, dtype=np.complex128 if self.is_complex else np.float64)
        xn[mp] = B
        yn = self.solve_nodal(xn)
        return yn[mp]

    def front(self, s):
        nd = self.tree.node(s)
        n, m = nd["size"], nd["size"] + nd["lower_size"]
        dt = np.complex128 if self.is_complex else np.float64
        L = np.zeros((m, n), dtype=dt, order="F")
        d = np.zeros(max(n, 1), dtype=dt)
        lib().or_get_front(self._h, s, _dp(L.reshape(-1, order="F").view(np.float64)),
                           _dp(d.view(np.float64)))
        return L, d[:n]


def multiply(indptr, indices, vals, x, alpha=1.0, beta=0.0, y=None):
    """y := alpha A x + beta y with the oracle's scalar loop."""
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    v, cplx = _as_f64(vals)
    dt = np.complex128 if cplx else np.float64
    X = np.array(x, dtype=dt, order="F")
    squeeze = X.ndim == 1
    if squeeze:
        X = X.reshape(-1, 1, order="F")
    N, k = X.shape
    Y = np.zeros((N, k), dtype=dt, order="F") if y is None else np.array(y, dtype=dt, order="F").reshape(N, k, order="F")
    xb = X.reshape(-1, order="F").view(np.float64)
    yb = Y.reshape(-1, order="F").view(np.float64)
    lib().or_multiply(N, _lp(indptr), _ip(indices), _dp(v), int(cplx), alpha, _dp(xb), N,
                      beta, _dp(yb), N, k)
    Y = yb.view(dt).reshape(N, k, order="F")
    return Y[:, 0] if squeeze else Y
