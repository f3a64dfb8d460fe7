"""Loader and build driver of libcliqg.so (the HIP extension, built in-tree)."""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcliqg.so")
_CSRC = os.path.join(_HERE, "csrc")
_lib = None

_ERRORS = {1: "invalid argument", 2: "singular pivot", 3: "refinement did not converge",
           4: "out of memory", 5: "HIP runtime error", 6: "wrong plan state"}


class CliqgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cliqg error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class ProfileEntry(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int),
                ("total_ms", ctypes.c_double), ("flops", ctypes.c_double),
                ("bytes", ctypes.c_double)]


class RefineInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("max_backward_error", ctypes.c_double)]


def build(jobs: int = 8) -> str:
    """Compile libcliqg.so for gfx950 with hipcc (always recompiles what changed)."""
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", _CSRC],
                          stdout=subprocess.DEVNULL)
    return LIB_PATH


# every symbol include/cliqg.h declares
EXPORTS = [
    "cliqg_last_error", "cliqg_version", "cliqg_device_malloc", "cliqg_device_free",
    "cliqg_memcpy", "cliqg_device_synchronize", "cliqg_symbolic_natural",
    "cliqg_symbolic_concat", "cliqg_symbolic_destroy", "cliqg_symbolic_num_nodes",
    "cliqg_symbolic_node", "cliqg_symbolic_node_array", "cliqg_symbolic_map",
    "cliqg_symbolic_work", "cliqg_plan_create", "cliqg_plan_destroy", "cliqg_plan_memory",
    "cliqg_plan_factor", "cliqg_plan_solve", "cliqg_plan_solve_nodal", "cliqg_plan_multiply",
    "cliqg_spmv",
    "cliqg_plan_refine", "cliqg_plan_adjoint", "cliqg_plan_get_front",
    "cliqg_plan_last_factor_ms", "cliqg_plan_set_profiling", "cliqg_plan_get_profile",
    "cliqg_plan_set_graphs", "cliqg_plan_solve_noalloc", "cliqg_plan_set_front",
    "cliqg_permute_rows", "cliqg_col_norms", "cliqg_plan_get_front_pivots",
    "cliqg_plan_autotune", "cliqg_autotune_table", "cliqg_symbolic_graph",
]


def lib():
    """The loaded library; raises if the HIP extension has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libcliqg.so is not built ({LIB_PATH}); run __graft_entry__.build() "
            "or `make -C clique_amd/csrc` -- there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, i64p, dp = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double))
    c_int, c_double, c_size = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    sig = {
        "cliqg_last_error": (ctypes.c_char_p, []),
        "cliqg_version": (c_int, [ip, ip]),
        "cliqg_device_malloc": (c_int, [ctypes.POINTER(vp), c_size]),
        "cliqg_device_free": (c_int, [vp]),
        "cliqg_memcpy": (c_int, [vp, vp, c_size, c_int]),
        "cliqg_device_synchronize": (c_int, []),
        "cliqg_symbolic_natural": (c_int, [c_int, c_int, c_int, c_int, i64p, ip, c_int,
                                           ctypes.POINTER(vp)]),
        "cliqg_symbolic_concat": (c_int, [ctypes.POINTER(vp), c_int, ctypes.POINTER(vp)]),
        "cliqg_symbolic_graph": (c_int, [c_int, i64p, ip, c_int, ctypes.POINTER(vp)]),
        "cliqg_symbolic_destroy": (c_int, [vp]),
        "cliqg_symbolic_num_nodes": (c_int, [vp, ip, ip]),
        "cliqg_symbolic_node": (c_int, [vp, c_int, ip]),
        "cliqg_symbolic_node_array": (c_int, [vp, c_int, c_int, ip]),
        "cliqg_symbolic_map": (c_int, [vp, ip]),
        "cliqg_symbolic_work": (c_int, [vp, c_int, dp, dp, dp]),
        "cliqg_plan_create": (c_int, [vp, i64p, ip, c_int, c_int, vp, ctypes.POINTER(vp)]),
        "cliqg_plan_destroy": (c_int, [vp]),
        "cliqg_plan_memory": (c_int, [vp, ctypes.POINTER(c_size)]),
        "cliqg_plan_factor": (c_int, [vp, vp, c_int, c_int]),
        "cliqg_plan_solve": (c_int, [vp, vp, c_int, c_int]),
        "cliqg_plan_solve_nodal": (c_int, [vp, vp, c_int, c_int]),
        "cliqg_plan_solve_noalloc": (c_int, [vp, vp, c_int, c_int]),
        "cliqg_plan_multiply": (c_int, [vp, c_double, vp, c_int, c_double, vp, c_int, c_int]),
        "cliqg_spmv": (c_int, [c_int, vp, vp, vp, c_int, c_double, vp, c_int, c_double, vp, c_int,
                                c_int, vp]),
        "cliqg_plan_refine": (c_int, [vp, vp, c_int, vp, c_int, c_int, c_double, c_int,
                                      ctypes.POINTER(RefineInfo)]),
        "cliqg_plan_adjoint": (c_int, [vp, vp, c_int, vp, c_int, c_int, vp, c_int, vp]),
        "cliqg_plan_get_front": (c_int, [vp, c_int, vp, vp]),
        "cliqg_plan_set_front": (c_int, [vp, c_int, vp]),
        "cliqg_plan_get_front_pivots": (c_int, [vp, c_int, vp, ip]),
        "cliqg_plan_autotune": (c_int, [vp, vp, c_int, c_int, c_int, ip, ip]),
        "cliqg_autotune_table": (c_int, [ctypes.c_char_p, c_size, ctypes.POINTER(c_size)]),
        "cliqg_permute_rows": (c_int, [c_int, c_int, vp, c_int, vp, c_int, vp, c_int, c_int, vp]),
        "cliqg_col_norms": (c_int, [c_int, c_int, vp, c_int, c_int, dp, vp]),
        "cliqg_plan_last_factor_ms": (c_int, [vp, dp]),
        "cliqg_plan_set_profiling": (c_int, [vp, c_int]),
        "cliqg_plan_set_graphs": (c_int, [vp, c_int]),
        "cliqg_plan_get_profile": (c_int, [vp, ctypes.POINTER(ProfileEntry), c_int, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != 0:
        msg = lib().cliqg_last_error()
        raise CliqgError(status, msg.decode() if msg else "")
