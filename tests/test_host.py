"""Host-side checks of the product library (CPU only, no GPU calls).

- every function include/cliqg.h declares is exported by libcliqg.so;
- the product's C++ symbolic analysis equals the oracle's, bit for bit;
- argument validation returns status codes instead of crashing.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import clique_amd as cq
import oracle
from clique_amd import _lib
from clique_amd import problems as P
from helpers import problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cliqg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cliqg_\w+)\s*\(", src)))


def test_header_functions_are_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == decl


@pytest.mark.parametrize("dims,cut", [((6, 6, 6), 8), ((7, 5, 3), 4), ((30, 30, 30), 128),
                                      ((17, 1, 1), 2), ((12, 9, 1), 10), ((2, 1, 1), 1),
                                      ((1, 1, 1), 1), ((11, 3, 2), 1)])
def test_product_symbolic_equals_oracle(dims, cut):
    ip, ind, v, d = P.stencil7(*dims, 6.0)
    s = cq.NaturalNestedDissection(*d, ip, ind, cutoff=cut)
    o = oracle.Tree(*d, ip, ind, cutoff=cut)
    assert s.num_nodes == o.num
    assert np.array_equal(s.map(), o.map())
    for k in range(s.num_nodes):
        assert s.node(k) == o.node(k)
        for w in range(6):
            assert np.array_equal(s.node_array(k, w), o.array(k, w))


def test_product_symbolic_equals_oracle_on_config_graphs():
    for kind, n in (("pml", 14), ("lap2d", (40, 33)), ("grf", 9)):
        ip, ind, v, d = problem(kind, n)
        s = cq.NaturalNestedDissection(*d, ip, ind, cutoff=32)
        o = oracle.Tree(*d, ip, ind, cutoff=32)
        assert np.array_equal(s.map(), o.map())
        for k in range(0, s.num_nodes, 3):
            assert s.node(k) == o.node(k)
            assert np.array_equal(s.node_array(k, 1), o.array(k, oracle.OR_LOWER))


def test_work_matches_oracle():
    ip, ind, v, d = P.laplacian7(20)
    s = cq.NaturalNestedDissection(*d, ip, ind, cutoff=64)
    o = oracle.Tree(*d, ip, ind, cutoff=64)
    for c in (False, True):
        assert np.allclose(s.work(c), o.work(c), rtol=1e-15)


def test_concat_offsets():
    a = cq.NaturalNestedDissection(*P.stencil7(5, 4, 1, 4.0)[3], *P.stencil7(5, 4, 1, 4.0)[:2],
                                   cutoff=4)
    ip, ind, v, d = P.stencil7(3, 3, 3, 6.0)
    b = cq.NaturalNestedDissection(*d, ip, ind, cutoff=4)
    c = cq.Symbolic.concat([a, b])
    assert c.N == a.N + b.N and c.num_nodes == a.num_nodes + b.num_nodes
    assert np.array_equal(c.map()[:a.N], a.map())
    assert np.array_equal(c.map()[a.N:], b.map() + a.N)
    nb = b.node(b.num_nodes - 1)
    nc = c.node(c.num_nodes - 1)
    assert nc["off"] == nb["off"] + a.N and nc["child0"] == nb["child0"] + a.num_nodes


def test_invalid_arguments_return_status():
    ip, ind, v, d = P.stencil7(4, 4, 4, 6.0)
    with pytest.raises(cq.CliqgError) as e:
        cq.NaturalNestedDissection(4, 4, 5, ip, ind)      # box does not match N
    assert e.value.code == 1
    bad = ind.copy()
    bad[3] = 10 ** 6
    with pytest.raises(cq.CliqgError) as e:
        cq.NaturalNestedDissection(4, 4, 4, ip, bad)
    assert e.value.code == 1
    with pytest.raises(cq.CliqgError):
        cq.NaturalNestedDissection(4, 4, 4, ip, ind, cutoff=0)
    assert _lib.lib().cliqg_plan_solve(None, None, 1, 0) == 1


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "clique_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in txt and "liboracle" not in txt, fn


@pytest.mark.parametrize("dims,cut", [((6, 6, 6), 8), ((30, 30, 30), 128), ((12, 9, 1), 10),
                                      ((11, 3, 2), 1)])
def test_relative_indices_increase(dims, cut):
    # the two-pass extend-adds (factor and forward solve) split each child's
    # update at its first index past the parent's own columns: the relative
    # indices of every child must increase, the own-column ones first
    ip, ind, v, d = P.stencil7(*dims, 6.0)
    s = cq.NaturalNestedDissection(*d, ip, ind, cutoff=cut)
    for k in range(s.num_nodes):
        nd = s.node(k)
        for w in (4, 5):
            rel = s.node_array(k, w)
            assert np.all(np.diff(rel) > 0), (k, w)
            if rel.size:
                own = rel < nd["size"]
                # a prefix of own-column indices, then the update rows
                assert not np.any(own[1:] & ~own[:-1]), (k, w)
