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


# ---------------------------------------------------------------------------
# nested dissection of general graphs (NestedDissection with the library's
# own bisection)
# ---------------------------------------------------------------------------
def irregular_graph(seed, n=600, deg=6):
    """A random geometric graph in the unit square (k nearest neighbours,
    symmetrised) plus a second, disconnected component: irregular degrees,
    no grid structure."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    pts = rng.uniform(size=(n, 2))
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    nbr = np.argsort(d2, axis=1)[:, 1:deg + 1]
    rows = np.repeat(np.arange(n), deg)
    A = sp.csr_matrix((np.ones(rows.size), (rows, nbr.ravel())), shape=(n, n))
    A = A + A.T
    ip2, ind2, _, _ = P.stencil7(7, 6, 1, 4.0)
    B = sp.csr_matrix((np.ones(ind2.size), ind2, ip2))
    M = (sp.block_diag([A, B]) + sp.identity(n + B.shape[0])).tocsr()
    M.has_sorted_indices = False
    M.sort_indices()
    return M.indptr.astype(np.int64), M.indices.astype(np.int32)


def dissection_checks(ip, ind, cut):
    s = cq.NestedDissection(ip, ind, cutoff=cut)
    N = ip.size - 1
    mp = s.map()
    assert np.array_equal(np.sort(mp), np.arange(N))
    nodes = [s.node(k) for k in range(s.num_nodes)]
    assert nodes[-1]["parent"] == -1
    # subtree nodal ranges: a node's subtree owns [first, off + size)
    first = {}
    for k, nd in enumerate(nodes):   # postorder: children come first
        kids = [c for c in (nd["child0"], nd["child1"]) if c >= 0]
        assert len(kids) in (0, 2)
        first[k] = min([first[c] for c in kids] + [nd["off"]])
        if not kids:
            assert 1 <= nd["size"] <= cut
        else:
            assert nd["size"] >= 1
            assert nodes[kids[0]]["parent"] == k and nodes[kids[1]]["parent"] == k
    # separator property: no edge joins the two children's subtrees
    rows = np.repeat(np.arange(N), np.diff(ip))
    a, b = mp[rows], mp[ind]
    for k, nd in enumerate(nodes):
        if nd["child0"] < 0:
            continue
        l, r = nd["child0"], nd["child1"]
        inL = lambda x: (x >= first[l]) & (x < nodes[l]["off"] + nodes[l]["size"])   # noqa: E731
        inR = lambda x: (x >= first[r]) & (x < nodes[r]["off"] + nodes[r]["size"])   # noqa: E731
        assert not np.any(inL(a) & inR(b))
    return s


@pytest.mark.parametrize("seed,cut", [(0, 16), (1, 40), (2, 128)])
def test_graph_dissection_separates_an_irregular_graph(seed, cut):
    ip, ind = irregular_graph(seed)
    dissection_checks(ip, ind, cut)


@pytest.mark.parametrize("dims,cut", [((6, 5, 4), 6), ((9, 8, 1), 5), ((30, 1, 1), 3)])
def test_graph_dissection_lower_struct_holds_symbolic_cholesky(dims, cut):
    """The analysis of a dissection built by the graph bisection holds the
    column structure of the Cholesky factor of P A P^T (brute force): every
    fill entry below a node lies in its lowerStruct. Equality is not implied:
    the supernodal rule (lowerStruct = children's structures and the node's
    own connections, less the node, LocalSymmetricAnalysis.cpp:15-114)
    treats a separator as one dense block, and a separator of a general
    graph need not be connected (METIS's need not either)."""
    from test_oracle import symbolic_cholesky
    ip, ind, v, d = P.stencil7(*dims, 6.0)
    s = dissection_checks(ip, ind, cut)
    A = P.to_scipy(ip, ind, v).toarray()
    mp = s.map()
    Ap = np.zeros_like(A)
    Ap[np.ix_(mp, mp)] = A
    Lpat = symbolic_cholesky(Ap)
    for k in range(s.num_nodes):
        nd = s.node(k)
        cols = range(nd["off"], nd["off"] + nd["size"])
        below = set()
        for c in cols:
            below |= set(np.nonzero(Lpat[:, c])[0].tolist())
        below -= set(cols)
        assert below <= set(s.node_array(k, 1).tolist())


def test_graph_dissection_fill_is_close_to_natural_on_grids():
    """On grids the graph bisection's factor is within a small factor of the
    natural (geometric) dissection's in entries and flops."""
    for dims in ((24, 24, 24), (120, 90, 1)):
        ip, ind, v, d = P.stencil7(*dims, 6.0)
        g = cq.NestedDissection(ip, ind, cutoff=64)
        n = cq.NaturalNestedDissection(*d, ip, ind, cutoff=64)
        fg, _, eg = g.work(False)
        fn, _, en = n.work(False)
        assert eg <= 1.6 * en and fg <= 2.5 * fn, (dims, eg / en, fg / fn)
