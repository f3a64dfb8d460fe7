// Natural nested dissection + local symmetric analysis (host, integer).
// See symbolic.hpp for the reference functions each piece restates.
#include "symbolic.hpp"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

namespace cliqg {

namespace {

// Node of the separator/elimination tree in the order the recursion creates
// it (root first, right subtree before left subtree), before ReverseOrder.
struct PreNode {
    int size = 0, off = 0, parent = -1;
    std::vector<int> children;
    std::vector<int> lowerStruct;   // sorted, unique
    std::vector<int> inds;          // original indices
};

void sort_unique(std::vector<int>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

void finalize_rows(Graph& g) {
    for (int s = 0; s < g.n; ++s) {
        auto b = g.tgt.begin() + g.off[s];
        auto e = g.tgt.begin() + g.off[s + 1];
        std::sort(b, e);
    }
    // de-duplicate in place, rebuilding offsets
    int64_t w = 0;
    int64_t start = 0;
    for (int s = 0; s < g.n; ++s) {
        const int64_t e = g.off[s + 1];
        const int64_t rowStart = w;
        for (int64_t k = start; k < e; ++k) {
            if (w > rowStart && g.tgt[w - 1] == g.tgt[k]) continue;
            g.tgt[w++] = g.tgt[k];
        }
        start = e;
        g.off[s + 1] = w;
        g.off[s] = rowStart;
    }
    g.tgt.resize(w);
}

// Vertices of the subgraph whose edges leave it: the set {off + t : t >= n}
// over the listed sources (natural_nested_dissection.hpp:91-106, 144-160).
std::vector<int> connected_ancestors(const Graph& g, const int* sources,
                                     int numSources, int off) {
    std::vector<int> out;
    for (int i = 0; i < numSources; ++i) {
        const int s = sources ? sources[i] : i;
        for (int64_t k = g.off[s]; k < g.off[s + 1]; ++k) {
            const int t = g.tgt[k];
            if (t >= g.n) out.push_back(off + t);
        }
    }
    sort_unique(out);
    return out;
}

// Split the longest box dimension (ties: x, then y, then z); the left box
// gets (d-1)/2 planes, the separator plane is numbered last
// (natural_nested_dissection.hpp:513-614).
int natural_bisect(int nx, int ny, int nz, int& nxL, int& nyL, int& nzL,
                   int& nxR, int& nyR, int& nzR, std::vector<int>& perm) {
    const int n = nx * ny * nz;
    perm.assign(n, -1);
    int L, R, S;
    if (nx >= ny && nx >= nz) {
        nxL = (nx - 1) / 2; nyL = ny; nzL = nz;
        nxR = nx - 1 - nxL; nyR = ny; nzR = nz;
        L = nxL * nyL * nzL; R = nxR * nyR * nzR; S = ny * nz;
        int o = 0;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nxL; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = nxL + 1; x < nx; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L + R;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) perm[nxL + y * nx + z * nx * ny] = o++;
    } else if (ny >= nx && ny >= nz) {
        nxL = nx; nyL = (ny - 1) / 2; nzL = nz;
        nxR = nx; nyR = ny - 1 - nyL; nzR = nz;
        L = nxL * nyL * nzL; R = nxR * nyR * nzR; S = nx * nz;
        int o = 0;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < nyL; ++y)
                for (int x = 0; x < nx; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L;
        for (int z = 0; z < nz; ++z)
            for (int y = nyL + 1; y < ny; ++y)
                for (int x = 0; x < nx; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L + R;
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) perm[x + nyL * nx + z * nx * ny] = o++;
    } else {
        nxL = nx; nyL = ny; nzL = (nz - 1) / 2;
        nxR = nx; nyR = ny; nzR = nz - 1 - nzL;
        L = nxL * nyL * nzL; R = nxR * nyR * nzR; S = nx * ny;
        int o = 0;
        for (int z = 0; z < nzL; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L;
        for (int z = nzL + 1; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) perm[x + y * nx + z * nx * ny] = o++;
        o = L + R;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) perm[x + y * nx + nzL * nx * ny] = o++;
    }
    (void)L; (void)R;
    return S;
}

// Child subgraphs of a bisection: internal targets are renumbered through
// perm, external ones (>= n) are kept, and the right child is shifted down by
// the left child's size (nested_dissection.hpp:888-963).
void build_children(const Graph& g, const std::vector<int>& perm,
                    const std::vector<int>& inversePerm, int L, Graph& left,
                    int R, Graph& right) {
    const int n = g.n;
    auto build = [&](Graph& c, int size, int shift) {
        c.n = size;
        c.off.assign(size + 1, 0);
        int64_t count = 0;
        for (int s = 0; s < size; ++s) {
            const int src = inversePerm[s + shift];
            count += g.off[src + 1] - g.off[src];
        }
        c.tgt.resize(count);
        int64_t w = 0;
        for (int s = 0; s < size; ++s) {
            c.off[s] = w;
            const int src = inversePerm[s + shift];
            for (int64_t k = g.off[src]; k < g.off[src + 1]; ++k) {
                const int t = g.tgt[k];
                const int mapped = (t < n ? perm[t] : t);
                c.tgt[w++] = mapped - shift;
            }
        }
        c.off[size] = w;
        finalize_rows(c);
    };
    build(left, L, 0);
    build(right, R, L);
}

// Vertex bisection of a general graph (the role METIS plays in the
// reference's Bisect, nested_dissection.hpp:535-596; METIS itself is not
// used): a level structure rooted at a pseudo-peripheral vertex (George and
// Liu: repeated breadth-first searches from a vertex of least degree in the
// deepest level until the depth stops growing); the level that best balances
// the vertices before and after it is the separator; separator vertices
// without neighbours on one side move to the other side. Vertices the search
// does not reach (other components) join the smaller side. Only edges inside
// the subgraph (targets < n) count. When no component holds half of the
// vertices, whole components are split between the sides. Returns the
// separator size; perm puts
// the left part first, then the right part, then the separator, each in
// increasing vertex order, as the natural bisection does.
int graph_bisect(const Graph& g, int& L, int& R, std::vector<int>& perm) {
    const int n = g.n;
    std::vector<int> level(n, -1), order, comp(n, -1);
    order.reserve(n);
    // breadth-first search from root inside its component; returns the depth
    auto bfs = [&](int root) {
        for (int s : order) level[s] = -1;
        order.clear();
        level[root] = 0;
        order.push_back(root);
        for (size_t h = 0; h < order.size(); ++h) {
            const int s = order[h];
            for (int64_t k = g.off[s]; k < g.off[s + 1]; ++k) {
                const int t = g.tgt[k];
                if (t < n && t != s && level[t] < 0) {
                    level[t] = level[s] + 1;
                    order.push_back(t);
                }
            }
        }
        return level[order.back()];
    };
    auto degree = [&](int s) {
        int d = 0;
        for (int64_t k = g.off[s]; k < g.off[s + 1]; ++k) d += (g.tgt[k] < n && g.tgt[k] != s);
        return d;
    };
    // connected components; the largest one is bisected, the others join a side
    std::vector<int> compSize, compRoot;
    for (int s = 0; s < n; ++s) {
        if (comp[s] >= 0) continue;
        bfs(s);
        int best = s;
        for (int t : order) {
            comp[t] = (int)compSize.size();
            if (degree(t) < degree(best)) best = t;
        }
        compSize.push_back((int)order.size());
        compRoot.push_back(best);
    }
    int big = 0;
    for (int c = 1; c < (int)compSize.size(); ++c)
        if (compSize[c] > compSize[big]) big = c;
    // part: 0 left, 1 right, 2 separator, 3 undecided (other components)
    std::vector<char> part(n, 3);
    if (2 * compSize[big] >= n) {
        // pseudo-peripheral root: repeated searches from a vertex of least
        // degree in the deepest level while the depth grows
        int root = compRoot[big];
        int depth = bfs(root);
        for (int it = 0; it < 8; ++it) {
            int cand = -1;
            for (int s : order)
                if (level[s] == depth && (cand < 0 || degree(s) < degree(cand))) cand = s;
            const int d = bfs(cand);
            if (d <= depth) {
                depth = bfs(root);
                break;
            }
            root = cand;
            depth = d;
        }
        const int reached = (int)order.size();
        std::vector<int> count(depth + 1, 0);
        for (int s : order) ++count[level[s]];
        // separator level: the one minimising |before - after| (ties to the
        // smaller level), strictly inside the structure when it has one
        const int lo = depth >= 2 ? 1 : 0, hi = depth >= 2 ? depth - 1 : depth;
        int sepLevel = lo, before = 0, bestGap = -1;
        for (int l = 0; l <= depth; ++l) {
            if (l >= lo && l <= hi) {
                const int after = reached - before - count[l];
                const int gap = before > after ? before - after : after - before;
                if (bestGap < 0 || gap < bestGap) {
                    bestGap = gap;
                    sepLevel = l;
                }
            }
            before += count[l];
        }
        for (int s : order) part[s] = level[s] < sepLevel ? 0 : (level[s] > sepLevel ? 1 : 2);
        for (int s : order) {
            if (part[s] != 2) continue;
            bool toL = false, toR = false;
            for (int64_t k = g.off[s]; k < g.off[s + 1]; ++k) {
                const int t = g.tgt[k];
                if (t >= n || t == s) continue;
                toL = toL || part[t] == 0;
                toR = toR || part[t] == 1;
            }
            if (!toR && toL) part[s] = 0;
            else if (!toL && toR) part[s] = 1;
        }
    }
    // the other components, largest first, each to the smaller side
    int cnt[3] = {0, 0, 0};
    for (int s = 0; s < n; ++s)
        if (part[s] < 3) ++cnt[(int)part[s]];
    std::vector<int> rest;
    for (int c = 0; c < (int)compSize.size(); ++c)
        if (!(2 * compSize[big] >= n && c == big)) rest.push_back(c);
    std::stable_sort(rest.begin(), rest.end(),
                     [&](int a, int b) { return compSize[a] > compSize[b]; });
    std::vector<char> side(compSize.size(), 0);
    for (int c : rest) {
        side[c] = cnt[0] <= cnt[1] ? 0 : 1;
        cnt[(int)side[c]] += compSize[c];
    }
    for (int s = 0; s < n; ++s)
        if (part[s] == 3) part[s] = side[comp[s]];
    L = R = 0;
    int S = 0;
    for (int s = 0; s < n; ++s) (part[s] == 0 ? L : part[s] == 1 ? R : S) += 1;
    if (S == 0) {
        // a separator of at least one vertex keeps every separator node
        // non-empty: the last vertex of the larger side
        const char bigSide = L >= R ? 0 : 1;
        for (int s = n - 1; s >= 0; --s)
            if (part[s] == bigSide) {
                part[s] = 2;
                break;
            }
        (bigSide == 0 ? L : R) -= 1;
        S = 1;
    }
    perm.assign(n, -1);
    int o[3] = {0, L, L + R};
    for (int s = 0; s < n; ++s) perm[s] = o[(int)part[s]]++;
    return S;
}

void recurse_graph(const Graph& g, const std::vector<int>& perm, std::vector<PreNode>& list,
                   int parent, int off, int cutoff) {
    const int n = g.n;
    if (n <= cutoff) {
        PreNode node;
        node.size = n;
        node.off = off;
        node.parent = parent;
        node.inds = perm;
        node.lowerStruct = connected_ancestors(g, nullptr, n, off);
        list.push_back(std::move(node));
        return;
    }
    int L, R;
    std::vector<int> map;
    const int sepSize = graph_bisect(g, L, R, map);
    std::vector<int> inverseMap(n);
    for (int s = 0; s < n; ++s) inverseMap[map[s]] = s;
    const int me = (int)list.size();
    {
        PreNode node;
        node.size = sepSize;
        node.off = off + (n - sepSize);
        node.parent = parent;
        node.children.assign(2, -1);
        std::vector<int> sepLocal(sepSize);
        for (int s = 0; s < sepSize; ++s) sepLocal[s] = inverseMap[s + (n - sepSize)];
        node.lowerStruct = connected_ancestors(g, sepLocal.data(), sepSize, off);
        node.inds.resize(sepSize);
        for (int s = 0; s < sepSize; ++s) node.inds[s] = perm[sepLocal[s]];
        list.push_back(std::move(node));
    }
    Graph left, right;
    build_children(g, map, inverseMap, L, left, R, right);
    std::vector<int> leftPerm(L), rightPerm(R);
    for (int s = 0; s < L; ++s) leftPerm[s] = perm[inverseMap[s]];
    for (int s = 0; s < R; ++s) rightPerm[s] = perm[inverseMap[s + L]];
    std::vector<int>().swap(inverseMap);
    std::vector<int>().swap(map);
    list[me].children[1] = (int)list.size();
    {
        Graph r = std::move(right);
        recurse_graph(r, rightPerm, list, me, off + L, cutoff);
    }
    std::vector<int>().swap(rightPerm);
    list[me].children[0] = (int)list.size();
    recurse_graph(left, leftPerm, list, me, off, cutoff);
}

void recurse(int nx, int ny, int nz, const Graph& g, const std::vector<int>& perm,
             std::vector<PreNode>& list, int parent, int off, int cutoff) {
    const int n = g.n;
    if (n <= cutoff) {
        PreNode node;
        node.size = n;
        node.off = off;
        node.parent = parent;
        node.inds = perm;
        node.lowerStruct = connected_ancestors(g, nullptr, n, off);
        list.push_back(std::move(node));
        return;
    }
    if (nx * ny * nz != n)
        throw std::invalid_argument("natural nested dissection: box " +
                                    std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                                    std::to_string(nz) + " does not match a subgraph of " +
                                    std::to_string(n) + " vertices");
    int nxL, nyL, nzL, nxR, nyR, nzR;
    std::vector<int> map;
    const int sepSize = natural_bisect(nx, ny, nz, nxL, nyL, nzL, nxR, nyR, nzR, map);
    std::vector<int> inverseMap(n);
    for (int s = 0; s < n; ++s) inverseMap[map[s]] = s;
    const int L = nxL * nyL * nzL;
    const int R = nxR * nyR * nzR;

    const int me = (int)list.size();
    {
        PreNode node;
        node.size = sepSize;
        node.off = off + (n - sepSize);
        node.parent = parent;
        node.children.assign(2, -1);
        std::vector<int> sepLocal(sepSize);
        for (int s = 0; s < sepSize; ++s) sepLocal[s] = inverseMap[s + (n - sepSize)];
        node.lowerStruct = connected_ancestors(g, sepLocal.data(), sepSize, off);
        node.inds.resize(sepSize);
        for (int s = 0; s < sepSize; ++s) node.inds[s] = perm[sepLocal[s]];
        list.push_back(std::move(node));
    }
    Graph left, right;
    build_children(g, map, inverseMap, L, left, R, right);
    std::vector<int> leftPerm(L), rightPerm(R);
    for (int s = 0; s < L; ++s) leftPerm[s] = perm[inverseMap[s]];
    for (int s = 0; s < R; ++s) rightPerm[s] = perm[inverseMap[s + L]];
    std::vector<int>().swap(inverseMap);
    std::vector<int>().swap(map);

    // right first so that, after reversal, the left child is ordered first
    list[me].children[1] = (int)list.size();
    {
        Graph r = std::move(right);
        recurse(nxR, nyR, nzR, r, rightPerm, list, me, off + L, cutoff);
    }
    std::vector<int>().swap(rightPerm);
    list[me].children[0] = (int)list.size();
    recurse(nxL, nyL, nzL, left, leftPerm, list, me, off, cutoff);
}

std::vector<int> union_sorted(const std::vector<int>& a, const std::vector<int>& b) {
    std::vector<int> out(a.size() + b.size());
    auto it = std::set_union(a.begin(), a.end(), b.begin(), b.end(), out.begin());
    out.resize(it - out.begin());
    return out;
}

std::vector<int> relative_indices(const std::vector<int>& sub, const std::vector<int>& full) {
    std::vector<int> rel(sub.size());
    auto it = full.begin();
    for (size_t i = 0; i < sub.size(); ++i) {
        it = std::lower_bound(it, full.end(), sub[i]);
        if (it == full.end() || *it != sub[i])
            throw std::logic_error("relative index not found in full structure");
        rel[i] = (int)(it - full.begin());
    }
    return rel;
}

}  // namespace

Graph graph_from_csr(int n, const int64_t* rowOff, const int* cols) {
    Graph g;
    g.n = n;
    g.off.assign(rowOff, rowOff + n + 1);
    const int64_t nnz = rowOff[n] - rowOff[0];
    g.tgt.assign(cols + rowOff[0], cols + rowOff[0] + nnz);
    if (rowOff[0] != 0)
        for (auto& o : g.off) o -= rowOff[0];
    for (int64_t k = 0; k < nnz; ++k)
        if (g.tgt[k] < 0 || g.tgt[k] >= n)
            throw std::invalid_argument("graph target out of range");
    finalize_rows(g);
    return g;
}

void analyze_postorder(Symbolic& sym) {
    const int num = (int)sym.nodes.size();
    int myOff = 0;
    for (int s = 0; s < num; ++s) {
        SymNode& node = sym.nodes[s];
        node.myOff = myOff;
        if (node.children.size() == 2) {
            SymNode& l = sym.nodes[node.children[0]];
            SymNode& r = sym.nodes[node.children[1]];
            l.onLeft = true;
            r.onLeft = false;
            const auto childrenStruct = union_sorted(l.lowerStruct, r.lowerStruct);
            const auto partial = union_sorted(node.origLowerStruct, childrenStruct);
            std::vector<int> nodeInds(node.size);
            std::iota(nodeInds.begin(), nodeInds.end(), node.off);
            const auto full = union_sorted(partial, nodeInds);
            node.origLowerRelInds = relative_indices(node.origLowerStruct, full);
            node.leftRelInds = relative_indices(l.lowerStruct, full);
            node.rightRelInds = relative_indices(r.lowerStruct, full);
            node.lowerStruct.assign(full.begin() + node.size, full.end());
        } else if (node.children.empty()) {
            node.lowerStruct = node.origLowerStruct;
            node.origLowerRelInds.resize(node.origLowerStruct.size());
            for (size_t i = 0; i < node.origLowerStruct.size(); ++i)
                node.origLowerRelInds[i] = (int)i + node.size;
        } else {
            throw std::logic_error("tree must be built from bisections");
        }
        myOff += node.size;
    }
}

namespace {
void finish_dissection(std::vector<PreNode>& pre, const Graph& g, Symbolic& out);
}

void natural_nested_dissection(int nx, int ny, int nz, const Graph& g, int cutoff,
                               Symbolic& out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || (int64_t)nx * ny * nz != g.n)
        throw std::invalid_argument("grid dimensions do not match the graph size");
    if (cutoff < 1) throw std::invalid_argument("cutoff must be positive");
    std::vector<PreNode> pre;
    std::vector<int> perm(g.n);
    std::iota(perm.begin(), perm.end(), 0);
    recurse(nx, ny, nz, g, perm, pre, -1, 0, cutoff);
    finish_dissection(pre, g, out);
}

void graph_nested_dissection(const Graph& g, int cutoff, Symbolic& out) {
    if (g.n <= 0) throw std::invalid_argument("empty graph");
    if (cutoff < 1) throw std::invalid_argument("cutoff must be positive");
    std::vector<PreNode> pre;
    std::vector<int> perm(g.n);
    std::iota(perm.begin(), perm.end(), 0);
    recurse_graph(g, perm, pre, -1, 0, cutoff);
    finish_dissection(pre, g, out);
}

namespace {
// ReverseOrder, BuildMap and the symbolic analysis of a dissection's tree
void finish_dissection(std::vector<PreNode>& pre, const Graph& g, Symbolic& out) {
    // ReverseOrder: node i of the creation order becomes node last-i.
    const int num = (int)pre.size();
    const int last = num - 1;
    out.N = g.n;
    out.nodes.assign(num, SymNode());
    for (int i = 0; i < num; ++i) {
        PreNode& p = pre[i];
        SymNode& s = out.nodes[last - i];
        s.size = p.size;
        s.off = p.off;
        s.parent = (p.parent < 0 ? -1 : last - p.parent);
        s.children.clear();
        for (int c : p.children) s.children.push_back(last - c);
        s.origLowerStruct = std::move(p.lowerStruct);
        s.inds = std::move(p.inds);
    }
    std::vector<PreNode>().swap(pre);

    // BuildMap at P=1
    out.map.assign(g.n, -1);
    for (const SymNode& s : out.nodes)
        for (int t = 0; t < s.size; ++t) out.map[s.inds[t]] = s.off + t;
    out.inverseMap.assign(g.n, -1);
    for (int i = 0; i < g.n; ++i) {
        if (out.map[i] < 0) throw std::logic_error("nested dissection map is not a permutation");
        out.inverseMap[out.map[i]] = i;
    }
    analyze_postorder(out);
}
}  // namespace

}  // namespace cliqg
