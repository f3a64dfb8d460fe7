// Host-side symbolic analysis for the multifrontal LDL^T engine.
//
// Restates, bit-exactly, Clique's natural (geometric) nested dissection and
// its local symmetric analysis at world size 1:
//   NaturalNestedDissection   include/clique/symbolic/natural_nested_dissection.hpp:447-488
//   NaturalBisect             natural_nested_dissection.hpp:490-619
//   recursion                 natural_nested_dissection.hpp:60-189, 313-445
//   BuildChildrenFromPerm     include/clique/symbolic/nested_dissection.hpp:888-963
//   ReverseOrder              nested_dissection.hpp:817-886
//   BuildMap                  nested_dissection.hpp:1206-1365 (P=1: map[inds[t]] = off+t)
//   LocalSymmetricAnalysis    src/symbolic/LocalSymmetricAnalysis.cpp:15-114
//   Union / RelativeIndices   include/clique/core/environment/impl.hpp:95-139
// The outputs (tree, separator index order, front index sets and the
// extend-add relative indices) are integers, so "bit-exact" is literal.
#pragma once
#include <cstdint>
#include <vector>

namespace cliqg {

// Compressed adjacency of an undirected graph: the targets of source s are
// tgt[off[s] .. off[s+1]), sorted and de-duplicated. Targets >= n are
// "external" vertices (ancestors in the dissection), encoded relative to the
// subgraph's offset exactly as the reference does.
struct Graph {
    int n = 0;
    std::vector<int64_t> off;   // n+1
    std::vector<int> tgt;
};

// One node of the separator/elimination tree after analysis. Field names
// follow SymmNodeInfo (include/clique/symbolic/dist_symm_info.hpp:17-36) plus
// the separator index list of SepOrLeaf (dist_separator_tree.hpp).
struct SymNode {
    int size = 0, off = 0, myOff = 0;
    int parent = -1;
    std::vector<int> children;          // 0 or 2 entries: [left, right]
    bool onLeft = false;
    std::vector<int> inds;              // original indices of this node's vertices
    std::vector<int> origLowerStruct;   // directly connected ancestors
    std::vector<int> lowerStruct;       // full lower structure (nodal indices)
    std::vector<int> origLowerRelInds;
    std::vector<int> leftRelInds, rightRelInds;
};

struct Symbolic {
    int N = 0;
    std::vector<SymNode> nodes;     // left-first postorder, root last
    std::vector<int> map;           // map[original] = nodal index
    std::vector<int> inverseMap;    // inverseMap[nodal] = original
};

// Builds the adjacency used by the dissection from a CSR pattern of A
// (rows need not be sorted; duplicates are removed).
Graph graph_from_csr(int n, const int64_t* rowOff, const int* cols);

// Natural nested dissection of an nx*ny*nz box graph (vertex x+y*nx+z*nx*ny)
// followed by the local symmetric analysis. Throws std::invalid_argument on
// inconsistent input.
void natural_nested_dissection(int nx, int ny, int nz, const Graph& g,
                               int cutoff, Symbolic& out);

// Nested dissection of a general graph (the reference's NestedDissection,
// include/clique/symbolic/nested_dissection.hpp:491-533, with its recursion
// :115-238) with this library's own vertex bisection in place of METIS
// (symbolic.cpp graph_bisect): the same tree layout, numbering, map and
// analysis; the separators themselves differ from METIS's.
void graph_nested_dissection(const Graph& g, int cutoff, Symbolic& out);

// Symbolic analysis for a user-provided separator tree given in the
// reference's pre-reversal form is not needed at P=1; trees can instead be
// supplied already in postorder through this entry (used for batching many
// independent problems into one forest).
void analyze_postorder(Symbolic& sym);

}  // namespace cliqg
