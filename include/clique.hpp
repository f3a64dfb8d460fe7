/*
 * clique.hpp -- header-only C++ drop-in surface of the MI355X multifrontal
 * LDL^T engine, namespace cliq, at world size 1.
 *
 * Mirrors the types and calls of Clique's own umbrella header
 * (include/clique.hpp:19-90 of poulson/Clique) for the hot path:
 *   DistGraph, DistSparseMatrix, DistMultiVec, DistMap, DistSeparatorTree,
 *   DistSymmInfo, NaturalNestedDissection, DistSymmFrontTree, LDL,
 *   DistNodalMultiVec (Pull/Push), Solve, SymmetricSolve, Multiply,
 *   FactorizationWork / SolveWork / MemoryInfo.
 * Every numeric call goes through the C ABI of libcliqg.so (cliqg.h): the
 * symbolic analysis runs on the host inside the library, the fronts, the
 * factorization and the solves live on the GPU. Containers keep host data
 * as the reference's do; device copies are made at the boundary.
 *
 * Errors keep the reference's exception types: LogicError for API misuse
 * (El::LogicError), RuntimeError otherwise; singular pivots raise
 * SingularMatrixException as Elemental's checkIfSingular path does.
 */
#pragma once
#ifndef CLIQ_MI355X_CLIQUE_HPP
#define CLIQ_MI355X_CLIQUE_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cliqg.h"

namespace cliq {

using Int = int;
template <typename T> using Complex = std::complex<T>;
template <typename T> struct Base { typedef T type; };
template <typename T> struct Base<std::complex<T>> { typedef T type; };
template <typename T> struct IsComplex { static const bool val = false; };
template <typename T> struct IsComplex<std::complex<T>> { static const bool val = true; };

struct LogicError : std::logic_error {
    explicit LogicError(const std::string& m) : std::logic_error(m) {}
};
struct RuntimeError : std::runtime_error {
    explicit RuntimeError(const std::string& m) : std::runtime_error(m) {}
};
struct SingularMatrixException : std::runtime_error {
    explicit SingularMatrixException(const std::string& m) : std::runtime_error(m) {}
};

namespace detail {

inline void check(cliqg_status st, const char* what) {
    if (st == CLIQG_OK) return;
    const std::string m = std::string(what) + ": " + cliqg_last_error();
    if (st == CLIQG_ERR_SINGULAR) throw SingularMatrixException(m);
    if (st == CLIQG_ERR_INVALID || st == CLIQG_ERR_STATE) throw LogicError(m);
    throw RuntimeError(m);
}

// Owning device allocation through the library's allocator.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { resize(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void resize(size_t bytes) {
        if (bytes == n_) return;
        release();
        if (bytes) check(cliqg_device_malloc(&p_, bytes), "device allocation");
        n_ = bytes;
    }
    void upload(const void* h, size_t bytes) {
        resize(bytes);
        if (bytes) check(cliqg_memcpy(p_, h, bytes, 0), "host to device copy");
    }
    void download(void* h, size_t bytes) const {
        if (bytes) check(cliqg_memcpy(h, p_, bytes, 1), "device to host copy");
    }
    void* get() const { return p_; }
    size_t size() const { return n_; }

private:
    void release() {
        if (p_) (void)cliqg_device_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void* p_ = nullptr;
    size_t n_ = 0;
};

template <typename F> struct Field;
template <> struct Field<double> { static const cliqg_field value = CLIQG_REAL; };
template <> struct Field<std::complex<double>> { static const cliqg_field value = CLIQG_COMPLEX; };

struct SymbolicDeleter {
    void operator()(cliqg_symbolic_s* s) const { if (s) (void)cliqg_symbolic_destroy(s); }
};
struct PlanDeleter {
    void operator()(cliqg_plan_s* p) const { if (p) (void)cliqg_plan_destroy(p); }
};

}  // namespace detail

// ---------------------------------------------------------------------------
// environment (src/core/global.cpp:75-110; a single process)
// ---------------------------------------------------------------------------
namespace mpi {
struct Comm {
    int id = 0;
};
const Comm COMM_WORLD{};
inline int Rank(Comm = COMM_WORLD) { return 0; }
inline int Size(Comm = COMM_WORLD) { return 1; }
inline void Barrier(Comm = COMM_WORLD) { detail::check(cliqg_device_synchronize(), "Barrier"); }
inline double Time() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace mpi

inline int& InitCounter() {
    static int n = 0;
    return n;
}
inline void Initialize(int&, char**&) { ++InitCounter(); }
inline void Finalize() {
    if (InitCounter() > 0) --InitCounter();
}
inline bool Initialized() { return InitCounter() > 0; }
inline int DistributedDepth(mpi::Comm = mpi::COMM_WORLD) { return 0; }

// ---------------------------------------------------------------------------
// minimal column-major matrix (the El::Matrix uses made by the hot path)
// ---------------------------------------------------------------------------
template <typename T>
class Matrix {
public:
    Matrix() = default;
    Matrix(int m, int n) { Resize(m, n); }
    void Resize(int m, int n) {
        m_ = m;
        n_ = n;
        ld_ = std::max(m, 1);
        a_.assign((size_t)ld_ * n, T(0));
    }
    void Empty() { Resize(0, 0); }
    int Height() const { return m_; }
    int Width() const { return n_; }
    int LDim() const { return ld_; }
    T Get(int i, int j) const { return a_[i + (size_t)j * ld_]; }
    void Set(int i, int j, T v) { a_[i + (size_t)j * ld_] = v; }
    void Update(int i, int j, T v) { a_[i + (size_t)j * ld_] += v; }
    T* Buffer() { return a_.data(); }
    const T* LockedBuffer() const { return a_.data(); }
    double MemorySize() const { return (double)a_.size(); }

private:
    int m_ = 0, n_ = 0, ld_ = 1;
    std::vector<T> a_;
};

// ---------------------------------------------------------------------------
// graphs and sparse matrices (core/dist_graph, core/dist_sparse_matrix)
// ---------------------------------------------------------------------------
class DistGraph {
public:
    explicit DistGraph(mpi::Comm comm = mpi::COMM_WORLD) : comm_(comm) {}
    explicit DistGraph(int numSources, mpi::Comm comm = mpi::COMM_WORLD)
        : numSources_(numSources), numTargets_(numSources), comm_(comm) {
        offsets_.assign(numSources + 1, 0);
    }
    int NumSources() const { return numSources_; }
    int NumTargets() const { return numTargets_; }
    int FirstLocalSource() const { return 0; }
    int NumLocalSources() const { return numSources_; }
    int Blocksize() const { return numSources_; }
    mpi::Comm Comm() const { return comm_; }
    void ResizeTo(int numSources) {
        numSources_ = numTargets_ = numSources;
        sources_.clear();
        targets_.clear();
        offsets_.assign(numSources + 1, 0);
    }
    void StartAssembly() {
        if (assembling_) throw LogicError("Already assembling");
        assembling_ = true;
    }
    void Reserve(int n) {
        sources_.reserve(n);
        targets_.reserve(n);
    }
    void Insert(int row, int col) {
        if (!assembling_) throw LogicError("Insert outside of assembly");
        if (row < 0 || row >= numSources_ || col < 0 || col >= numTargets_)
            throw LogicError("Index out of bounds");
        sources_.push_back(row);
        targets_.push_back(col);
    }
    void StopAssembly() {
        if (!assembling_) throw LogicError("Cannot stop assembly without starting");
        assembling_ = false;
        std::vector<std::pair<int, int>> e(sources_.size());
        for (size_t k = 0; k < e.size(); ++k) e[k] = {sources_[k], targets_[k]};
        std::sort(e.begin(), e.end());
        e.erase(std::unique(e.begin(), e.end()), e.end());
        sources_.resize(e.size());
        targets_.resize(e.size());
        for (size_t k = 0; k < e.size(); ++k) {
            sources_[k] = e[k].first;
            targets_[k] = e[k].second;
        }
        ComputeOffsets();
    }
    int NumLocalEdges() const { return (int)sources_.size(); }
    int Source(int e) const { return sources_[e]; }
    int Target(int e) const { return targets_[e]; }
    int LocalEdgeOffset(int s) const { return (int)offsets_[s]; }
    int EdgeOffset(int s) const { return (int)offsets_[s]; }
    int NumConnections(int s) const { return (int)(offsets_[s + 1] - offsets_[s]); }
    const std::vector<int64_t>& Offsets() const { return offsets_; }
    const std::vector<int>& Targets() const { return targets_; }

    // used by DistSparseMatrix, whose entries are sorted and de-duplicated
    // alongside the values
    void SetSorted(std::vector<int>&& s, std::vector<int>&& t) {
        sources_ = std::move(s);
        targets_ = std::move(t);
        assembling_ = false;
        ComputeOffsets();
    }

private:
    void ComputeOffsets() {
        offsets_.assign(numSources_ + 1, 0);
        for (int s : sources_) ++offsets_[s + 1];
        for (int i = 0; i < numSources_; ++i) offsets_[i + 1] += offsets_[i];
    }
    int numSources_ = 0, numTargets_ = 0;
    mpi::Comm comm_;
    bool assembling_ = false;
    std::vector<int> sources_, targets_;
    std::vector<int64_t> offsets_{0};
};

template <typename T>
class DistSparseMatrix {
public:
    explicit DistSparseMatrix(mpi::Comm comm = mpi::COMM_WORLD) : graph_(comm) {}
    explicit DistSparseMatrix(int N, mpi::Comm comm = mpi::COMM_WORLD) : graph_(N, comm) {}
    int Height() const { return graph_.NumSources(); }
    int Width() const { return graph_.NumTargets(); }
    int LocalHeight() const { return graph_.NumSources(); }
    int FirstLocalRow() const { return 0; }
    int Blocksize() const { return graph_.NumSources(); }
    int NumLocalEntries() const { return (int)vals_.size(); }
    mpi::Comm Comm() const { return graph_.Comm(); }
    void ResizeTo(int N) {
        graph_.ResizeTo(N);
        vals_.clear();
        pending_.clear();
        ++version_;
    }
    void StartAssembly() {
        if (assembling_) throw LogicError("Already assembling");
        assembling_ = true;
    }
    void Reserve(int n) { pending_.reserve(n); }
    // Update appends; StopAssembly sorts by (i, j) and sums duplicates
    // (include/clique/core/dist_sparse_matrix/impl.hpp:209-257)
    void Update(int row, int col, T value) {
        if (!assembling_) throw LogicError("Update outside of assembly");
        if (row < 0 || row >= Height() || col < 0 || col >= Width())
            throw LogicError("Index out of bounds");
        pending_.push_back({row, col, value});
    }
    void StopAssembly() {
        if (!assembling_) throw LogicError("Cannot stop assembly without starting");
        assembling_ = false;
        for (int e = 0; e < NumLocalEntries(); ++e)
            pending_.push_back({graph_.Source(e), graph_.Target(e), vals_[e]});
        std::stable_sort(pending_.begin(), pending_.end(), [](const Entry& a, const Entry& b) {
            return a.i < b.i || (a.i == b.i && a.j < b.j);
        });
        std::vector<int> s, t;
        vals_.clear();
        for (const Entry& e : pending_) {
            if (!s.empty() && s.back() == e.i && t.back() == e.j) {
                vals_.back() += e.v;
            } else {
                s.push_back(e.i);
                t.push_back(e.j);
                vals_.push_back(e.v);
            }
        }
        pending_.clear();
        graph_.SetSorted(std::move(s), std::move(t));
        ++version_;
    }
    int Row(int e) const { return graph_.Source(e); }
    int Col(int e) const { return graph_.Target(e); }
    T Value(int e) const { return vals_[e]; }
    void SetValue(int e, T v) { vals_[e] = v; ++version_; }
    int LocalEntryOffset(int i) const { return graph_.LocalEdgeOffset(i); }
    int NumConnections(int i) const { return graph_.NumConnections(i); }
    const cliq::DistGraph& DistGraph() const { return graph_; }
    const cliq::DistGraph& LockedDistGraph() const { return graph_; }
    const std::vector<T>& Values() const { return vals_; }
    unsigned Version() const { return version_; }
    // CSR arrays in device memory, uploaded once per assembly or value change
    const int64_t* DeviceRowOffsets() const { sync(); return (const int64_t*)dOff_.get(); }
    const int* DeviceCols() const { sync(); return (const int*)dCols_.get(); }
    const void* DeviceValues() const { sync(); return dVals_.get(); }

private:
    struct Entry {
        int i, j;
        T v;
    };
    void sync() const {
        if (devVersion_ == version_ + 1) return;
        const cliq::DistGraph& g = graph_;
        dOff_.upload(g.Offsets().data(), g.Offsets().size() * sizeof(int64_t));
        dCols_.upload(g.Targets().data(), g.Targets().size() * sizeof(int));
        dVals_.upload(vals_.data(), vals_.size() * sizeof(T));
        devVersion_ = version_ + 1;
    }
    cliq::DistGraph graph_;
    std::vector<T> vals_;
    std::vector<Entry> pending_;
    bool assembling_ = false;
    unsigned version_ = 0;
    mutable unsigned devVersion_ = 0;   // version_ + 1 when the device copy is current
    mutable detail::DeviceBuffer dOff_, dCols_, dVals_;
};

// ---------------------------------------------------------------------------
// multi-vectors (core/dist_multi_vec): column-major N x k
// ---------------------------------------------------------------------------
template <typename T>
class DistMultiVec {
public:
    explicit DistMultiVec(mpi::Comm comm = mpi::COMM_WORLD) : comm_(comm) {}
    DistMultiVec(int height, int width, mpi::Comm comm = mpi::COMM_WORLD) : comm_(comm) {
        ResizeTo(height, width);
    }
    DistMultiVec(const DistMultiVec& o) : comm_(o.comm_), h_(o.h_), w_(o.w_) {
        o.host();
        a_ = o.a_;
        where_ = HOST;
    }
    DistMultiVec& operator=(const DistMultiVec& o) {
        if (this != &o) {
            o.host();
            comm_ = o.comm_;
            h_ = o.h_;
            w_ = o.w_;
            a_ = o.a_;
            where_ = HOST;
        }
        return *this;
    }
    void ResizeTo(int height, int width) {
        h_ = height;
        w_ = width;
        a_.assign((size_t)height * width, T(0));
        where_ = HOST;
    }
    int Height() const { return h_; }
    int Width() const { return w_; }
    int LocalHeight() const { return h_; }
    int FirstLocalRow() const { return 0; }
    int LDim() const { return std::max(h_, 1); }
    mpi::Comm Comm() const { return comm_; }
    // Host access: the values come back from the device copy when that is
    // the newer one (after a device-side Push, Multiply or Refine).
    T GetLocal(int i, int j) const { host(); return a_[i + (size_t)j * h_]; }
    void SetLocal(int i, int j, T v) { hostMut(); a_[i + (size_t)j * h_] = v; }
    void UpdateLocal(int i, int j, T v) { hostMut(); a_[i + (size_t)j * h_] += v; }
    T* Buffer() { hostMut(); return a_.data(); }
    const T* LockedBuffer() const { host(); return a_.data(); }
    // Device access (HBM, column-major, ld LDim()): the host values go up
    // once when they are the newer ones; DeviceBuffer() marks the device
    // copy as the newer one.
    const void* LockedDeviceBuffer() const { device(); return d_.get(); }
    void* DeviceBuffer() { device(); where_ = DEVICE; return d_.get(); }
    bool DeviceIsCurrent() const { return where_ != HOST; }

private:
    enum Where { BOTH, HOST, DEVICE };   // which copy holds the newest values
    size_t bytes() const { return (size_t)h_ * w_ * sizeof(T); }
    void host() const {
        if (where_ == DEVICE) {
            d_.download(a_.data(), bytes());
            where_ = BOTH;
        }
    }
    void hostMut() { host(); where_ = HOST; }
    void device() const {
        if (d_.size() != bytes()) {
            d_.resize(bytes());
            if (where_ != HOST && bytes()) where_ = HOST;   // fresh device memory
        }
        if (where_ == HOST) {
            if (bytes()) d_.upload(a_.data(), bytes());
            where_ = BOTH;
        }
    }
    mpi::Comm comm_;
    int h_ = 0, w_ = 0;
    mutable std::vector<T> a_;
    mutable detail::DeviceBuffer d_;
    mutable Where where_ = HOST;
};

template <typename T> void Zero(DistMultiVec<T>& X) { X.ResizeTo(X.Height(), X.Width()); }

namespace detail {
inline std::mt19937_64& rng() {
    static std::mt19937_64 g(0);
    return g;
}
template <typename T> T sample_ball(double radius);
template <> inline double sample_ball<double>(double radius) {
    std::uniform_real_distribution<double> u(-radius, radius);
    return u(rng());
}
template <> inline std::complex<double> sample_ball<std::complex<double>>(double radius) {
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double r = radius * std::sqrt(u(rng())), th = 2 * M_PI * u(rng());
    return {r * std::cos(th), r * std::sin(th)};
}
}  // namespace detail

// uniform samples from the ball of the given radius (El::SampleBall)
template <typename T> void MakeUniform(DistMultiVec<T>& X, double radius = 1.0) {
    for (int j = 0; j < X.Width(); ++j)
        for (int i = 0; i < X.Height(); ++i) X.SetLocal(i, j, detail::sample_ball<T>(radius));
}

// column two-norms; on the device when the device copy is current
template <typename T>
void Norms(const DistMultiVec<T>& X, std::vector<typename Base<T>::type>& norms) {
    norms.assign(X.Width(), 0);
    if (X.DeviceIsCurrent() && X.Height() > 0) {
        std::vector<double> d(X.Width());
        detail::check(cliqg_col_norms(X.Height(), X.Width(), X.LockedDeviceBuffer(), X.LDim(),
                                      detail::Field<T>::value, d.data(), nullptr),
                      "Norms");
        for (int j = 0; j < X.Width(); ++j) norms[j] = d[j];
        return;
    }
    for (int j = 0; j < X.Width(); ++j) {
        double s = 0;
        for (int i = 0; i < X.Height(); ++i) s += std::norm(X.GetLocal(i, j));
        norms[j] = std::sqrt(s);
    }
}

template <typename T> void Axpy(T alpha, const DistMultiVec<T>& X, DistMultiVec<T>& Y) {
    if (X.Height() != Y.Height() || X.Width() != Y.Width())
        throw LogicError("X and Y must be the same size");
    for (int j = 0; j < X.Width(); ++j)
        for (int i = 0; i < X.Height(); ++i) Y.UpdateLocal(i, j, alpha * X.GetLocal(i, j));
}

// ---------------------------------------------------------------------------
// maps (core/dist_map): a permutation at world size 1
// ---------------------------------------------------------------------------
class DistMap {
public:
    explicit DistMap(mpi::Comm comm = mpi::COMM_WORLD) : comm_(comm) {}
    explicit DistMap(int numSources, mpi::Comm comm = mpi::COMM_WORLD)
        : comm_(comm), map_(numSources, 0) {}
    int NumSources() const { return (int)map_.size(); }
    int NumLocalSources() const { return (int)map_.size(); }
    int FirstLocalSource() const { return 0; }
    int Blocksize() const { return (int)map_.size(); }
    mpi::Comm Comm() const { return comm_; }
    void Resize(int n) { map_.assign(n, 0); stale_ = true; }
    int GetLocal(int i) const { return map_[i]; }
    void SetLocal(int i, int v) { map_[i] = v; stale_ = true; }
    std::vector<int>& Map() { stale_ = true; return map_; }
    const std::vector<int>& Map() const { return map_; }
    // the map in device memory (uploaded again after any change)
    const int* DeviceMap() const {
        if (stale_ || dev_.size() != map_.size() * sizeof(int)) {
            dev_.upload(map_.data(), map_.size() * sizeof(int));
            stale_ = false;
        }
        return static_cast<const int*>(dev_.get());
    }
    void Translate(std::vector<int>& inds) const {
        for (int& i : inds) i = map_[i];
    }
    void FormInverse(DistMap& inv) const {
        inv.comm_ = comm_;
        inv.map_.assign(map_.size(), 0);
        for (size_t i = 0; i < map_.size(); ++i) inv.map_[map_[i]] = (int)i;
        inv.stale_ = true;
    }
    // firstMap := this o firstMap (dist_map/impl.hpp Extend)
    void Extend(DistMap& firstMap) const {
        Translate(firstMap.map_);
        firstMap.stale_ = true;
    }

private:
    mpi::Comm comm_;
    std::vector<int> map_;
    mutable detail::DeviceBuffer dev_;
    mutable bool stale_ = true;
};

// ---------------------------------------------------------------------------
// separator tree and symbolic info (symbolic/dist_separator_tree.hpp,
// symbolic/dist_symm_info.hpp)
// ---------------------------------------------------------------------------
struct SepOrLeaf {
    int parent;
    int off;
    std::vector<int> inds;
};
struct DistSeparator {
    mpi::Comm comm;
    int off;
    std::vector<int> inds;
};
struct DistSeparatorTree {
    std::vector<SepOrLeaf*> localSepsAndLeaves;
    std::vector<DistSeparator> distSeps;
    DistSeparatorTree() = default;
    DistSeparatorTree(const DistSeparatorTree&) = delete;
    DistSeparatorTree& operator=(const DistSeparatorTree&) = delete;
    ~DistSeparatorTree() { Clear(); }
    void Clear() {
        for (SepOrLeaf* p : localSepsAndLeaves) delete p;
        localSepsAndLeaves.clear();
    }
};

struct SymmNodeInfo {
    int size, off;
    int parent;
    std::vector<int> children;
    std::vector<int> origLowerStruct;
    bool onLeft;
    int myOff;
    std::vector<int> lowerStruct;
    std::vector<int> origLowerRelInds;
    std::vector<int> leftRelInds, rightRelInds;
};
struct DistSymmNodeInfo {
    int size, off;
    std::vector<int> origLowerStruct;
    bool onLeft;
    mpi::Comm comm;
    int myOff, leftSize, rightSize;
    std::vector<int> lowerStruct;
    std::vector<int> origLowerRelInds;
    std::vector<int> leftRelInds, rightRelInds;
};
struct DistSymmInfo {
    std::vector<SymmNodeInfo> localNodes;
    std::vector<DistSymmNodeInfo> distNodes;
    std::shared_ptr<cliqg_symbolic_s> symbolic;   // the library's analysis
};

namespace detail {
inline std::vector<int> node_array(cliqg_symbolic h, int s, int which, int n) {
    std::vector<int> v(n);
    if (n) check(cliqg_symbolic_node_array(h, s, which, v.data()), "symbolic node");
    return v;
}
}  // namespace detail

namespace detail {
// map, sepTree and info of a finished symbolic analysis h (N vertices)
inline void take_symbolic(cliqg_symbolic h, int N, DistMap& map, DistSeparatorTree& sepTree,
                          DistSymmInfo& info, const char* what) {
    info.symbolic.reset(h, detail::SymbolicDeleter());
    int num = 0;
    detail::check(cliqg_symbolic_num_nodes(h, &num, nullptr), what);
    map.Resize(N);
    detail::check(cliqg_symbolic_map(h, map.Map().data()), what);
    sepTree.Clear();
    sepTree.distSeps.clear();
    info.localNodes.assign(num, SymmNodeInfo());
    for (int s = 0; s < num; ++s) {
        int in[8];
        detail::check(cliqg_symbolic_node(h, s, in), what);
        SymmNodeInfo& nd = info.localNodes[s];
        nd.size = in[0];
        nd.off = in[1];
        nd.myOff = in[1];
        nd.parent = in[2];
        nd.children.clear();
        if (in[3] >= 0) nd.children = {in[3], in[4]};
        nd.onLeft = in[7] != 0;
        nd.lowerStruct = detail::node_array(h, s, CLIQG_NODE_LOWER_STRUCT, in[5]);
        nd.origLowerStruct = detail::node_array(h, s, CLIQG_NODE_ORIG_LOWER_STRUCT, in[6]);
        nd.origLowerRelInds = detail::node_array(h, s, CLIQG_NODE_ORIG_LOWER_REL, in[6]);
        nd.leftRelInds.clear();
        nd.rightRelInds.clear();
        if (in[3] >= 0) {
            nd.leftRelInds = detail::node_array(h, s, CLIQG_NODE_LEFT_REL,
                                                (int)info.localNodes[in[3]].lowerStruct.size());
            nd.rightRelInds = detail::node_array(h, s, CLIQG_NODE_RIGHT_REL,
                                                 (int)info.localNodes[in[4]].lowerStruct.size());
        }
        SepOrLeaf* sl = new SepOrLeaf;
        sl->parent = nd.parent;
        sl->off = nd.off;
        sl->inds = detail::node_array(h, s, CLIQG_NODE_INDS, nd.size);
        sepTree.localSepsAndLeaves.push_back(sl);
    }
    // at world size 1 the single distributed node mirrors the local root
    // (src/symbolic/DistSymmetricAnalysis.cpp:377-392)
    info.distNodes.assign(1, DistSymmNodeInfo());
    if (num) {
        const SymmNodeInfo& r = info.localNodes.back();
        DistSymmNodeInfo& d = info.distNodes[0];
        d.size = r.size;
        d.off = r.off;
        d.myOff = r.myOff;
        d.onLeft = r.onLeft;
        d.lowerStruct = r.lowerStruct;
        d.origLowerStruct = r.origLowerStruct;
        d.origLowerRelInds = r.origLowerRelInds;
        d.leftRelInds = r.leftRelInds;
        d.rightRelInds = r.rightRelInds;
        d.leftSize = d.rightSize = -1;
    }
}
}  // namespace detail

// Natural (geometric) nested dissection of an nx x ny x nz grid graph
// (include/clique/symbolic/natural_nested_dissection.hpp:17-26, 447-488).
inline void NaturalNestedDissection(int nx, int ny, int nz, const DistGraph& graph,
                                    DistMap& map, DistSeparatorTree& sepTree,
                                    DistSymmInfo& info, int cutoff = 128,
                                    bool storeFactRecvInds = false) {
    (void)storeFactRecvInds;
    const int N = graph.NumSources();
    cliqg_symbolic h = nullptr;
    detail::check(cliqg_symbolic_natural(nx, ny, nz, N, graph.Offsets().data(),
                                         graph.Targets().data(), cutoff, &h),
                  "NaturalNestedDissection");
    detail::take_symbolic(h, N, map, sepTree, info, "NaturalNestedDissection");
}

// Nested dissection of a general graph (symbolic/nested_dissection.hpp:35-44,
// 491-533): the reference's recursion, numbering, map and analysis, with the
// library's own vertex bisection in place of METIS (which may not be
// redistributed), so the separators differ from the reference's. The
// distribution arguments have no effect at world size 1.
inline void NestedDissection(const DistGraph& graph, DistMap& map, DistSeparatorTree& sepTree,
                             DistSymmInfo& info, bool sequential = true, int numDistSeps = 1,
                             int numSeqSeps = 1, int cutoff = 128,
                             bool storeFactRecvInds = false) {
    (void)sequential;
    (void)numDistSeps;
    (void)numSeqSeps;
    (void)storeFactRecvInds;
    const int N = graph.NumSources();
    cliqg_symbolic h = nullptr;
    detail::check(cliqg_symbolic_graph(N, graph.Offsets().data(), graph.Targets().data(), cutoff,
                                       &h),
                  "NestedDissection");
    detail::take_symbolic(h, N, map, sepTree, info, "NestedDissection");
}

// ---------------------------------------------------------------------------
// front tree (numeric/dist_symm_front_tree)
// ---------------------------------------------------------------------------
// SymmFrontType and its predicates and conversions
// (numeric/dist_symm_front_tree/decl.hpp:17-168)
enum SymmFrontType {
    SYMM_1D, SYMM_2D,
    LDL_1D, LDL_2D,
    LDL_SELINV_1D, LDL_SELINV_2D,
    LDL_INTRAPIV_1D, LDL_INTRAPIV_2D,
    LDL_INTRAPIV_SELINV_1D, LDL_INTRAPIV_SELINV_2D,
    BLOCK_LDL_1D, BLOCK_LDL_2D,
    BLOCK_LDL_INTRAPIV_1D, BLOCK_LDL_INTRAPIV_2D
};
inline bool Unfactored(SymmFrontType t) { return t == SYMM_1D || t == SYMM_2D; }
inline bool FrontsAre1d(SymmFrontType t) {
    return t == SYMM_1D || t == LDL_1D || t == LDL_SELINV_1D || t == LDL_INTRAPIV_1D ||
           t == LDL_INTRAPIV_SELINV_1D || t == BLOCK_LDL_1D || t == BLOCK_LDL_INTRAPIV_1D;
}
inline bool BlockFactorization(SymmFrontType t) {
    return t == BLOCK_LDL_1D || t == BLOCK_LDL_2D || t == BLOCK_LDL_INTRAPIV_1D ||
           t == BLOCK_LDL_INTRAPIV_2D;
}
inline bool SelInvFactorization(SymmFrontType t) {
    return t == LDL_SELINV_1D || t == LDL_SELINV_2D || t == LDL_INTRAPIV_SELINV_1D ||
           t == LDL_INTRAPIV_SELINV_2D;
}
inline bool PivotedFactorization(SymmFrontType t) {
    return t == LDL_INTRAPIV_1D || t == LDL_INTRAPIV_2D || t == LDL_INTRAPIV_SELINV_1D ||
           t == LDL_INTRAPIV_SELINV_2D || t == BLOCK_LDL_INTRAPIV_1D ||
           t == BLOCK_LDL_INTRAPIV_2D;
}
// the 1D and 2D members of each pair are adjacent, 1D first
inline SymmFrontType ConvertTo2d(SymmFrontType t) {
    if (t < SYMM_1D || t > BLOCK_LDL_INTRAPIV_2D) throw LogicError("Invalid front type");
    return SymmFrontType(t | 1);
}
inline SymmFrontType ConvertTo1d(SymmFrontType t) {
    if (t < SYMM_1D || t > BLOCK_LDL_INTRAPIV_2D) throw LogicError("Invalid front type");
    return SymmFrontType(t & ~1);
}
inline SymmFrontType AppendSelInv(SymmFrontType t) {
    switch (t) {
        case LDL_1D: return LDL_SELINV_1D;
        case LDL_2D: return LDL_SELINV_2D;
        case LDL_INTRAPIV_1D: return LDL_INTRAPIV_SELINV_1D;
        case LDL_INTRAPIV_2D: return LDL_INTRAPIV_SELINV_2D;
        default: throw LogicError("Sel-inv does not make sense for this type");
    }
}
inline SymmFrontType RemoveSelInv(SymmFrontType t) {
    switch (t) {
        case LDL_SELINV_1D: return LDL_1D;
        case LDL_SELINV_2D: return LDL_2D;
        case LDL_INTRAPIV_SELINV_1D: return LDL_INTRAPIV_1D;
        case LDL_INTRAPIV_SELINV_2D: return LDL_INTRAPIV_2D;
        default: throw LogicError("This type did not involve selective inversion");
    }
}
// the factorization LDL runs before converting to the requested type
// (dist_symm_front_tree/decl.hpp:158-168)
inline SymmFrontType InitialFactorType(SymmFrontType t) {
    if (Unfactored(t)) throw LogicError("Front type does not require factorization");
    if (BlockFactorization(t)) return ConvertTo2d(t);
    if (PivotedFactorization(t)) return LDL_INTRAPIV_2D;
    return LDL_2D;
}

// Arithmetic of the numeric factorization (no counterpart in the reference,
// which factors in the matrix's own precision).
enum FactorPrecision { FACTOR_FP64 = 0, FACTOR_FP32 = 1 };

template <typename F>
struct SymmFront {
    Matrix<F> frontL;
    Matrix<F> diag;
    Matrix<F> subdiag;
    Matrix<int> piv;
    mutable Matrix<F> work;
};
// A distributed front (dist_symm_front_tree/decl.hpp:187-204). At world size
// 1 there is one, the leaf of the distributed tree, which is the local root:
// front1dL and front2dL both hold that front.
template <typename F>
struct DistSymmFront {
    Matrix<F> front1dL, front2dL;
    Matrix<F> diag, subdiag;
    Matrix<int> piv;
};

template <typename F> class DistSymmFrontTree;
template <typename F> void ChangeFrontType(DistSymmFrontTree<F>& L, SymmFrontType frontType);
template <typename F>
void LDL(DistSymmInfo& info, DistSymmFrontTree<F>& L, SymmFrontType newType = LDL_2D,
         FactorPrecision precision = FACTOR_FP64);

template <typename F>
class DistSymmFrontTree {
public:
    bool isHermitian = false;
    SymmFrontType frontType = SYMM_2D;
    // Host mirrors of the device fronts, filled by MirrorFronts(). After
    // ChangeFrontType(L, SYMM_1D/SYMM_2D) on a factored tree they are the
    // values the next LDL factors (edits to frontL go to the device then).
    mutable std::vector<SymmFront<F>> localFronts;
    mutable std::vector<DistSymmFront<F>> distFronts;

    DistSymmFrontTree() = default;
    DistSymmFrontTree(const DistSparseMatrix<F>& A, const DistMap& reordering,
                      const DistSeparatorTree& sepTree, const DistSymmInfo& info,
                      bool conjugate = false) {
        Initialize(A, reordering, sepTree, info, conjugate);
    }

    // (dist_symm_front_tree/impl.hpp:22-384) The scatter of A into the
    // fronts runs on the device at the start of LDL; this records the
    // pattern, builds the plan and uploads the values. Calling it again
    // with new values (same pattern) prepares a refactorization.
    void Initialize(const DistSparseMatrix<F>& A, const DistMap&, const DistSeparatorTree&,
                    const DistSymmInfo& info, bool conjugate = false) {
        if (!info.symbolic) throw LogicError("DistSymmInfo holds no symbolic analysis");
        const cliq::DistGraph& g = A.LockedDistGraph();
        if (!plan_ || sym_ != info.symbolic || conjugate != isHermitian) {
            cliqg_plan p = nullptr;
            detail::check(cliqg_plan_create(info.symbolic.get(), g.Offsets().data(),
                                            g.Targets().data(), detail::Field<F>::value,
                                            conjugate ? 1 : 0, nullptr, &p),
                          "DistSymmFrontTree::Initialize");
            plan_.reset(p, detail::PlanDeleter());
            sym_ = info.symbolic;
        }
        vals_.upload(A.Values().data(), A.Values().size() * sizeof(F));
        isHermitian = conjugate;
        frontType = SYMM_2D;
        hostFronts_ = false;
        localFronts.clear();
        distFronts.clear();
    }

    // Copies every front (frontL, diag) of the factored tree to the host;
    // distFronts[0] mirrors the root.
    void MirrorFronts() const {
        const int num = (int)sym_nodes();
        localFronts.assign(num, SymmFront<F>());
        for (int s = 0; s < num; ++s) {
            int in[8];
            detail::check(cliqg_symbolic_node(sym_.get(), s, in), "MirrorFronts");
            const int n = in[0], m = in[0] + in[5];
            localFronts[s].frontL.Resize(m, n);
            // block fronts keep no diagonal (ldl/local.hpp:106-107)
            localFronts[s].diag.Resize(BlockFactorization(frontType) ? 0 : n, 1);
            detail::check(cliqg_plan_get_front(plan_.get(), s, localFronts[s].frontL.Buffer(),
                                               BlockFactorization(frontType)
                                                   ? nullptr
                                                   : localFronts[s].diag.Buffer()),
                          "MirrorFronts");
            // pivoted fronts also keep D's subdiagonal and the permutation
            // (ldl/local.hpp:108-115)
            if (PivotedFactorization(frontType) && !BlockFactorization(frontType)) {
                localFronts[s].subdiag.Resize(n, 1);
                localFronts[s].piv.Resize(n, 1);
                detail::check(cliqg_plan_get_front_pivots(plan_.get(), s,
                                                          localFronts[s].subdiag.Buffer(),
                                                          localFronts[s].piv.Buffer()),
                              "MirrorFronts");
            }
        }
        distFronts.assign(1, DistSymmFront<F>());
        if (num) {
            distFronts[0].front1dL = localFronts.back().frontL;
            distFronts[0].front2dL = localFronts.back().frontL;
            distFronts[0].diag = localFronts.back().diag;
            distFronts[0].subdiag = localFronts.back().subdiag;
            distFronts[0].piv = localFronts.back().piv;
        }
    }

    // Sends the host fronts (localFronts[s].frontL, unfactored values) to
    // the plan; the next LDL factors them instead of A's values.
    void PushFronts() const {
        for (int s = 0; s < (int)localFronts.size(); ++s)
            detail::check(cliqg_plan_set_front(plan_.get(), s, localFronts[s].frontL.LockedBuffer()),
                          "PushFronts");
    }

    void MemoryInfo(double& numLocalEntries, double& minLocalEntries, double& maxLocalEntries,
                    double& numGlobalEntries) const {
        double f, s, e;
        detail::check(cliqg_symbolic_work(sym_.get(), 0, &f, &s, &e), "MemoryInfo");
        int N = 0;
        detail::check(cliqg_symbolic_num_nodes(sym_.get(), nullptr, &N), "MemoryInfo");
        numLocalEntries = minLocalEntries = maxLocalEntries = numGlobalEntries = e + N;
    }
    // (dist_symm_front_tree/impl.hpp:540-590)
    void FactorizationWork(double& numLocalFlops, double& minLocalFlops, double& maxLocalFlops,
                           double& numGlobalFlops, bool selInv = false) const {
        (void)selInv;
        double f, s, e;
        detail::check(cliqg_symbolic_work(sym_.get(), IsComplex<F>::val, &f, &s, &e),
                      "FactorizationWork");
        numLocalFlops = minLocalFlops = maxLocalFlops = numGlobalFlops = f;
    }
    // (dist_symm_front_tree/impl.hpp:592-639)
    void SolveWork(double& numLocalFlops, double& minLocalFlops, double& maxLocalFlops,
                   double& numGlobalFlops, int numRhs = 1) const {
        double f, s, e;
        detail::check(cliqg_symbolic_work(sym_.get(), IsComplex<F>::val, &f, &s, &e),
                      "SolveWork");
        numLocalFlops = minLocalFlops = maxLocalFlops = numGlobalFlops = s * numRhs;
    }

    cliqg_plan Plan() const { return plan_.get(); }
    const void* DeviceValues() const { return vals_.get(); }
    int Size() const {
        int N = 0;
        detail::check(cliqg_symbolic_num_nodes(sym_.get(), nullptr, &N), "Size");
        return N;
    }
    double LastFactorMilliseconds() const {
        double ms = 0;
        detail::check(cliqg_plan_last_factor_ms(plan_.get(), &ms), "LastFactorMilliseconds");
        return ms;
    }

private:
    template <typename G> friend void ChangeFrontType(DistSymmFrontTree<G>&, SymmFrontType);
    template <typename G>
    friend void LDL(DistSymmInfo&, DistSymmFrontTree<G>&, SymmFrontType, FactorPrecision);
    size_t sym_nodes() const {
        int num = 0;
        detail::check(cliqg_symbolic_num_nodes(sym_.get(), &num, nullptr), "front count");
        return (size_t)num;
    }
    std::shared_ptr<cliqg_symbolic_s> sym_;
    std::shared_ptr<cliqg_plan_s> plan_;
    detail::DeviceBuffer vals_;
    bool hostFronts_ = false;   // localFronts hold the values the next LDL factors
};

// (numeric/change_front_type.hpp:26-109) At world size 1 the 1D and 2D
// layouts of the single distributed front are the same storage, and
// selective inversion touches distributed fronts s >= 1 only, of which there
// are none: those changes relabel the tree. Resetting a factored tree to
// SYMM_1D/SYMM_2D (tests/Refactor.cpp:150) hands its fronts to the caller
// as localFronts, which the next LDL factors; the reference has no such
// branch and throws "Unavailable front type change" there.
template <typename F>
inline void ChangeFrontType(DistSymmFrontTree<F>& L, SymmFrontType frontType) {
    const SymmFrontType old = L.frontType;
    if (frontType == old) return;
    if (frontType == ConvertTo2d(old) || frontType == ConvertTo1d(old)) {
        L.frontType = frontType;
    } else if (SelInvFactorization(frontType) &&
               ConvertTo2d(frontType) == ConvertTo2d(AppendSelInv(old))) {
        L.frontType = frontType;
    } else if (Unfactored(frontType) && !Unfactored(old)) {
        if (L.localFronts.empty()) L.MirrorFronts();
        L.hostFronts_ = true;
        L.frontType = frontType;
    } else {
        throw LogicError("Unavailable front type change");
    }
}

// (numeric/ldl.hpp:57-75) Factors the tree in place: the initial
// factorization (InitialFactorType) in the requested precision, then the
// conversion to the requested front type. LDL_* and LDL_SELINV_* give
// unit-L fronts, BLOCK_LDL_* inverted diagonal blocks. A tree is refactored
// by calling Initialize again (new values) or ChangeFrontType to SYMM_*
// (the edited localFronts).
template <typename F>
inline void LDL(DistSymmInfo& info, DistSymmFrontTree<F>& L, SymmFrontType newType,
                FactorPrecision precision) {
    (void)info;
    if (!Unfactored(L.frontType)) throw LogicError("Matrix is already factored");
    ChangeFrontType(L, SYMM_2D);
    const SymmFrontType initial = InitialFactorType(newType);
    if (L.hostFronts_) L.PushFronts();
    const bool blk = BlockFactorization(initial), piv = PivotedFactorization(initial);
    const cliqg_front_type ft = blk ? (piv ? CLIQG_FRONT_BLOCK_LDL_INTRAPIV : CLIQG_FRONT_BLOCK_LDL)
                                    : (piv ? CLIQG_FRONT_LDL_INTRAPIV : CLIQG_FRONT_LDL);
    detail::check(cliqg_plan_factor(L.Plan(), L.DeviceValues(),
                                    precision == FACTOR_FP32 ? CLIQG_F32 : CLIQG_F64, ft),
                  "LDL");
    L.hostFronts_ = false;
    L.localFronts.clear();
    L.distFronts.clear();
    L.frontType = initial;
    ChangeFrontType(L, newType);
}

// ---------------------------------------------------------------------------
// nodal multi-vectors (numeric/dist_nodal_multi_vec): device-resident
// ---------------------------------------------------------------------------
template <typename F>
class DistNodalMultiVec {
public:
    DistNodalMultiVec() = default;
    DistNodalMultiVec(const DistMap& inverseMap, const DistSymmInfo& info,
                      const DistMultiVec<F>& X) {
        Pull(inverseMap, info, X);
    }
    // X_nodal(p, j) = X(inverseMap[p], j)   (impl.hpp:67-217), on the device
    void Pull(const DistMap& inverseMap, const DistSymmInfo&, const DistMultiVec<F>& X) {
        h_ = X.Height();
        w_ = X.Width();
        if (inverseMap.NumSources() != h_) throw LogicError("map and vector do not match");
        buf_.resize((size_t)h_ * w_ * sizeof(F));
        if (h_ && w_)
            detail::check(cliqg_permute_rows(h_, w_, inverseMap.DeviceMap(), 0,
                                             X.LockedDeviceBuffer(), X.LDim(), buf_.get(), h_,
                                             detail::Field<F>::value, nullptr),
                          "Pull");
    }
    // X(inverseMap[p], j) = X_nodal(p, j)   (impl.hpp:219-356), on the device
    void Push(const DistMap& inverseMap, const DistSymmInfo&, DistMultiVec<F>& X) const {
        if (inverseMap.NumSources() != h_) throw LogicError("map and vector do not match");
        if (X.Height() != h_ || X.Width() != w_) X.ResizeTo(h_, w_);
        if (h_ && w_)
            detail::check(cliqg_permute_rows(h_, w_, inverseMap.DeviceMap(), 1, buf_.get(), h_,
                                             X.DeviceBuffer(), X.LDim(),
                                             detail::Field<F>::value, nullptr),
                          "Push");
    }
    int Height() const { return h_; }
    int Width() const { return w_; }
    void* DeviceBuffer() const { return buf_.get(); }
    // copy of another nodal multi-vector (device to device)
    void Assign(const DistNodalMultiVec<F>& X) {
        h_ = X.h_;
        w_ = X.w_;
        const size_t bytes = (size_t)h_ * w_ * sizeof(F);
        buf_.resize(bytes);
        if (bytes)
            detail::check(cliqg_memcpy(buf_.get(), X.buf_.get(), bytes, 2), "DistNodalMultiVec copy");
    }

private:
    int h_ = 0, w_ = 0;
    detail::DeviceBuffer buf_;
};

// (numeric/dist_nodal_matrix/decl.hpp:19-48) The right-hand sides for 2D
// front types. At world size 1 the [MC,MR] distribution of each node's block
// is the whole block, so the storage is the nodal multi-vector's: one
// device-resident N x k block in nodal order.
template <typename F>
class DistNodalMatrix {
public:
    DistNodalMatrix() = default;
    DistNodalMatrix(const DistMap& inverseMap, const DistSymmInfo& info,
                    const DistMultiVec<F>& X) {
        Pull(inverseMap, info, X);
    }
    explicit DistNodalMatrix(const DistNodalMultiVec<F>& X) { *this = X; }
    const DistNodalMatrix<F>& operator=(const DistNodalMultiVec<F>& X) {
        v_.Assign(X);
        return *this;
    }
    void Pull(const DistMap& inverseMap, const DistSymmInfo& info, const DistMultiVec<F>& X) {
        v_.Pull(inverseMap, info, X);
    }
    void Push(const DistMap& inverseMap, const DistSymmInfo& info, DistMultiVec<F>& X) const {
        v_.Push(inverseMap, info, X);
    }
    int Height() const { return v_.Height(); }
    int Width() const { return v_.Width(); }
    void* DeviceBuffer() const { return v_.DeviceBuffer(); }

private:
    DistNodalMultiVec<F> v_;
};

namespace detail {
template <typename F>
inline void solve_nodal(const DistSymmFrontTree<F>& L, void* buf, int h, int w) {
    if (Unfactored(L.frontType)) throw LogicError("Nonsensical front type for solve");
    detail::check(cliqg_plan_solve_nodal(L.Plan(), buf, std::max(h, 1), w), "Solve");
}
}  // namespace detail

// (numeric/solve.hpp:40-92) 1D front types solve against a
// DistNodalMultiVec, 2D types against a DistNodalMatrix; block front types
// skip the diagonal solve inside the engine. As in the reference, 1D block
// fronts are rejected by the distributed sweep (lower_solve/dist.hpp:54-56).
template <typename F>
inline void Solve(const DistSymmInfo& info, const DistSymmFrontTree<F>& L,
                  DistNodalMultiVec<F>& X) {
    (void)info;
    if (!FrontsAre1d(L.frontType)) throw LogicError("Invalid front type for 1D solve");
    if (Unfactored(L.frontType)) throw LogicError("Nonsensical front type for solve");
    if (L.frontType == BLOCK_LDL_1D || L.frontType == BLOCK_LDL_INTRAPIV_1D)
        throw LogicError("1D block solves not yet supported by this routine");
    detail::solve_nodal(L, X.DeviceBuffer(), X.Height(), X.Width());
}
template <typename F>
inline void Solve(const DistSymmInfo& info, const DistSymmFrontTree<F>& L,
                  DistNodalMatrix<F>& X) {
    (void)info;
    if (FrontsAre1d(L.frontType)) throw LogicError("Invalid front type for 2D solve");
    detail::solve_nodal(L, X.DeviceBuffer(), X.Height(), X.Width());
}

// FP64 iterative refinement of A X = B (no reference counterpart): X is
// overwritten with the refined solution (device-resident: B and X stay in
// HBM, nothing crosses to the host); returns the refinement report.
template <typename F>
inline cliqg_refine_info Refine(const DistSymmFrontTree<F>& L, const DistMultiVec<F>& B,
                                DistMultiVec<F>& X, double tol = 1e-10, int maxIters = 20) {
    if (X.Height() != B.Height() || X.Width() != B.Width()) X.ResizeTo(B.Height(), B.Width());
    cliqg_refine_info r{};
    const cliqg_status st = cliqg_plan_refine(L.Plan(), B.LockedDeviceBuffer(), B.LDim(),
                                              X.DeviceBuffer(), X.LDim(), B.Width(), tol, maxIters,
                                              &r);
    if (st != CLIQG_ERR_NONCONVERGED) detail::check(st, "Refine");
    return r;
}

// (numeric/multiply.hpp:27-166) Y := alpha A X + beta Y on the device (A's
// CSR and both multi-vectors stay in HBM).
template <typename T>
inline void Multiply(T alpha, const DistSparseMatrix<T>& A, const DistMultiVec<T>& X, T beta,
                     DistMultiVec<T>& Y) {
    if (A.Height() != Y.Height() || A.Width() != X.Height() || X.Width() != Y.Width())
        throw LogicError("A, X, and Y did not conform");
    if (std::imag(std::complex<double>(alpha)) != 0 || std::imag(std::complex<double>(beta)) != 0)
        throw LogicError("Multiply supports real alpha and beta");
    const int N = A.Height(), k = X.Width();
    if (N == 0 || k == 0) return;
    detail::check(cliqg_spmv(N, A.DeviceRowOffsets(), A.DeviceCols(), A.DeviceValues(),
                             detail::Field<T>::value, std::real(std::complex<double>(alpha)),
                             X.LockedDeviceBuffer(), X.LDim(),
                             std::real(std::complex<double>(beta)), Y.DeviceBuffer(), Y.LDim(), k,
                             nullptr),
                  "Multiply");
}

// The reference's black-box solve (numeric/solve.hpp:94-116): nested
// dissection of A's graph, LDL with intra-front pivoting (LDL_INTRAPIV_1D,
// numeric/solve.hpp:110), Pull, Solve, Push.
template <typename F>
inline void SymmetricSolve(const DistSparseMatrix<F>& A, DistMultiVec<F>& X, bool conjugate = false,
                           bool sequential = true, int numDistSeps = 1, int numSeqSeps = 1,
                           int cutoff = 128) {
    DistSymmInfo info;
    DistSeparatorTree sepTree;
    DistMap map, inverseMap;
    NestedDissection(A.LockedDistGraph(), map, sepTree, info, sequential, numDistSeps, numSeqSeps,
                     cutoff);
    map.FormInverse(inverseMap);
    DistSymmFrontTree<F> frontTree(A, map, sepTree, info, conjugate);
    LDL(info, frontTree, LDL_INTRAPIV_1D);
    DistNodalMultiVec<F> XNodal;
    XNodal.Pull(inverseMap, info, X);
    Solve(info, frontTree, XNodal);
    XNodal.Push(inverseMap, info, X);
}
template <typename F>
inline void HermitianSolve(const DistSparseMatrix<F>& A, DistMultiVec<F>& X, bool sequential = true,
                           int numDistSeps = 1, int numSeqSeps = 1, int cutoff = 128) {
    SymmetricSolve(A, X, true, sequential, numDistSeps, numSeqSeps, cutoff);
}

// The same on a grid graph: natural dissection of the nx x ny x nz grid
// (the geometric ordering the reference's tests use), LDL (intra-front
// pivoting unless frontType says otherwise) and Solve.
template <typename F>
inline void SymmetricSolve(const DistSparseMatrix<F>& A, DistMultiVec<F>& X, int nx, int ny,
                           int nz, bool conjugate = false, int cutoff = 128,
                           SymmFrontType frontType = LDL_INTRAPIV_1D) {
    DistSymmInfo info;
    DistSeparatorTree sepTree;
    DistMap map, inverseMap;
    NaturalNestedDissection(nx, ny, nz, A.LockedDistGraph(), map, sepTree, info, cutoff);
    map.FormInverse(inverseMap);
    DistSymmFrontTree<F> frontTree(A, map, sepTree, info, conjugate);
    LDL(info, frontTree, frontType);
    DistNodalMultiVec<F> XNodal;
    XNodal.Pull(inverseMap, info, X);
    Solve(info, frontTree, XNodal);
    XNodal.Push(inverseMap, info, X);
}
template <typename F>
inline void HermitianSolve(const DistSparseMatrix<F>& A, DistMultiVec<F>& X, int nx, int ny,
                           int nz, int cutoff = 128) {
    SymmetricSolve(A, X, nx, ny, nz, true, cutoff);
}

}  // namespace cliq

#endif
