// Device plan: front layout in HBM, static level schedule, numeric drivers.
#pragma once
#include <map>
#include <string>
#include <memory>
#include <vector>

#include "kernels.hpp"
#include "symbolic.hpp"

namespace cliqg {

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    void alloc(size_t n);
    void release();
    template <class X> X* as() const { return static_cast<X*>(p); }
};

// Per-front layout (host copy).
struct FrontLayout {
    int n, u, m, off;
    int parent, slot, height, nchild;
    int child[2];
    int64_t L;      // factor pool offset (elements)
    int64_t U;      // update pool offset
    int64_t rel;    // offset of relInds into the parent's front (length u)
    int64_t low;    // offset of lowerStruct (length u)
    int64_t W;      // solve workspace row offset (prefix of u)
    int64_t Linv;   // inverse diagonal block pool offset
    int64_t LinvW;  // wide inverse diagonal block pool offset (blocks of SB)
};

// Default width of the diagonal blocks whose inverses the LDL solves apply
// (a multiple of the 64-wide factor panels).
constexpr int SB_DEFAULT = 128;

// Opaque per-working-type numeric state.
struct NumericBase {
    virtual ~NumericBase() = default;
    virtual size_t device_bytes() const = 0;
};

// The autotuner's table as text, one line per tuned bucket: the bucket
// (t<type>.r<role>.l<lower>.T<log2 tiles>.K<log2 mean K>, and for solve
// products .N<16-column blocks of right-hand sides>), the chosen tile
// shape and the milliseconds each shape (32x32, 64x64, 128x128 four-wave,
// 128x128 eight-wave, 32x32 with 64-deep chunks, streaming) took over the
// bucket's launches, -1 where not built.
std::string tune_table_text();

class Plan {
public:
    Plan(std::shared_ptr<const Symbolic> sym, const int64_t* rowOff, const int* cols,
         bool complex, bool conjugate, hipStream_t stream);
    ~Plan();

    void factor(const void* d_vals, int precision, int frontType);
    void solve(void* d_x, int64_t ldx, int k, bool nodal);
    // solve() without allocating device memory (all workspace in the
    // library's static scratch array of the plan's device)
    void solve_noalloc(void* d_x, int64_t ldx, int k);
    void multiply(double alpha, const void* d_x, int64_t ldx, double beta, void* d_y,
                  int64_t ldy, int k);
    void refine(const void* d_b, int64_t ldb, void* d_x, int64_t ldx, int k, double tol,
                int maxIters, int* iters, int* converged, double* maxEta);
    void adjoint(const void* d_g, int64_t ldg, const void* d_x, int64_t ldx, int k,
                 void* d_lambda, int64_t ldl, void* d_dA);
    void get_front(int s, void* h_L, void* h_diag);
    // pivoted factorizations: D's subdiagonal (n FP64 field values) and the
    // pivot permutation (n local indices) of front s
    void get_pivots(int s, void* h_sub, int* h_piv);
    // stage front s's unfactored values (m x n, FP64 field type) for the
    // next factorization, which uses them in place of A's assembled values
    void set_front(int s, const void* h_L);
    size_t device_bytes() const;
    double last_factor_ms() const { return lastFactorMs_; }

    // shared layout, read by the numeric engines
    std::shared_ptr<const Symbolic> sym;
    bool complex, conjugate;
    int N = 0, numFronts = 0, numLevels = 0;
    int64_t nnz = 0;
    std::vector<FrontLayout> fronts;
    std::vector<std::vector<int>> levels;   // fronts of each height
    int64_t factorEntries = 0, updateEntries = 0, linvEntries = 0, linvWEntries = 0, wRows = 0;
    int sb = SB_DEFAULT;   // solve block width (CLIQG_SOLVE_BLOCK overrides; multiple of 64)
    bool lookahead = true; // side-stream outer-panel updates (CLIQG_LOOKAHEAD=0 disables)
    std::vector<int64_t> hostRowOff;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    int device = 0;   // the device current at creation (the plan's memory lives there)
    // side stream (default priority) for work that overlaps the main
    // sequence, and the events that order the two
    hipStream_t stream2 = nullptr;
    std::vector<hipEvent_t> depEvents;
    void ensure_dep_events(int n);

    DevBuf dRowOff, dCols, dAsmDst, dRel, dLow, dMap, dVals, dStatus, dNorms;
    std::vector<int> relHost;   // relative indices of every front (host copy of dRel)
    int precision = -1, frontType = 0;   // frontType: cliqg_front_type
    bool factored = false;
    double normA = 0.0;

    std::unique_ptr<NumericBase> num;
    std::map<int, std::vector<char>> stagedFronts;   // front -> m x n FP64 field values
    int check_status(const char* what);

    // opt-in per-launch profiling (HIP events around every scheduled launch)
    struct Pending {
        const char* name;
        double flops, bytes;
        hipEvent_t e0, e1;
        int ndesc, nwork, variant, stream;
    };
    struct ProfileEntry {
        int launches = 0;
        double ms = 0.0, flops = 0.0, bytes = 0.0;
    };
    bool profiling = false;
    bool useGraphs = false;  // replay the factor and solve launch sequences as HIP graphs (opt-in)
    std::vector<Pending> pending;
    std::map<std::string, ProfileEntry> profile;
    hipEvent_t next_event();
    void reset_events() { evNext_ = 0; }
    void resolve_profile();
    // Times every admissible GEMM tile shape of the factorization (front
    // type ftype, precision prec) of d_vals and, with rhs > 0, of a solve
    // with rhs right-hand sides, records the fastest per launch bucket in the
    // process-wide tuning table and factors d_vals again with the tuned
    // shapes. nBuckets: buckets measured; nChanged: buckets whose shape
    // changed.
    void autotune(const void* d_vals, int prec, int ftype, int rhs, int* nBuckets, int* nChanged);

private:
    double lastFactorMs_ = 0.0;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr, evCaller_ = nullptr;
    // Order the plan's own stream after the caller's prior work on the
    // default stream (the plan creates a non-blocking stream when given none).
    void after_caller();
    std::vector<hipEvent_t> evPool_;
    size_t evNext_ = 0;
};

}  // namespace cliqg
