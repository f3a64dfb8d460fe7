// Device plan of the multifrontal LDL^T engine (host orchestration).
//
// Memory layout in HBM (one allocation per pool):
//   factor pool  : front s is an m x n column-major block (ld m) at L_s,
//                  the reference's frontL (dist_symm_front_tree/decl.hpp:174-185)
//   update pool  : front s's u x u Schur complement (ld u) at U_s, the
//                  reference's `work` matrix during LDL (ldl/local.hpp:42-51)
//   pivot pool   : D in nodal order, FP64 field type for every precision
//   inverse pool : per 64-wide diagonal block of every front, the inverse of
//                  its unit lower triangle (64 x 64, ld 64)
//   workspace    : per front an m x k solve workspace W_s (ld m)
// The schedule is level-synchronous: fronts of equal height are processed by
// the same grouped launches; its descriptor arrays are built once.
#include "plan.hpp"

#include "../../include/cliqg.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>

namespace cliqg {

namespace {

// Scratch memory of the allocation-free solve (cliqg_plan_solve_noalloc),
// one copy per device, held until the process ends. Host code never uses
// the array's name for its memory: scratch_address() looks up each device's
// address once, with hipGetSymbolAddress, and every kernel argument, copy
// and clear uses that address.
constexpr size_t NOALLOC_SCRATCH_BYTES = CLIQG_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex scratchMutex;
std::map<int, char*> scratchAddress;   // device -> its copy of the array

char* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(scratchMutex);
    auto it = scratchAddress.find(device);
    if (it != scratchAddress.end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(5, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                           hipGetErrorString(e));
    scratchAddress[device] = static_cast<char*>(p);
    return static_cast<char*>(p);
}

}  // namespace

void DevBuf::alloc(size_t n) {
    release();
    if (n) CQ_HIP(hipMalloc(&p, n));
    bytes = n;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

namespace {

template <class X>
void upload(DevBuf& b, const std::vector<X>& v) {
    b.alloc(std::max<size_t>(v.size(), 1) * sizeof(X));
    if (!v.empty()) CQ_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
}

// Descriptor arrays of a schedule, staged on the host and uploaded at once
// (or, with `base` set, copied by the caller to device memory at `base`
// before every use: the allocation-free solve).
struct Arena {
    std::vector<char> host;
    DevBuf dev;
    const char* base = nullptr;
    template <class X>
    size_t add(const std::vector<X>& v) {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + std::max<size_t>(v.size(), 1) * sizeof(X));
        if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(X));
        return off;
    }
    void finish() {
        dev.alloc(std::max<size_t>(host.size(), 1));
        if (!host.empty())
            CQ_HIP(hipMemcpy(dev.p, host.data(), host.size(), hipMemcpyHostToDevice));
        std::vector<char>().swap(host);
    }
    template <class X> const X* at(size_t off) const {
        return reinterpret_cast<const X*>((base ? base : static_cast<const char*>(dev.p)) + off);
    }
};

// First-fit allocator over a virtual range with coalescing free list.
struct IntervalAlloc {
    std::map<int64_t, int64_t> freeList;   // offset -> size
    int64_t top = 0;
    int64_t alloc(int64_t n) {
        if (n <= 0) return 0;
        for (auto it = freeList.begin(); it != freeList.end(); ++it) {
            if (it->second >= n) {
                const int64_t off = it->first, rest = it->second - n;
                freeList.erase(it);
                if (rest > 0) freeList[off + n] = rest;
                return off;
            }
        }
        // grow, absorbing a free block that ends at the top
        if (!freeList.empty()) {
            auto last = std::prev(freeList.end());
            if (last->first + last->second == top) {
                const int64_t off = last->first;
                freeList.erase(last);
                top = off + n;
                return off;
            }
        }
        const int64_t off = top;
        top += n;
        return off;
    }
    void release(int64_t off, int64_t n) {
        if (n <= 0) return;
        auto it = freeList.emplace(off, n).first;
        if (it != freeList.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                freeList.erase(it);
                it = prev;
            }
        }
        auto next = std::next(it);
        if (next != freeList.end() && it->first + it->second == next->first) {
            it->second += next->second;
            freeList.erase(next);
        }
    }
};

enum StepKind { S_EXTEND, S_DIAG, S_GEMM, S_GATHER, S_UNUSED, S_SEXTEND, S_ZERO, S_SOLVE_GEMM, S_COPY,
                S_RECORD, S_WAIT, S_REDUCE, S_ZLOWER, S_BK, S_TRINV, S_GCOLS, S_QSCALE, S_PERM,
                S_QROWS, S_PSYM, S_PSTEP, S_WINV };

struct Step {
    int kind;
    size_t desc, start;   // arena offsets
    int ndesc, nwork;
    bool backward;
    double flops = 0.0;   // algorithmic flops of the launch (reference convention)
    double bytes = 0.0;   // algorithmic bytes of the launch (HBM-bound steps)
    int variant = 0;      // GEMM tile shape
    // GEMM tile shapes the launch may take (the autotuner picks among them):
    // shape, arena offset of its tile-start array and its tile count
    int nvar = 0;
    int vars[NUM_GEMM_VARIANTS] = {};
    size_t vstart[NUM_GEMM_VARIANTS] = {};
    int vtiles[NUM_GEMM_VARIANTS] = {};
    int bucket = -1;      // the launch's autotuner bucket (index into the tuning table)
    bool scaled = false;  // GEMM descriptors carry pivots (scaleK)
    bool update = false;  // a factor update (profiled as gemm_update, role 3 kernels)
    int stream = 0;       // 0: the plan's stream, 1: its side stream
    int event = -1;       // S_RECORD / S_WAIT: index into the plan's dependency events
};

// ---------------------------------------------------------------------------
// GEMM tile autotuner. Launches fall into buckets by working type, role
// (TRSM-like product, solve product, pivot-scaled or factor update), lower
// masking, and the base-2 logarithms of their 64x64 tile count and mean K.
// Plan::autotune times every admissible tile shape of every grouped GEMM of
// a factorization (and of a solve) on the device and records per bucket the
// shape with the least total time; schedules built afterwards, and the
// tuned plan's own, use it. The table is shared by all plans of the process
// and, with CLIQG_TUNE_CACHE=<file>, read from and written to that file
// (one "bucket shape" line per entry).
// ---------------------------------------------------------------------------
struct TuneTable {
    std::mutex mu;
    std::map<std::string, int> index;
    std::vector<std::string> keys;
    std::vector<int> choice;   // tuned shape per bucket, -1 if none
    // milliseconds per tile shape (GEMM_VARIANTS order) summed over the
    // bucket's launches in the last tuning pass, -1 where not measured
    std::vector<std::array<double, NUM_GEMM_VARIANTS>> ms;
    bool loaded = false;
};
TuneTable& tune_table() {
    static TuneTable t;
    return t;
}
const char* tune_cache_path() {
    static const char* p = std::getenv("CLIQG_TUNE_CACHE");
    return p && *p ? p : nullptr;
}
int tune_bucket_locked(TuneTable& t, const std::string& key) {
    auto it = t.index.find(key);
    if (it != t.index.end()) return it->second;
    const int b = (int)t.keys.size();
    t.index.emplace(key, b);
    t.keys.push_back(key);
    t.choice.push_back(-1);
    std::array<double, NUM_GEMM_VARIANTS> none;
    none.fill(-1.0);
    t.ms.push_back(none);
    return b;
}
void tune_load_locked(TuneTable& t) {
    if (t.loaded) return;
    t.loaded = true;
    const char* path = tune_cache_path();
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char key[128];
    int v;
    while (std::fscanf(f, "%127s %d", key, &v) == 2) {
        bool known = false;
        for (int g : GEMM_VARIANTS) known = known || g == v;
        if (known) t.choice[tune_bucket_locked(t, key)] = v;
    }
    std::fclose(f);
}
int tune_bucket(int type, int role, bool lower, double tiles64, double meanK, int n) {
    auto lg = [](double x) { return x >= 1.0 ? (int)std::lround(std::log2(x)) : 0; };
    char key[128];
    // solve products also by their right-hand-side count (16-column blocks)
    if (role == 1)
        std::snprintf(key, sizeof(key), "t%d.r%d.l%d.T%d.K%d.N%d", type, role, lower ? 1 : 0,
                      lg(tiles64), lg(meanK), (std::min(n, 64) + 15) / 16);
    else
        std::snprintf(key, sizeof(key), "t%d.r%d.l%d.T%d.K%d", type, role, lower ? 1 : 0,
                      lg(tiles64), lg(meanK));
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    tune_load_locked(t);
    return tune_bucket_locked(t, key);
}
// CLIQG_GEMM_TILE=<shape>: every launch that admits the shape takes it
// (0: 64x64, 4: 128x128, 5: 32x32, 6: 128x128 eight-wave, 7: 32x32 with
// 64-deep K chunks, 8: streaming solve tiles), ahead of the
// tuned and the modelled choice
int forced_variant() {
    static const int v = [] {
        const char* e = std::getenv("CLIQG_GEMM_TILE");
        if (!e || !*e) return -1;
        const int t = std::atoi(e);
        for (int g : GEMM_VARIANTS)
            if (g == t) return t;
        return -1;
    }();
    return v;
}
int tuned_variant(int bucket) {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    return (bucket >= 0 && bucket < (int)t.choice.size()) ? t.choice[bucket] : -1;
}
void tune_save_locked(TuneTable& t) {
    const char* path = tune_cache_path();
    if (!path) return;
    FILE* f = std::fopen(path, "w");
    if (!f) return;
    for (size_t b = 0; b < t.keys.size(); ++b)
        if (t.choice[b] >= 0) std::fprintf(f, "%s %d\n", t.keys[b].c_str(), t.choice[b]);
    std::fclose(f);
}
int variant_slot(int v) {
    for (int i = 0; i < NUM_GEMM_VARIANTS; ++i)
        if (GEMM_VARIANTS[i] == v) return i;
    return -1;
}
template <class T> constexpr int type_code() {
    return Scalar<T>::complex ? (sizeof(T) == 8 ? 2 : 3) : (sizeof(T) == 4 ? 0 : 1);
}
bool step_has_variant(const Step& s, int v) {
    for (int i = 0; i < s.nvar; ++i)
        if (s.vars[i] == v) return true;
    return false;
}
// the launch takes tile shape v (one of its admissible shapes)
void set_variant(Step& s, int v) {
    for (int i = 0; i < s.nvar; ++i)
        if (s.vars[i] == v) {
            s.variant = v;
            s.start = s.vstart[i];
            s.nwork = s.vtiles[i];
            return;
        }
    throw Error(5, "a GEMM launch was given a tile shape it was not built for");
}

const char* step_name(int kind) {
    switch (kind) {
        case 0: return "extend_add";
        case 1: return "diag_ldl";
        case 2: return "gemm";
        case 7: return "solve_gemm";
        case 3: return "solve_gather";
        case 5: return "solve_extend_add";
        case 6: return "zero_update";
        case 8: return "copy";
        case 11: return "solve_reduce";
        case 12: return "zero_update";
        case 13: return "bk_panel";
        case 14: return "unit_lower_inverse";
        case 15: return "permute_cols";
        case 16: return "quasi_scale";
        case 17: return "solve_permute";
        case 18: return "quasi_rows";
        case 19: return "permute_sym";
        case 20: return "solve_panel";
        case 21: return "wide_inverse";
        default: return "other";
    }
}

template <class T>
struct StepBuilder {
    Arena& arena;
    std::vector<Step>& steps;
    int cur = 0;   // stream of the steps added next

    void on(int stream) { cur = stream; }
    void record(int ev) {
        Step x{S_RECORD, 0, 0, 0, 0, false};
        x.stream = cur;
        x.event = ev;
        steps.push_back(x);
    }
    void wait(int ev) {
        Step x{S_WAIT, 0, 0, 0, 0, false};
        x.stream = cur;
        x.event = ev;
        steps.push_back(x);
    }

    void zero(const std::vector<ZeroSeg<T>>& zs) {
        std::vector<int> start;
        int chunks = 0;
        std::vector<ZeroSeg<T>> keep;
        for (const auto& z : zs) {
            if (z.n <= 0) continue;
            keep.push_back(z);
            start.push_back(chunks);
            chunks += (int)((z.n + ZERO_CHUNK - 1) / ZERO_CHUNK);
        }
        if (keep.empty()) return;
        steps.push_back(Step{S_ZERO, arena.add(keep), arena.add(start), (int)keep.size(), chunks,
                             false});
        steps.back().stream = cur;
    }

    void copy(const std::vector<CopyDesc<T>>& cs) {
        std::vector<int> start;
        int chunks = 0;
        double by = 0.0;
        std::vector<CopyDesc<T>> keep;
        for (const auto& c : cs) {
            const int64_t e = (int64_t)c.rows * c.cols;
            if (e <= 0) continue;
            keep.push_back(c);
            start.push_back(chunks);
            chunks += (int)((e + 1023) / 1024);
            by += 2.0 * e * sizeof(T);
        }
        if (keep.empty()) return;
        steps.push_back(Step{S_COPY, arena.add(keep), arena.add(start), (int)keep.size(), chunks,
                             false, 0.0, by});
        steps.back().stream = cur;
    }

    void reduce(const std::vector<ReduceDesc<T>>& rs) {
        std::vector<int> start;
        int chunks = 0;
        double by = 0.0;
        for (const auto& r : rs) {
            const int64_t e = (int64_t)r.rows * r.cols;
            start.push_back(chunks);
            chunks += (int)((e + 1023) / 1024);
            by += (r.slices + ((r.flags & G_ACCUM) ? 2.0 : 1.0)) * e * sizeof(T);
        }
        if (rs.empty()) return;
        steps.push_back(Step{S_REDUCE, arena.add(rs), arena.add(start), (int)rs.size(), chunks,
                             false, 0.0, by});
        steps.back().stream = cur;
    }

    // triB: one operand is triangular (an inverted diagonal block, or a
    // triangular op(A) whose K range the kernel clamps), so only half of the
    // K products count as algorithmic work.
    // a factor update (named and instantiated apart from the TRSM-like products)
    void update(const std::vector<GemmDesc<T>>& in) { gemm_role(in, false, true); }

    bool solve = false;   // the steps are a solve's (GEMM role 1)

    void gemm(const std::vector<GemmDesc<T>>& in, bool triB = false) { gemm_role(in, triB, false); }

    // A grouped GEMM launch: its tile-start arrays for every admissible tile
    // shape go to the arena, the shape used is the autotuner's choice for
    // the launch's bucket or else the cost model's (choose_variant).
    void gemm_role(const std::vector<GemmDesc<T>>& in, bool triB, bool upd) {
        std::vector<GemmDesc<T>> ds;
        double flops = 0.0, bytes = 0.0, ksum = 0.0;
        bool inplaceWide = false, lower = false;
        for (const auto& d : in) {
            if (d.M <= 0 || d.N <= 0) continue;
            if (d.K == 0 && (d.flags & G_ACCUM)) continue;
            ds.push_back(d);
            const double M = d.M, N = d.N, K = d.K;
            const double entries = (d.flags & G_LOWER) ? M * N - N * (N - 1) / 2 : M * N;
            const double keff = triB ? (K + 1) / 2 : K;
            flops += 2.0 * entries * keff * (Scalar<T>::complex ? 4 : 1);
            bytes += (M * K + K * N + entries * ((d.flags & G_ACCUM) ? 2 : 1)) * sizeof(T);
            ksum += K;
            lower = lower || (d.flags & G_LOWER);
        }
        if (ds.empty()) return;
        const bool scaled = ds[0].scaleK != nullptr;
        for (const auto& d : ds)
            if ((d.scaleK != nullptr) != scaled)
                throw Error(1, "a grouped GEMM mixes scaled and unscaled descriptors");
        Step s{S_GEMM, arena.add(ds), 0, (int)ds.size(), 0, false, flops, bytes};
        s.scaled = scaled;
        s.update = upd;
        s.stream = cur;
        double tiles64 = 0.0;
        for (const auto& d : ds) tiles64 += gemm_tiles(d.M, d.N, d.flags, TILE_64);
        for (int v : GEMM_VARIANTS) {
            if (!gemm_variant_allowed<T>(v)) continue;
            // an in-place product (C is A or B, the TRSM of the factor) needs
            // every column of a row block in one tile: a second tile column
            // would read A after the first one has overwritten it
            bool ok = true;
            for (const auto& d : ds)
                if ((d.C == d.A || d.C == d.B) && d.N > gemm_tile_n(v)) ok = false;
            // a triangular K range (G_KLO, G_KHI) is widened to whole
            // chunks: 64-deep chunks read up to 63 entries of op(A) past
            // the triangle, which only the solves' inverted blocks hold
            // as zeros
            if (v == TILE_32_K64 && !solve)
                for (const auto& d : ds)
                    if (d.flags & (G_KLO | G_KHI)) ok = false;
            // streaming tiles: solve products of at most 64 right-hand
            // sides, plain or transposed A (a triangular K range only on the
            // solves' inverted blocks), no output masks, scales or aliasing
            if (v == TILE_STREAM) {
                ok = ok && solve && !scaled;
                for (const auto& d : ds)
                    if (d.N > 64 || (d.flags & ~STREAM_FLAGS) || d.scaleN || d.C == d.A ||
                        d.C == d.B)
                        ok = false;
            }
            if (!ok) continue;
            std::vector<int> start;
            int total = 0;
            for (const auto& d : ds) {
                start.push_back(total);
                total += gemm_tiles(d.M, d.N, d.flags, v);
            }
            s.vars[s.nvar] = v;
            s.vstart[s.nvar] = arena.add(start);
            s.vtiles[s.nvar] = total;
            ++s.nvar;
        }
        if (s.nvar == 0) throw Error(1, "an in-place GEMM needs all of its columns in one tile");
        const int role = scaled ? 2 : (upd ? 3 : (solve ? 1 : 0));
        int maxN = 0;
        for (const auto& d : ds) maxN = std::max(maxN, d.N);
        s.bucket = tune_bucket(type_code<T>(), role, lower, tiles64, ksum / ds.size(), maxN);
        const int forced = forced_variant();
        const int tuned = tuned_variant(s.bucket);
        int model = choose_variant(ds);
        // solve products of at most 32 right-hand sides stream op(A) (the
        // 120^3 PML solve with 32: 35.7 ms with every such product
        // streaming, 38.4 with the tile shapes)
        if (solve && maxN <= 32 && step_has_variant(s, TILE_STREAM)) model = TILE_STREAM;
        set_variant(s, forced >= 0 && step_has_variant(s, forced)  ? forced
                       : tuned >= 0 && step_has_variant(s, tuned) ? tuned
                                                                   : model);
        steps.push_back(s);
    }

    static constexpr double SMALL_TILE_LIMIT = 384.0;

    // Tile shape of a grouped launch: 64x64 tiles (four waves of 32x32) or
    // 128x128 tiles (four waves of 64x64, two blocks per CU). Cost = rounds of
    // 256 CUs the launch needs x MFMA work of one tile x the measured relative
    // cost per flop of the shape on gfx950 (real FP64 0.83, real FP32 0.95 for
    // the large tile). Complex FP32 always takes 64x64 tiles, whose
    // three-multiplication products beat the large tile's four; complex FP64
    // has no large tile (LDS).
    static int choose_variant(const std::vector<GemmDesc<T>>& in) {
        // Launches that fill at most about one round of 64x64 tiles are
        // latency-bound: 32x32 tiles put four times as many workgroups on
        // the CUs, each with a quarter of the MFMA chain (measured on complex
        // FP32: 128x64 with K = 128 8 us instead of 15, 14400x64 with
        // K = 128 16 instead of 20, 20000x64 with K = 7200 0.69 ms instead
        // of 1.12).
        static const double smallLimit = [] {
            const char* e = std::getenv("CLIQG_SMALL_TILES");
            const double v = e ? std::atof(e) : SMALL_TILE_LIMIT;
            return (v >= 0.0 && v <= 65536.0) ? v : SMALL_TILE_LIMIT;
        }();
        if (smallLimit > 0.0) {
            // An in-place product (C is A or B, the TRSM of the factor) needs
            // every column of a row block in one tile: a second tile column
            // would read A after the first one has overwritten it.
            double tiles0 = 0.0;
            bool inplace = false;
            for (const auto& d : in) {
                tiles0 += gemm_tiles(d.M, d.N, d.flags, TILE_64);
                inplace = inplace || ((d.C == d.A || d.C == d.B) && d.N > gemm_tile_n(TILE_32));
            }
            if (!inplace && tiles0 > 0.0 && tiles0 <= smallLimit) return TILE_32;
        }
        if (!gemm_variant_allowed<T>(TILE_128)) return TILE_64;
        const double rel4 = sizeof(T) == 4 ? 0.95 : 0.83;
        double cost[2] = {0.0, 0.0};
        const int vs[2] = {TILE_64, TILE_128};
        for (int c = 0; c < 2; ++c) {
            const int v = vs[c];
            double tiles = 0.0, work = 0.0;
            for (const auto& d : in) {
                const int t = gemm_tiles(d.M, d.N, d.flags, v);
                tiles += t;
                work += (double)t * std::max(d.K, 1);
            }
            if (tiles == 0) return TILE_64;
            // average K per tile times the tile area, times rounds over the CUs
            const double rounds = std::ceil(tiles / 256.0);
            cost[c] = (c ? rel4 : 1.0) * rounds * (work / tiles) * gemm_tile_m(v) * gemm_tile_n(v);
        }
        return cost[1] < cost[0] ? TILE_128 : TILE_64;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Numeric state for one working type T (float, double, cfloat, cdouble)
// ---------------------------------------------------------------------------
// Panels per super-panel of the LDL solves' fused panel steps
// (CLIQG_SOLVE_CHAIN; 0 keeps two grouped products per panel).
constexpr int CHAIN_DEFAULT = 8;
constexpr long CHAIN_MAX_TASKS = 512;
static int solve_chain() {
    static const int v = [] {
        const char* e = std::getenv("CLIQG_SOLVE_CHAIN");
        const int t = e ? std::atoi(e) : CHAIN_DEFAULT;
        return (t >= 0 && t <= 64) ? t : CHAIN_DEFAULT;
    }();
    return v;
}

// ---------------------------------------------------------------------------
// Outer panel width of the front factorization (a multiple of NB).
constexpr int OB_DEFAULT = 256;
static int outer_panel() {
    static const int ob = [] {
        const char* e = std::getenv("CLIQG_OUTER_PANEL");
        const int v = e ? std::atoi(e) : OB_DEFAULT;
        return (v >= NB && v % NB == 0) ? v : OB_DEFAULT;
    }();
    return ob;
}

template <class T>
struct Numeric : NumericBase {
    using V = typename Scalar<T>::wide;
    Plan& P;
    DevBuf fpool, upool, dpool, linv, linvw, wtmp;
    // W = L*D below the diagonal blocks of the current level's fronts (the
    // TRSM writes it next to L), the left operand of every update
    DevBuf wbuf;
    // factor schedules by front type (cliqg_front_type: LDL, block LDL,
    // pivoted LDL, pivoted block LDL), built at first use
    Arena farenas[4];
    std::vector<Step> fsched[4];
    bool fbuilt[4] = {false, false, false, false};
    int nextEvent = 0;                  // dependency events used by the schedules
    // block LDL work buffers, reused level by level: the fronts' original
    // A21, the unit-lower inverses of L11 and the n x 64 product panels
    DevBuf bsave, bsinv, btmp;
    // pivoted factorization (LDL_INTRAPIV): D's subdiagonal and each front's
    // pivot permutation, in nodal order, and its schedule
    DevBuf subpool, pivpool;

    // A captured launch sequence (HIP graph), replayed instead of the
    // individual launches.
    struct Graph {
        hipGraphExec_t exec = nullptr;
        bool tried = false;
        ~Graph() {
            if (exec) (void)hipGraphExecDestroy(exec);
        }
    };

    struct SolveSched {
        Arena arena;
        std::vector<Step> fwd, bwd;
        // Workspace the descriptors point into: buffers owned here for the
        // allocating entry points, the library's static scratch array for
        // the allocation-free solve.
        DevBuf wpool, xn, xn2, tmp;
        DevBuf splitWs;          // K slices of the split update-row products
        T* wp = nullptr;         // update rows W_s of every front, (sum of u) x k
        T* x1 = nullptr;         // nodal right-hand sides N x k: sweep input and result
        T* x2 = nullptr;         // second nodal buffer of the LDL sweeps (x1 for block fronts)
        int* status = nullptr;   // zero-pivot report of the diagonal solve
        int64_t wsUsed = 0;      // split-K workspace entries the schedule uses
        Graph graph;   // forward, diagonal and backward sweeps
    };
    Graph fgraph[4];   // the factorizations, by front type
    std::map<int, std::unique_ptr<SolveSched>> solves;
    // schedules of the allocation-free solve (descriptors into the scratch array)
    std::map<int, std::unique_ptr<SolveSched>> noallocSolves;
    DevBuf rbuf, identity;
    int rbufCols = 0;

    explicit Numeric(Plan& p) : P(p) {
        fpool.alloc(std::max<int64_t>(P.factorEntries, 1) * sizeof(T));
        upool.alloc(std::max<int64_t>(P.updateEntries, 1) * sizeof(T));
        dpool.alloc(std::max(P.N, 1) * sizeof(V));
        subpool.alloc(std::max(P.N, 1) * sizeof(V));
        pivpool.alloc(std::max(P.N, 1) * sizeof(int));
        linv.alloc(std::max<int64_t>(P.linvEntries, 1) * sizeof(T));
        linvw.alloc(std::max<int64_t>(P.linvWEntries, 1) * sizeof(T));
        {
            // n x 64 panel products of the wide inverses, reused level by level
            int64_t mx = 0;
            for (const auto& lv : P.levels) {
                int64_t a = 0;
                for (int s : lv) a += (int64_t)P.fronts[s].n * NB;
                mx = std::max(mx, a);
            }
            wtmp.alloc(std::max<int64_t>(mx, 1) * sizeof(T));
            int64_t mw = 0;
            for (const auto& lv : P.levels) {
                int64_t a = 0;
                for (int s : lv) a += (int64_t)P.fronts[s].m * P.fronts[s].n;
                mw = std::max(mw, a);
            }
            wbuf.alloc(std::max<int64_t>(mw, 1) * sizeof(T));
        }
        build_factor_schedule(farenas[0], fsched[0], false);
        fbuilt[0] = true;
        P.ensure_dep_events(nextEvent);
    }

    bool block() const { return P.frontType == 1 || P.frontType == 3; }
    bool pivoted() const { return P.frontType == 2 || P.frontType == 3; }

    size_t device_bytes() const override {
        size_t b = fpool.bytes + upool.bytes + dpool.bytes + linv.bytes + bsave.bytes +
                   bsinv.bytes + btmp.bytes + rbuf.bytes + identity.bytes + linvw.bytes +
                   wtmp.bytes + wbuf.bytes + subpool.bytes + pivpool.bytes;
        for (const Arena& a : farenas) b += a.dev.bytes;
        for (auto& kv : solves)
            b += kv.second->arena.dev.bytes + kv.second->wpool.bytes + kv.second->xn.bytes +
                 kv.second->xn2.bytes + kv.second->tmp.bytes + kv.second->splitWs.bytes;
        return b;
    }

    T* L(int s) const { return fpool.as<T>() + P.fronts[s].L; }
    T* U(int s) const { return upool.as<T>() + P.fronts[s].U; }
    V* D(int s) const { return dpool.as<V>() + P.fronts[s].off; }
    V* Sub(int s) const { return subpool.as<V>() + P.fronts[s].off; }
    int* Piv(int s) const { return pivpool.as<int>() + P.fronts[s].off; }
    T* Li(int s, int k0) const { return linv.as<T>() + P.fronts[s].Linv + (int64_t)(k0 / NB) * NB * NB; }
    // wide inverse block starting at column w0 (a multiple of SB), ld min(SB, n - w0)
    T* LW(int s, int w0) const {
        const int SB = P.sb;
        return linvw.as<T>() + P.fronts[s].LinvW + (int64_t)(w0 / SB) * SB * SB;
    }

    // Factor schedule. Block LDL (ldl/local_front_block.hpp:25-73) adds, per
    // level: save A21 after the extend-add; after the Schur update, the
    // unit-lower inverse X = L11^{-1} built block column by block column from
    // the last (X21_j = -(X22 L21_j) Linv_jj, X_jj = Linv_jj) in a separate
    // n x n buffer; ATL := X^T D^{-1} X (X^H for Hermitian) written to both
    // triangles; and A21 restored.
    void build_factor_schedule(Arena& arena, std::vector<Step>& steps, bool blockLdl,
                               bool pivot = false) {
        StepBuilder<T> sb{arena, steps};
        const int herm = P.conjugate ? G_CONJB : 0;
        std::vector<int64_t> saveOff(P.numFronts, 0), sinvOff(P.numFronts, 0),
            tmpOff(P.numFronts, 0);
        if (blockLdl) {
            int64_t maxSave = 0, maxSinv = 0, maxTmp = 0;
            for (int h = 0; h < P.numLevels; ++h) {
                int64_t a = 0, b = 0, c = 0;
                for (int s : P.levels[h]) {
                    const FrontLayout& f = P.fronts[s];
                    saveOff[s] = a; a += (int64_t)f.u * f.n;
                    sinvOff[s] = b; b += (int64_t)f.n * f.n;
                    tmpOff[s] = c; c += (int64_t)f.n * NB;
                }
                maxSave = std::max(maxSave, a);
                maxSinv = std::max(maxSinv, b);
                maxTmp = std::max(maxTmp, c);
            }
            bsave.alloc(std::max<int64_t>(maxSave, 1) * sizeof(T));
            bsinv.alloc(std::max<int64_t>(maxSinv, 1) * sizeof(T));
            btmp.alloc(std::max<int64_t>(maxTmp, 1) * sizeof(T));
        }
        auto Sv = [&](int s) { return bsave.as<T>() + saveOff[s]; };
        auto Xi = [&](int s) { return bsinv.as<T>() + sinvOff[s]; };
        auto Tp = [&](int s) { return btmp.as<T>() + tmpOff[s]; };
        for (int h = 0; h < P.numLevels; ++h) {
            const auto& lv = P.levels[h];
            // Extend-add in two passes (ldl/local.hpp:60-102 adds each child's
            // update matrix into the parent front in one): the child columns
            // that map into the parent's own columns (L part, a prefix since
            // the relative indices increase) before the panels, the rest
            // (update part) after the Schur update, which then writes the
            // update matrix instead of accumulating into a cleared one.
            auto extend = [&](bool upart) {
                for (int slot = 0; slot < 2; ++slot) {
                    std::vector<ExtendDesc<T>> ds;
                    std::vector<int> start;
                    int total = 0;
                    double by = 0.0;
                    for (int p : lv) {
                        const FrontLayout& fp = P.fronts[p];
                        if (fp.nchild <= slot) continue;
                        const int c = fp.child[slot];
                        const FrontLayout& fc = P.fronts[c];
                        if (fc.u == 0) continue;
                        const int* r = P.relHost.data() + fc.rel;
                        const int js = (int)(std::lower_bound(r, r + fc.u, fp.n) - r);
                        const int j0 = upart ? js : 0, nc = upart ? fc.u - js : js;
                        if (nc == 0) continue;
                        ds.push_back(ExtendDesc<T>{U(c), P.dRel.as<int>() + fc.rel, L(p), U(p),
                                                   fc.u, fp.n, fp.m, fp.u, j0, nc});
                        start.push_back(total);
                        total += (nc + 1) / 2;   // column pairs
                        // entries of columns j0..j0+nc-1 of a lower triangle of order uc
                        const double e = (double)nc * fc.u - ((double)(j0 + nc) * (j0 + nc - 1) -
                                                               (double)j0 * (j0 - 1)) / 2.0;
                        by += 3.0 * e * sizeof(T);
                    }
                    if (!ds.empty())
                        steps.push_back(Step{S_EXTEND, arena.add(ds), arena.add(start),
                                              (int)ds.size(), total, false, 0.0, by});
                }
            };
            {
                // update matrices of fronts without columns get no Schur
                // update: clear them (only their lower triangle is read)
                std::vector<ZeroLowerDesc<T>> zs;
                std::vector<int> start;
                int cols = 0;
                double by = 0.0;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    if (f.u == 0 || f.n > 0) continue;
                    zs.push_back(ZeroLowerDesc<T>{U(s), f.u, f.u, f.u});
                    start.push_back(cols);
                    cols += f.u;
                    by += 0.5 * f.u * (f.u + 1.0) * sizeof(T);
                }
                if (!zs.empty())
                    steps.push_back(Step{S_ZLOWER, arena.add(zs), arena.add(start), (int)zs.size(),
                                          cols, false, 0.0, by});
            }
            extend(false);
            if (blockLdl) {
                std::vector<CopyDesc<T>> cs;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    cs.push_back(CopyDesc<T>{L(s) + f.n, Sv(s), f.m, f.u, f.u, f.n});
                }
                sb.copy(cs);
            }
            int maxn = 0;
            for (int s : lv) maxn = std::max(maxn, P.fronts[s].n);
            int pendingB2 = -1;   // side-stream update the next panel's B1 must wait for
            std::map<int, int64_t> wOff;
            {
                int64_t a = 0;
                for (int s : lv) {
                    wOff[s] = a;
                    a += (int64_t)P.fronts[s].m * P.fronts[s].n;
                }
            }
            auto Wf = [&](int s) { return wbuf.as<T>() + wOff[s]; };
            // Two-level panels: 64-wide diagonal blocks whose updates stay
            // inside an OB-wide outer panel, then one rank-OB update of the
            // front's trailing columns (K = OB instead of K = 64 passes).
            const int OB = outer_panel();
            if (pivot) pivoted_level(sb, arena, steps, lv, Wf);
            for (int k0 = 0; k0 < (pivot ? 0 : maxn); k0 += OB) {
                for (int k = k0; k < std::min(maxn, k0 + OB); k += NB) {
                    std::vector<DiagDesc<T>> dd;
                    std::vector<GemmDesc<T>> trsm, upd;
                    for (int s : lv) {
                        const FrontLayout& f = P.fronts[s];
                        if (f.n <= k) continue;
                        const int b = std::min(NB, f.n - k);
                        const int c1 = std::min(f.n, k0 + OB);   // end of the outer panel
                        T* base = L(s);
                        dd.push_back(DiagDesc<T>{base + k + (int64_t)k * f.m, D(s) + k, Li(s, k), f.m,
                                                 b, s, k});
                        T* a21 = base + (k + b) + (int64_t)k * f.m;
                        T* w21 = Wf(s) + (k + b) + (int64_t)k * f.m;
                        if (f.m - k - b > 0) {
                            // L21 = W21 D^{-1} into the front, W21 = A21 Linv^T next to it
                            GemmDesc<T> t{a21, Li(s, k), a21, nullptr, D(s) + k, f.m, NB, f.m,
                                          f.m - k - b, b, b, G_TRANSB | herm};
                            t.C2 = w21;
                            trsm.push_back(t);
                        }
                        if (c1 - k - b > 0) {
                            T* a22 = base + (k + b) + (int64_t)(k + b) * f.m;
                            upd.push_back(GemmDesc<T>{w21, a21, a22, nullptr, nullptr, f.m, f.m, f.m,
                                                      f.m - k - b, c1 - k - b, b,
                                                      G_TRANSB | G_LOWER | G_ACCUM | G_NEG | herm});
                        }
                    }
                    if (!dd.empty()) {
                        double fl = 0.0;
                        for (const auto& x : dd) fl += (double)x.b * x.b * x.b / 3.0;
                        steps.push_back(Step{S_DIAG, arena.add(dd), 0, (int)dd.size(), 0, false,
                                             fl * (Scalar<T>::complex ? 4 : 1), 0.0});
                    }
                    sb.gemm(trsm, true);
                    sb.update(upd);
                }
                // Outer update with a lookahead of one panel: the next outer
                // panel's columns (B1) on the plan's stream, the columns after
                // it (B2) on the side stream, overlapping the next panel's
                // diagonal blocks and TRSMs. B2 of this panel must finish
                // before B1 of the next one writes the same columns.
                std::vector<GemmDesc<T>> b1, b2;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    const int c1 = std::min(f.n, k0 + OB);
                    if (f.n <= k0 || f.n - c1 <= 0) continue;
                    const int c2 = std::min(f.n, c1 + OB);
                    T* l1 = L(s) + c1 + (int64_t)k0 * f.m;
                    T* w1 = Wf(s) + c1 + (int64_t)k0 * f.m;
                    b1.push_back(GemmDesc<T>{w1, l1, L(s) + c1 + (int64_t)c1 * f.m, nullptr, nullptr,
                                             f.m, f.m, f.m, f.m - c1, c2 - c1, c1 - k0,
                                             G_TRANSB | G_LOWER | G_ACCUM | G_NEG | herm});
                    if (f.n - c2 > 0) {
                        T* l2 = L(s) + c2 + (int64_t)k0 * f.m;
                        T* w2 = Wf(s) + c2 + (int64_t)k0 * f.m;
                        b2.push_back(GemmDesc<T>{w2, l2, L(s) + c2 + (int64_t)c2 * f.m, nullptr,
                                                 nullptr, f.m, f.m, f.m, f.m - c2, f.n - c2, c1 - k0,
                                                 G_TRANSB | G_LOWER | G_ACCUM | G_NEG | herm});
                    }
                }
                int evB2 = -1;
                if (!b2.empty() && !P.lookahead) {
                    sb.update(b2);
                } else if (!b2.empty()) {
                    // B2 needs this panel's final L columns, and runs after
                    // the previous B2 on the side stream
                    const int evL = nextEvent++;
                    evB2 = nextEvent++;
                    sb.record(evL);
                    sb.on(1);
                    sb.wait(evL);
                    sb.update(b2);
                    sb.record(evB2);
                    sb.on(0);
                }
                if (pendingB2 >= 0) sb.wait(pendingB2);
                sb.update(b1);
                pendingB2 = evB2;
            }
            if (pendingB2 >= 0) {
                sb.wait(pendingB2);
                pendingB2 = -1;
            }
            std::vector<GemmDesc<T>> syrk;
            for (int s : lv) {
                const FrontLayout& f = P.fronts[s];
                if (f.u == 0 || f.n == 0) continue;
                T* l21 = L(s) + f.n;
                syrk.push_back(GemmDesc<T>{Wf(s) + f.n, l21, U(s), nullptr, nullptr, f.m, f.m, f.u,
                                           f.u, f.u, f.n, G_TRANSB | G_LOWER | G_NEG | herm});
            }
            sb.update(syrk);
            extend(true);
            if (blockLdl) block_inverse(sb, lv, Sv, Xi, Tp, pivot, Wf);
            else wide_inverse(sb, lv);
        }
        arena.finish();
    }

    // Panel width of the pivoted LDL^T of ATL (CLIQG_BK_PANEL, 8..64): the
    // panel's delayed updates are a matrix-vector product per column inside
    // one workgroup, the trailing updates MFMA launches of rank width + 1.
    static int bk_panel() {
        static const int v = [] {
            const char* e = std::getenv("CLIQG_BK_PANEL");
            const int t = e ? std::atoi(e) : 32;
            return (t >= 8 && t <= BK_MAX_PANEL) ? t : 32;
        }();
        return v;
    }

    // FrontLDLIntraPiv (ldl/local_front.hpp:73-94) for the fronts of one
    // level, between the extend-add of the children's L parts and the Schur
    // update: the pivoted LDL^T of ATL panel by panel (k_bk_panel, then the
    // rank nbp + 1 update of ATL's trailing lower triangle), the inverses of
    // L11's 64-wide diagonal blocks, ABL's columns permuted into the W buffer
    // (El::PermuteCols), W21 := ABL L11^{-T} (^{-H} if Hermitian) by 64-wide
    // blocks inside OB-wide outer panels as in the unpivoted path, and L21 :=
    // W21 D^{-1} with 2x2 pivots (El::QuasiDiagonalSolve). The Schur update
    // ABR -= W21 L21^T that follows is the unpivoted path's.
    template <class FW>
    void pivoted_level(StepBuilder<T>& sb, Arena& arena, std::vector<Step>& steps,
                       const std::vector<int>& lv, FW Wf) {
        const int herm = P.conjugate ? G_CONJB : 0;
        const int nbp = bk_panel();
        int maxn = 0;
        for (int s : lv) maxn = std::max(maxn, P.fronts[s].n);
        for (int k0 = 0; k0 < maxn; k0 += nbp) {
            std::vector<BkDesc<T>> bd;
            std::vector<GemmDesc<T>> trail;
            double by = 0.0;
            for (int s : lv) {
                const FrontLayout& f = P.fronts[s];
                if (f.n <= k0) continue;
                bd.push_back(BkDesc<T>{L(s), Wf(s), D(s), Sub(s), Piv(s), f.m, f.n, k0, s});
                by += 3.0 * (double)(f.n - k0) * nbp * sizeof(T);
                const int e = k0 + nbp;
                if (e + 1 < f.n)
                    trail.push_back(GemmDesc<T>{Wf(s) + (e + 1) + (int64_t)k0 * f.m,
                                                L(s) + (e + 1) + (int64_t)k0 * f.m,
                                                L(s) + (e + 1) + (int64_t)(e + 1) * f.m, nullptr,
                                                nullptr, f.m, f.m, f.m, f.n - e - 1, f.n - e - 1,
                                                e - k0 + 1,
                                                G_TRANSB | G_LOWER | G_ACCUM | G_NEG | herm});
            }
            if (bd.empty()) continue;
            Step st{S_BK, arena.add(bd), 0, (int)bd.size(), 0, false, 0.0, by};
            st.variant = nbp;
            steps.push_back(st);
            sb.update(trail);
        }
        std::vector<TriInvDesc<T>> ti;
        std::vector<GatherColsDesc<T>> gc;
        std::vector<QuasiDesc<T>> qd;
        std::vector<int> gstart, qstart;
        int gblocks = 0;
        double gby = 0.0;
        for (int s : lv) {
            const FrontLayout& f = P.fronts[s];
            for (int k = 0; k < f.n; k += NB)
                ti.push_back(TriInvDesc<T>{L(s) + k + (int64_t)k * f.m, Li(s, k), f.m,
                                           std::min(NB, f.n - k)});
            if (f.u == 0 || f.n == 0) continue;
            gc.push_back(GatherColsDesc<T>{L(s) + f.n, Wf(s) + f.n, Piv(s), f.m, f.m, f.u, f.n});
            qd.push_back(QuasiDesc<T>{Wf(s) + f.n, L(s) + f.n, D(s), Sub(s), f.m, f.u, f.n});
            gstart.push_back(gblocks);
            gblocks += f.n * ((f.u + 255) / 256);
            gby += 2.0 * (double)f.u * f.n * sizeof(T);
        }
        if (!ti.empty())
            steps.push_back(Step{S_TRINV, arena.add(ti), 0, (int)ti.size(), 0, false, 0.0,
                                 2.0 * ti.size() * NB * NB * sizeof(T)});
        if (gc.empty()) return;
        const size_t gstartOff = arena.add(gstart);
        steps.push_back(Step{S_GCOLS, arena.add(gc), gstartOff, (int)gc.size(), gblocks, false, 0.0,
                             gby});
        // W21 := W21 L11^{-T}, 64-wide blocks inside OB-wide outer panels
        const int OB = outer_panel();
        for (int k0 = 0; k0 < maxn; k0 += OB) {
            for (int k = k0; k < std::min(maxn, k0 + OB); k += NB) {
                std::vector<GemmDesc<T>> trsm, upd;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    if (f.n <= k || f.u == 0) continue;
                    const int b = std::min(NB, f.n - k);
                    const int c1 = std::min(f.n, k0 + OB);
                    T* wb = Wf(s) + f.n;
                    trsm.push_back(GemmDesc<T>{wb + (int64_t)k * f.m, Li(s, k), wb + (int64_t)k * f.m,
                                               nullptr, nullptr, f.m, NB, f.m, f.u, b, b,
                                               G_TRANSB | herm});
                    if (c1 - k - b > 0)
                        upd.push_back(GemmDesc<T>{wb + (int64_t)k * f.m,
                                                  L(s) + (k + b) + (int64_t)k * f.m,
                                                  wb + (int64_t)(k + b) * f.m, nullptr, nullptr, f.m,
                                                  f.m, f.m, f.u, c1 - k - b, b,
                                                  G_TRANSB | G_ACCUM | G_NEG | herm});
                }
                sb.gemm(trsm, true);
                sb.update(upd);
            }
            std::vector<GemmDesc<T>> outer;
            for (int s : lv) {
                const FrontLayout& f = P.fronts[s];
                const int c1 = std::min(f.n, k0 + OB);
                if (f.n <= k0 || f.n - c1 <= 0 || f.u == 0) continue;
                T* wb = Wf(s) + f.n;
                outer.push_back(GemmDesc<T>{wb + (int64_t)k0 * f.m, L(s) + c1 + (int64_t)k0 * f.m,
                                            wb + (int64_t)c1 * f.m, nullptr, nullptr, f.m, f.m, f.m,
                                            f.u, f.n - c1, c1 - k0,
                                            G_TRANSB | G_ACCUM | G_NEG | herm});
            }
            sb.update(outer);
        }
        const size_t qstartOff = arena.add(gstart);
        steps.push_back(Step{S_QSCALE, arena.add(qd), qstartOff, (int)qd.size(), gblocks, false, 0.0,
                             gby});
    }

    // Inverses of the SB-wide diagonal blocks of each L11 (used by the LDL
    // solves), from the 64-wide ones: inside a wide block, sub-blocks from the
    // last, X_jj = Linv_jj and X21_j = -(X22 L21_j) Linv_jj.
    void wide_inverse(StepBuilder<T>& sb, const std::vector<int>& lv) {
        const int SB = P.sb;
        struct WB { int s, w0, sbw; int64_t tmp; };
        std::vector<WB> wbs;
        std::vector<ZeroSeg<T>> zs;
        // fronts of at most min(SB, 2 NB) columns: their one wide block in
        // one workgroup each (launch_wide_inverse_small)
        std::vector<WideInvDesc<T>> small;
        int64_t tmpTotal = 0;
        for (int s : lv) {
            const FrontLayout& f = P.fronts[s];
            if (f.n > 0 && f.n <= std::min(SB, 2 * NB)) {
                small.push_back(WideInvDesc<T>{Li(s, 0), f.n > NB ? Li(s, NB) : Li(s, 0),
                                               L(s) + NB, LW(s, 0), f.m, f.n});
                continue;
            }
            int64_t entries = 0;
            for (int w0 = 0; w0 < f.n; w0 += SB) {
                const int sbw = std::min(SB, f.n - w0);
                wbs.push_back(WB{s, w0, sbw, tmpTotal});
                tmpTotal += (int64_t)sbw * NB;
                entries += (int64_t)sbw * sbw;
            }
            if (entries) zs.push_back(ZeroSeg<T>{LW(s, 0), entries});
        }
        if (!small.empty())
            sb.steps.push_back(Step{S_WINV, sb.arena.add(small), 0, (int)small.size(), 0, false});
        if (tmpTotal * (int64_t)sizeof(T) > (int64_t)wtmp.bytes)
            throw Error(5, "wide inverse workspace smaller than a level needs");
        sb.zero(zs);
        for (int t = 0; t < SB / NB; ++t) {
            std::vector<CopyDesc<T>> diagBlocks;
            std::vector<GemmDesc<T>> trmm, scal;
            for (const WB& w : wbs) {
                const FrontLayout& f = P.fronts[w.s];
                const int q = (w.sbw + NB - 1) / NB;
                if (q <= t) continue;
                const int lo = (q - 1 - t) * NB, j0 = w.w0 + lo;
                const int b = std::min(NB, w.sbw - lo), r = w.sbw - lo - b;
                const int64_t ld = w.sbw;
                T* X = LW(w.s, w.w0);
                diagBlocks.push_back(CopyDesc<T>{Li(w.s, j0), X + lo + lo * ld, NB, ld, b, b});
                if (r > 0) {
                    T* tp = wtmp.as<T>() + w.tmp;
                    trmm.push_back(GemmDesc<T>{X + (lo + b) + (lo + b) * ld,
                                               L(w.s) + (j0 + b) + (int64_t)j0 * f.m, tp, nullptr,
                                               nullptr, ld, f.m, ld, r, b, r, G_KHI});
                    scal.push_back(GemmDesc<T>{tp, Li(w.s, j0), X + (lo + b) + lo * ld, nullptr,
                                               nullptr, ld, NB, ld, r, b, b, G_NEG});
                }
            }
            sb.copy(diagBlocks);
            sb.gemm(trmm, true);
            sb.gemm(scal, true);
        }
    }

    // With pivoting (FrontBlockLDL's intraPiv branch, ldl/local_front_block.hpp:
    // 35-57) the last product is ATL := P^T (X^T D^{-1} X) P with D's 2x2
    // blocks: Z = D^{-1} X into the dead L11, Y = op(X)^T Z into the W
    // buffer, then ATL(piv[i], piv[j]) = Y(i, j).
    template <class FS, class FX, class FT, class FW>
    void block_inverse(StepBuilder<T>& sb, const std::vector<int>& lv, FS Sv, FX Xi, FT Tp,
                       bool pivot, FW Wf) {
        const int cA = P.conjugate ? G_CONJA : 0;
        std::vector<ZeroSeg<T>> zs;
        int maxq = 0;
        for (int s : lv) {
            const FrontLayout& f = P.fronts[s];
            zs.push_back(ZeroSeg<T>{Xi(s), (int64_t)f.n * f.n});
            maxq = std::max(maxq, (f.n + NB - 1) / NB);
        }
        sb.zero(zs);
        for (int t = 0; t < maxq; ++t) {
            std::vector<CopyDesc<T>> diagBlocks;
            std::vector<GemmDesc<T>> trmm, scal;
            for (int s : lv) {
                const FrontLayout& f = P.fronts[s];
                const int q = (f.n + NB - 1) / NB;
                if (q <= t) continue;
                const int j0 = (q - 1 - t) * NB, b = std::min(NB, f.n - j0), r = f.n - j0 - b;
                const int64_t n = f.n;
                T* X = Xi(s);
                diagBlocks.push_back(CopyDesc<T>{Li(s, j0), X + j0 + j0 * n, NB, n, b, b});
                if (r > 0) {
                    T* x22 = X + (j0 + b) + (j0 + b) * n;
                    const T* l21 = L(s) + (j0 + b) + (int64_t)j0 * f.m;
                    trmm.push_back(GemmDesc<T>{x22, l21, Tp(s), nullptr, nullptr, n, f.m, n, r, b,
                                               r, G_KHI});
                    scal.push_back(GemmDesc<T>{Tp(s), Li(s, j0), X + (j0 + b) + j0 * n, nullptr,
                                               nullptr, n, NB, n, r, b, b, G_NEG});
                }
            }
            sb.copy(diagBlocks);
            sb.gemm(trmm, true);
            sb.gemm(scal, true);
        }
        std::vector<GemmDesc<T>> inv;
        std::vector<CopyDesc<T>> restore;
        std::vector<QuasiRowsDesc<T>> qr;
        std::vector<PermSymDesc<T>> ps;
        std::vector<int> qstart;
        int qblocks = 0;
        for (int s : lv) {
            const FrontLayout& f = P.fronts[s];
            if (f.n == 0) continue;
            if (pivot) {
                qr.push_back(QuasiRowsDesc<T>{Xi(s), L(s), D(s), Sub(s), f.n, f.m, f.n});
                inv.push_back(GemmDesc<T>{Xi(s), L(s), Wf(s), nullptr, nullptr, f.n, f.m, f.m, f.n,
                                          f.n, f.n, G_TRANSA | G_KLO | cA});
                ps.push_back(PermSymDesc<T>{Wf(s), L(s), Piv(s), f.m, f.m, f.n});
                qstart.push_back(qblocks);
                qblocks += f.n * ((f.n + 255) / 256);
            } else {
                inv.push_back(GemmDesc<T>{Xi(s), Xi(s), L(s), D(s), nullptr, f.n, f.n, f.m, f.n,
                                          f.n, f.n,
                                          G_TRANSA | G_LOWER | G_SYMM | G_INVK | G_KLO | cA});
            }
            restore.push_back(CopyDesc<T>{Sv(s), L(s) + f.n, f.u, f.m, f.u, f.n});
        }
        if (pivot && !qr.empty()) {
            const size_t st = sb.arena.add(qstart);
            sb.steps.push_back(Step{S_QROWS, sb.arena.add(qr), st, (int)qr.size(), qblocks, false});
            sb.gemm(inv, true);
            sb.steps.push_back(Step{S_PSYM, sb.arena.add(ps), st, (int)ps.size(), qblocks, false});
        } else {
            sb.gemm(inv, true);
        }
        sb.copy(restore);
    }

    // Split K of tall, skinny products with a long K (the update-row products
    // of the solves, M x k with K = u or n): a launch with fewer than
    // split_tiles() 64x64 tiles (2,048; CLIQG_SPLIT_TILES) leaves CUs idle or
    // unevenly loaded while each tile walks K in 16-deep steps, so every plain
    // descriptor with K >= 2 SPLIT_MIN_K is cut into up to
    // 2 split_tiles() / tiles slices, written to ws and summed in slice order
    // by a second launch (measured on the 120^3 PML solves: no split 209 ms
    // of solve products per five solves, 512: 200, 1,024: 195, 2,048: 192).
    // Returns the workspace entries used; with sb == nullptr nothing is
    // scheduled.
    static double split_tiles() {
        static const double v = [] {
            const char* e = std::getenv("CLIQG_SPLIT_TILES");
            const double t = e ? std::atof(e) : 2048.0;
            return (t >= 0.0 && t <= 8192.0) ? t : 2048.0;
        }();
        return v;
    }
    static constexpr int SPLIT_MIN_K = 512;
    static int64_t split_bound() { return (int64_t)4096 * (int64_t)(3.0 * split_tiles() + 1.0); }

    int64_t split_gemm(StepBuilder<T>* sb, const std::vector<GemmDesc<T>>& g, T* ws) {
        double tiles = 0.0;
        for (const auto& d : g) tiles += gemm_tiles(d.M, d.N, d.flags, TILE_64);
        const double lim = split_tiles();
        const int want = (tiles > 0.0 && tiles < lim) ? (int)std::ceil(2.0 * lim / tiles) : 1;
        std::vector<GemmDesc<T>> parts;
        std::vector<ReduceDesc<T>> red;
        int64_t off = 0;
        for (const auto& d : g) {
            const bool plain = !(d.flags & (G_KLO | G_KHI | G_LOWER | G_SYMM | G_INVK)) &&
                               !d.scaleK && !d.scaleN && !d.C2 && d.C != d.A && d.C != d.B;
            int S = 1;
            if (want > 1 && plain && d.K >= 2 * SPLIT_MIN_K) S = std::min(want, d.K / SPLIT_MIN_K);
            if (S <= 1) {
                parts.push_back(d);
                continue;
            }
            const int kc = (d.K + S - 1) / S;
            S = (d.K + kc - 1) / kc;
            for (int sl = 0; sl < S; ++sl) {
                const int k0 = sl * kc;
                GemmDesc<T> p = d;
                p.A = (d.flags & G_TRANSA) ? d.A + k0 : d.A + (int64_t)k0 * d.lda;
                p.B = (d.flags & G_TRANSB) ? d.B + (int64_t)k0 * d.ldb : d.B + k0;
                p.C = ws + off + (int64_t)sl * d.M * d.N;
                p.ldc = d.M;
                p.K = std::min(kc, d.K - k0);
                p.flags = d.flags & (G_TRANSA | G_TRANSB | G_CONJA | G_CONJB);
                parts.push_back(p);
            }
            red.push_back(ReduceDesc<T>{ws + off, d.C, d.ldc, d.M, d.N, S,
                                        d.flags & (G_ACCUM | G_NEG)});
            off += (int64_t)S * d.M * d.N;
        }
        if (off > split_bound()) throw Error(5, "split-K workspace smaller than a launch needs");
        if (sb) {
            sb->gemm(parts);
            sb->reduce(red);
        }
        return off;
    }

    // Solve schedule for k right-hand sides. The front's own rows are the
    // nodal right-hand sides X (ld N); W_s (u x k, ld u) holds its update rows.
    // Forward, per level: clear W, extend-add child slot 0 then 1, then per
    // 64-wide panel X_k := Linv_k X_k and X(k+b:n) -= L(k+b:n, k) X_k, then
    // W -= L21 X_s in one product.
    // Backward, per level from the root: W := X(lowerStruct), X_s -= L21^T W,
    // then per panel from the last X_k := Linv_k^T X_k and
    // X(0:k) -= L(k:k+b, 0:k)^T X_k.
    // per-level n x k workspace of the block solves (entries)
    int64_t solve_tmp_entries(int k) const {
        if (!block()) return 0;
        int64_t mx = 0;
        for (int h = 0; h < P.numLevels; ++h) {
            int64_t a = 0;
            for (int s : P.levels[h]) a += (int64_t)P.fronts[s].n * k;
            mx = std::max(mx, a);
        }
        return mx;
    }

    SolveSched& schedule(int k) {
        const bool blk = block();
        const int key = 4 * k + P.frontType;
        auto it = solves.find(key);
        if (it != solves.end()) return *it->second;
        auto sp = std::make_unique<SolveSched>();
        SolveSched& S = *sp;
        S.wpool.alloc(std::max<int64_t>(P.wRows * k, 1) * sizeof(T));
        S.xn.alloc(std::max<int64_t>((int64_t)P.N * k, 1) * sizeof(T));
        // LDL sweeps ping-pong between xn and xn2: the forward sweep takes its
        // input in xn and leaves X in xn2, the backward sweep the reverse
        if (!blk) S.xn2.alloc(std::max<int64_t>((int64_t)P.N * k, 1) * sizeof(T));
        S.splitWs.alloc(split_bound() * sizeof(T));
        S.tmp.alloc(std::max<int64_t>(solve_tmp_entries(k), 1) * sizeof(T));
        S.wp = S.wpool.template as<T>();
        S.x1 = S.xn.template as<T>();
        S.x2 = blk ? S.x1 : S.xn2.template as<T>();
        S.status = P.dStatus.as<int>();
        build_solve(S, k, S.tmp.template as<T>(), S.splitWs.template as<T>());
        S.arena.finish();
        SolveSched& ref = S;
        solves.emplace(key, std::move(sp));
        return ref;
    }

    // Schedule of the allocation-free solve for k right-hand sides: every
    // buffer inside the scratch array at `base` (cap bytes), laid out as
    // [descriptor arena | status | x1 | x2 | W pool | block workspace |
    // split-K workspace]. Built twice: once to learn the arena's size and
    // the split workspace the launches need, once with the final addresses.
    SolveSched& noalloc_schedule(int k, char* base, size_t cap) {
        const bool blk = block();
        const int key = 4 * k + P.frontType;
        auto it = noallocSolves.find(key);
        if (it != noallocSolves.end()) return *it->second;
        auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t nk = (size_t)P.N * k * sizeof(T);
        auto place = [&](SolveSched& S, char* b, size_t arenaBytes, int64_t wsEntries,
                         T** tp, T** ws) {
            size_t off = al(arenaBytes);
            S.status = reinterpret_cast<int*>(b + off);
            off += 256;
            S.x1 = reinterpret_cast<T*>(b + off);
            off += al(nk);
            S.x2 = S.x1;
            if (!blk) {
                S.x2 = reinterpret_cast<T*>(b + off);
                off += al(nk);
            }
            S.wp = reinterpret_cast<T*>(b + off);
            off += al((size_t)P.wRows * k * sizeof(T));
            *tp = reinterpret_cast<T*>(b + off);
            off += al((size_t)solve_tmp_entries(k) * sizeof(T));
            *ws = reinterpret_cast<T*>(b + off);
            off += (size_t)wsEntries * sizeof(T);
            return off;
        };
        T *tp = nullptr, *ws = nullptr;
        size_t arenaBytes = 0;
        int64_t wsEntries = 0;
        {
            SolveSched probe;
            place(probe, base, 0, 0, &tp, &ws);
            build_solve(probe, k, tp, ws);
            arenaBytes = probe.arena.host.size();
            wsEntries = probe.wsUsed;
        }
        auto sp = std::make_unique<SolveSched>();
        SolveSched& S = *sp;
        const size_t need = place(S, base, arenaBytes, wsEntries, &tp, &ws);
        if (need > cap)
            throw Error(1, "the allocation-free solve needs " + std::to_string(need >> 20) +
                               " MiB of scratch for this plan and " + std::to_string(k) +
                               " right-hand sides; its scratch array holds " +
                               std::to_string(cap >> 20) + " MiB");
        build_solve(S, k, tp, ws);
        if (S.arena.host.size() != arenaBytes || S.wsUsed > wsEntries)
            throw Error(5, "allocation-free solve schedule changed between its two builds");
        S.arena.base = base;
        SolveSched& ref = S;
        noallocSolves.emplace(key, std::move(sp));
        return ref;
    }

    // Solve schedule for k right-hand sides into S's workspace (tp: block
    // workspace, ws: split-K workspace). The front's own rows are the
    // nodal right-hand sides X (ld N); W_s (u x k, ld u) holds its update rows.
    // Forward, per level: clear W, extend-add child slot 0 then 1, then per
    // SB-wide panel X_k := Linv_k X_k and X(k+b:n) -= L(k+b:n, k) X_k, then
    // W -= L21 X_s in one product.
    // Backward, per level from the root: W := X(lowerStruct), X_s -= L21^T W,
    // then per panel from the last X_k := Linv_k^T X_k and
    // X(0:k) -= L(k:k+b, 0:k)^T X_k.
    void build_solve(SolveSched& S, int k, T* tp, T* ws) {
        const bool blk = block();
        const int SB = P.sb;
        // per-front n x k workspace of the block solves
        std::vector<int64_t> tmpOff(P.numFronts, 0);
        for (int h = 0; h < P.numLevels; ++h) {
            int64_t a = 0;
            for (int s : P.levels[h]) {
                tmpOff[s] = a;
                a += (int64_t)(blk ? P.fronts[s].n : 0) * k;
            }
        }
        T* wp = S.wp;
        T* xn = S.x1;
        T* xn2 = S.x2;
        auto Tm = [&](int s) { return tp + tmpOff[s]; };
        const int64_t N = P.N;
        auto Wp = [&](int s) { return wp + P.fronts[s].W * k; };
        auto Xs = [&](int s) { return xn + P.fronts[s].off; };
        auto X2 = [&](int s) { return xn2 + P.fronts[s].off; };
        auto split = [&](StepBuilder<T>& sbl, const std::vector<GemmDesc<T>>& g) {
            S.wsUsed = std::max(S.wsUsed, split_gemm(&sbl, g, ws));
        };
        const int cA = P.conjugate ? G_CONJA : 0;
        const int chainQ = (SB <= PANEL_MAX) ? solve_chain() * SB : 0;
        for (int pass = 0; pass < 2; ++pass) {
            const bool bwd = pass == 1;
            std::vector<Step>& steps = bwd ? S.bwd : S.fwd;
            StepBuilder<T> sb{S.arena, steps};
            sb.solve = true;
            // Fused panel steps pay where a level's chain of panels is
            // latency-bound: a few wide fronts, at most about one workgroup
            // per CU in the first step (at most 512 workgroups; every workgroup forms the panel's
            // product again; on wider levels that work outweighs the saved
            // launch)
            auto chain_level = [&](const std::vector<int>& lv, int maxn) {
                if (chainQ <= 0 || maxn <= SB) return false;
                const int cg = (k + 15) / 16;
                long tasks = 0;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    if (f.n == 0) continue;
                    const int sbw = std::min(SB, f.n);
                    tasks += (long)cg * (1 + (std::min(chainQ, f.n) - sbw + 63) / 64);
                }
                return tasks <= CHAIN_MAX_TASKS;
            };
            // one fused panel step of the fronts of level lv wider than w0
            // (launch_panel_step): the panel's rows and the band rows of
            // super-panel q0
            auto panel_step = [&](const std::vector<int>& lv, int w0, int q0, bool back) {
                std::vector<PanelStepDesc<T>> ds;
                std::vector<int> st;
                int tasks = 0;
                double fl = 0.0;
                const int cg = (k + 15) / 16;
                const double cm = Scalar<T>::complex ? 4.0 : 1.0;
                for (int s : lv) {
                    const FrontLayout& f = P.fronts[s];
                    if (f.n <= w0) continue;
                    const int sbw = std::min(SB, f.n - w0);
                    const int b0 = back ? q0 : w0 + sbw;
                    const int b1 = back ? w0 : std::min(q0 + chainQ, f.n);
                    const int band = b1 > b0 ? b1 - b0 : 0;
                    ds.push_back(PanelStepDesc<T>{LW(s, w0),
                                                  back ? L(s) + w0 : L(s) + (int64_t)w0 * f.m,
                                                  back ? X2(s) : Xs(s), back ? Xs(s) : X2(s),
                                                  N, f.m, w0, sbw, band ? b0 : 0, band ? b1 : 0,
                                                  k});
                    st.push_back(tasks);
                    tasks += cg * (1 + (band + 63) / 64);
                    fl += cm * 2.0 * k * ((double)sbw * (sbw + 1) / 2 + (double)band * sbw);
                }
                if (ds.empty()) return;
                Step x{S_PSTEP, S.arena.add(ds), S.arena.add(st), (int)ds.size(), tasks, back,
                       fl};
                steps.push_back(x);
            };
            for (int hh = 0; hh < P.numLevels; ++hh) {
                const int h = bwd ? P.numLevels - 1 - hh : hh;
                const auto& lv = P.levels[h];
                int maxn = 0;
                for (int s : lv) maxn = std::max(maxn, P.fronts[s].n);
                // Forward extend-add in two passes (lower_solve/local.hpp:86-104
                // adds a child's update rows into the parent in one): the rows
                // that land in the parent's own rows (a prefix: the relative
                // indices increase) before the parent's panels, the rest after
                // its update-row product, which then writes W_s instead of
                // accumulating into a cleared one.
                auto sextend = [&](bool wpart) {
                    for (int slot = 0; slot < 2; ++slot) {
                        std::vector<SolveExtend> ds;
                        std::vector<int> st;
                        int tot = 0;
                        for (int p : lv) {
                            const FrontLayout& fp = P.fronts[p];
                            if (fp.nchild <= slot) continue;
                            const FrontLayout& fc = P.fronts[fp.child[slot]];
                            if (fc.u == 0) continue;
                            const int* r = P.relHost.data() + fc.rel;
                            const int is = (int)(std::lower_bound(r, r + fc.u, fp.n) - r);
                            const int i0 = wpart ? is : 0, nr = wpart ? fc.u - is : is;
                            if (nr == 0) continue;
                            ds.push_back(SolveExtend{fc.W, fp.W, fc.rel, fc.u, fp.n, fp.off, fp.u,
                                                     i0});
                            st.push_back(tot);
                            tot += nr;
                        }
                        if (!ds.empty())
                            steps.push_back(Step{S_SEXTEND, S.arena.add(ds), S.arena.add(st),
                                                 (int)ds.size(), tot, false});
                    }
                };
                // PermuteRows of the fronts' own rows before the forward solve
                // and InversePermuteRows after the backward solve
                // (lower_solve/local_front.hpp:53-62, 84-93): the rows go to
                // the second nodal buffer permuted and are copied back
                auto permute = [&](bool scatter) {
                    std::vector<PermSeg> ps;
                    std::vector<int> st;
                    std::vector<CopyDesc<T>> back;
                    int tot = 0;
                    for (int s : lv) {
                        const FrontLayout& f = P.fronts[s];
                        if (f.n == 0) continue;
                        ps.push_back(PermSeg{f.off, f.n});
                        st.push_back(tot);
                        tot += f.n;
                        back.push_back(CopyDesc<T>{X2(s), Xs(s), N, N, f.n, k});
                    }
                    if (ps.empty()) return;
                    Step x{S_PERM, S.arena.add(ps), S.arena.add(st), (int)ps.size(), tot, bwd};
                    x.variant = scatter ? 1 : 0;
                    x.bytes = 2.0 * tot * k * sizeof(T);
                    steps.push_back(x);
                    sb.copy(back);
                };
                if (!bwd) {
                    // W_s of fronts without columns gets no update-row product
                    std::vector<ZeroSeg<T>> zs;
                    std::vector<int> zst;
                    int chunks = 0;
                    for (int s : lv) {
                        const int64_t n = (int64_t)P.fronts[s].u * k;
                        if (n == 0 || P.fronts[s].n > 0) continue;
                        zs.push_back(ZeroSeg<T>{Wp(s), n});
                        zst.push_back(chunks);
                        chunks += (int)((n + ZERO_CHUNK - 1) / ZERO_CHUNK);
                    }
                    if (!zs.empty())
                        steps.push_back(Step{S_ZERO, S.arena.add(zs), S.arena.add(zst),
                                             (int)zs.size(), chunks, false});
                    sextend(false);
                    if (pivoted() && !blk) permute(false);
                    if (blk) {
                        // FrontBlockLowerForwardSolve (lower_solve/local_front_block.hpp:28-48):
                        // X_s := inv(ATL) X_s through a copy of X_s
                        std::vector<CopyDesc<T>> cs;
                        std::vector<GemmDesc<T>> g;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.n == 0) continue;
                            cs.push_back(CopyDesc<T>{Xs(s), Tm(s), N, f.n, f.n, k});
                            g.push_back(GemmDesc<T>{L(s), Tm(s), Xs(s), nullptr, nullptr, f.m, f.n,
                                                    N, f.n, k, f.n, 0});
                        }
                        sb.copy(cs);
                        split(sb, g);
                    }
                    // Fused panel steps for levels with fronts wider than one
                    // panel: per super-panel of Q columns, one launch per panel
                    // (X2_p := inv(L_pp) X_p, band rows of the super-panel
                    // X(r) -= L(r, p) X2_p), then one product for the rows past
                    // the super-panel with K = Q
                    const bool fused = !blk && chain_level(lv, maxn);
                    for (int q0 = 0; fused && q0 < maxn; q0 += chainQ) {
                        for (int w0 = q0; w0 < std::min(q0 + chainQ, maxn); w0 += SB)
                            panel_step(lv, w0, q0, false);
                        std::vector<GemmDesc<T>> bulk;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.n <= q0 + chainQ) continue;
                            bulk.push_back(GemmDesc<T>{L(s) + (q0 + chainQ) + (int64_t)q0 * f.m,
                                                       X2(s) + q0, Xs(s) + q0 + chainQ, nullptr,
                                                       nullptr, f.m, N, N, f.n - q0 - chainQ, k,
                                                       chainQ, G_ACCUM | G_NEG});
                        }
                        split(sb, bulk);
                    }
                    // SB-wide panels: X2_p := inv(L_pp) X_p (into the second
                    // buffer), X(below) -= L(below, p) X2_p
                    for (int w0 = 0; w0 < ((blk || fused) ? 0 : maxn); w0 += SB) {
                        std::vector<GemmDesc<T>> tr, up;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.n <= w0) continue;
                            const int sbw = std::min(SB, f.n - w0);
                            T* x = Xs(s) + w0;
                            T* y = X2(s) + w0;
                            tr.push_back(GemmDesc<T>{LW(s, w0), x, y, nullptr, nullptr, sbw, N, N,
                                                     sbw, k, sbw, G_KHI});
                            if (f.n - w0 - sbw > 0)
                                up.push_back(GemmDesc<T>{L(s) + (w0 + sbw) + (int64_t)w0 * f.m, y,
                                                         x + sbw, nullptr, nullptr, f.m, N, N,
                                                         f.n - w0 - sbw, k, sbw, G_ACCUM | G_NEG});
                        }
                        sb.gemm(tr, true);
                        sb.gemm(up);
                    }
                    // the front's update rows in one product with K = n
                    std::vector<GemmDesc<T>> g;
                    for (int s : lv) {
                        const FrontLayout& f = P.fronts[s];
                        if (f.u == 0 || f.n == 0) continue;
                        g.push_back(GemmDesc<T>{L(s) + f.n, X2(s), Wp(s), nullptr, nullptr, f.m, N,
                                                f.u, f.u, k, f.n, G_NEG});
                    }
                    split(sb, g);
                    sextend(true);
                } else {
                    std::vector<SolveSeg> segs;
                    std::vector<int> rs;
                    int rows = 0;
                    std::vector<GemmDesc<T>> g;
                    for (int s : lv) {
                        const FrontLayout& f = P.fronts[s];
                        if (f.u == 0) continue;
                        segs.push_back(SolveSeg{f.W, f.low, f.u, 0});
                        rs.push_back(rows);
                        rows += f.u;
                        if (f.n > 0)
                            g.push_back(GemmDesc<T>{L(s) + f.n, Wp(s), X2(s), nullptr, nullptr, f.m,
                                                    f.u, N, f.n, k, f.u,
                                                    G_TRANSA | G_ACCUM | G_NEG | cA});
                    }
                    if (!segs.empty())
                        steps.push_back(Step{S_GATHER, S.arena.add(segs), S.arena.add(rs),
                                             (int)segs.size(), rows, true});
                    if (blk) {
                        // FrontBlockLowerBackwardSolve (lower_solve/local_front_block.hpp:50-72):
                        // Y := A21^T W (A21^H if Hermitian), X_s -= inv(ATL) Y
                        std::vector<GemmDesc<T>> y, xs;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.u == 0 || f.n == 0) continue;
                            y.push_back(GemmDesc<T>{L(s) + f.n, Wp(s), Tm(s), nullptr, nullptr, f.m,
                                                    f.u, f.n, f.n, k, f.u, G_TRANSA | cA});
                            xs.push_back(GemmDesc<T>{L(s), Tm(s), Xs(s), nullptr, nullptr, f.m, f.n,
                                                     N, f.n, k, f.n, G_ACCUM | G_NEG});
                        }
                        split(sb, y);
                        split(sb, xs);
                        continue;
                    }
                    split(sb, g);
                    // Fused panel steps (as in the forward sweep), super-panels
                    // and their panels from the last: X_p := inv(L_pp)^T X2_p,
                    // band rows X2(r) -= L(p, r)^T X_p for r in the super-panel,
                    // then X2(0:q0) -= L(q0:q1, 0:q0)^T X(q0:q1)
                    const bool fused = chain_level(lv, maxn);
                    const int maxq = fused ? (maxn + chainQ - 1) / chainQ : 0;
                    for (int qi = maxq - 1; qi >= 0; --qi) {
                        const int q0 = qi * chainQ, q1 = std::min(q0 + chainQ, maxn);
                        for (int w0 = q0 + ((q1 - q0 - 1) / SB) * SB; w0 >= q0; w0 -= SB)
                            panel_step(lv, w0, q0, true);
                        if (q0 == 0) continue;
                        std::vector<GemmDesc<T>> bulk;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.n <= q0) continue;
                            bulk.push_back(GemmDesc<T>{L(s) + q0, Xs(s) + q0, X2(s), nullptr,
                                                       nullptr, f.m, N, N, q0, k,
                                                       std::min(f.n, q0 + chainQ) - q0,
                                                       G_TRANSA | G_ACCUM | G_NEG | cA});
                        }
                        split(sb, bulk);
                    }
                    // SB-wide panels from the last: X_p := inv(L_pp)^T X2_p (^H if
                    // Hermitian; back into xn), X2(0:w0) -= L(p, 0:w0)^T X_p
                    const int maxw = fused ? 0 : (maxn + SB - 1) / SB;
                    for (int pw = maxw - 1; pw >= 0; --pw) {
                        const int w0 = pw * SB;
                        std::vector<GemmDesc<T>> tr, up;
                        for (int s : lv) {
                            const FrontLayout& f = P.fronts[s];
                            if (f.n <= w0) continue;
                            const int sbw = std::min(SB, f.n - w0);
                            T* x = Xs(s) + w0;
                            T* y = X2(s) + w0;
                            tr.push_back(GemmDesc<T>{LW(s, w0), y, x, nullptr, nullptr, sbw, N, N,
                                                     sbw, k, sbw, G_TRANSA | G_KLO | cA});
                            if (w0 > 0)
                                up.push_back(GemmDesc<T>{L(s) + w0, x, X2(s), nullptr, nullptr, f.m,
                                                         N, N, w0, k, sbw,
                                                         G_TRANSA | G_ACCUM | G_NEG | cA});
                        }
                        sb.gemm(tr, true);
                        sb.gemm(up);
                    }
                    if (pivoted() && !blk) permute(true);
                }
            }
        }
    }

    // one scheduled launch of step s on stream st
    void launch_one(const Step& s, const Arena& a, SolveSched* S, int k, hipStream_t st) {
        switch (s.kind) {
            case S_ZERO:
                launch_zero_segments<T>(a.at<ZeroSeg<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                        s.nwork, st);
                break;
            case S_ZLOWER:
                launch_zero_lower<T>(a.at<ZeroLowerDesc<T>>(s.desc), a.at<int>(s.start),
                                     s.ndesc, s.nwork, st);
                break;
            case S_EXTEND:
                launch_extend_add<T>(a.at<ExtendDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                     s.nwork, st);
                break;
            case S_DIAG:
                launch_diag_ldl<T>(a.at<DiagDesc<T>>(s.desc), s.ndesc, P.conjugate,
                                   P.dStatus.as<int>(), st);
                break;
            case S_GEMM:
                launch_gemm<T>(a.at<GemmDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc, s.nwork,
                               s.variant, s.scaled ? 2 : (s.update ? 3 : (S ? 1 : 0)), st);
                break;
            case S_GATHER:
                launch_gather<T>(a.at<SolveSeg>(s.desc), a.at<int>(s.start), s.ndesc, s.nwork,
                                 S->wp, S->x1, P.N, k, P.dLow.as<int>(), st);
                break;
            case S_REDUCE:
                launch_reduce_slices<T>(a.at<ReduceDesc<T>>(s.desc), a.at<int>(s.start),
                                        s.ndesc, s.nwork, st);
                break;
            case S_COPY:
                launch_copy2d<T>(a.at<CopyDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                 s.nwork, st);
                break;
            case S_SEXTEND:
                launch_solve_extend<T>(a.at<SolveExtend>(s.desc), a.at<int>(s.start), s.ndesc,
                                       s.nwork, S->wp, S->x1, P.N, P.dRel.as<int>(), k, st);
                break;
            case S_BK:
                launch_bk_panel<T>(a.at<BkDesc<T>>(s.desc), s.ndesc, s.variant, P.conjugate,
                                   P.dStatus.as<int>(), st);
                break;
            case S_TRINV:
                launch_unit_lower_inverse<T>(a.at<TriInvDesc<T>>(s.desc), s.ndesc, st);
                break;
            case S_GCOLS:
                launch_gather_cols<T>(a.at<GatherColsDesc<T>>(s.desc), a.at<int>(s.start),
                                      s.ndesc, s.nwork, st);
                break;
            case S_QSCALE:
                launch_quasi_scale<T>(a.at<QuasiDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                      s.nwork, P.conjugate, st);
                break;
            case S_QROWS:
                launch_quasi_rows<T>(a.at<QuasiRowsDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                     s.nwork, P.conjugate, st);
                break;
            case S_PSYM:
                launch_permute_sym<T>(a.at<PermSymDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                      s.nwork, st);
                break;
            case S_WINV:
                launch_wide_inverse_small<T>(a.at<WideInvDesc<T>>(s.desc), s.ndesc, st);
                break;
            case S_PSTEP:
                launch_panel_step<T>(a.at<PanelStepDesc<T>>(s.desc), a.at<int>(s.start), s.ndesc,
                                     s.nwork, s.backward, P.conjugate, st);
                break;
            case S_PERM:
                // gather (variant 0) or scatter (1) of the fronts' rows from
                // the first nodal buffer into the second
                launch_permute_front_rows<T>(a.at<PermSeg>(s.desc), a.at<int>(s.start), s.ndesc,
                                             s.nwork, S->x1, S->x2, P.N, k, pivpool.as<int>(),
                                             s.variant != 0, st);
                break;
        }
    }

    // Autotuning pass: every launch on the plan's stream, each grouped GEMM
    // once per admissible tile shape between events (its results are
    // meaningless: the products are applied several times).
    struct TuneSample {
        int bucket, variant;
        hipEvent_t e0, e1;
        int step;   // running index of the launch in the pass
    };
    std::vector<TuneSample>* tuning = nullptr;
    int tuneStep = 0;

    void run(const std::vector<Step>& steps, const Arena& a, SolveSched* S, int k) {
        for (const Step& s : steps) {
            if (tuning) {
                if (s.kind == S_RECORD || s.kind == S_WAIT) continue;
                if (s.kind != S_GEMM) {
                    launch_one(s, a, S, k, P.stream);
                    continue;
                }
                const int id = tuneStep++;
                for (int i = 0; i < s.nvar; ++i) {
                    Step t = s;
                    set_variant(t, s.vars[i]);
                    TuneSample x{s.bucket, s.vars[i], P.next_event(), nullptr, id};
                    CQ_HIP(hipEventRecord(x.e0, P.stream));
                    launch_one(t, a, S, k, P.stream);
                    x.e1 = P.next_event();
                    CQ_HIP(hipEventRecord(x.e1, P.stream));
                    tuning->push_back(x);
                }
                continue;
            }
            // a launch's profiling events go on its own stream: the interval is
            // the kernel's duration as it runs beside the other stream's work
            hipStream_t st = s.stream ? P.stream2 : P.stream;
            if (s.kind == S_RECORD || s.kind == S_WAIT) {
                if (s.kind == S_RECORD)
                    CQ_HIP(hipEventRecord(P.depEvents[s.event], st));
                else
                    CQ_HIP(hipStreamWaitEvent(st, P.depEvents[s.event], 0));
                continue;
            }
            hipEvent_t e0 = nullptr;
            const bool prof = P.profiling && !a.base;   // never inside the allocation-free solve
            if (prof) {
                e0 = P.next_event();
                CQ_HIP(hipEventRecord(e0, st));
            }
            launch_one(s, a, S, k, st);
            if (prof) {
                hipEvent_t e1 = P.next_event();
                CQ_HIP(hipEventRecord(e1, st));
                const int kind = (s.kind == S_GEMM && S) ? 7 : s.kind;
                // factor GEMMs by instantiation: pivot-scaled updates (Schur,
                // panel and outer-panel updates) apart from the TRSM-like rest
                // (each tile shape's updates apart: different kernels)
                const char* name = (s.kind == S_GEMM && !S)
                                       ? ((s.update || s.scaled)
                                              ? (s.variant == TILE_32       ? "gemm_update_small"
                                                 : s.variant == TILE_128    ? "gemm_update_large"
                                                 : s.variant == TILE_128_8W ? "gemm_update_w8"
                                                 : s.variant == TILE_32_K64 ? "gemm_update_k64"
                                                                            : "gemm_update")
                                              : "gemm_trsm")
                                   : (s.kind == S_ZERO && S) ? "solve_zero"
                                   : (kind == 7 && s.variant == TILE_STREAM) ? "solve_gemm_stream"
                                                             : step_name(kind);
                P.pending.push_back(Plan::Pending{name, s.flops, s.bytes, e0, e1, s.ndesc, s.nwork,
                                                  s.variant, s.stream ? 1 : 0});
            }
        }
        CQ_HIP(hipGetLastError());
    }

    // Record body's launches on the plan's stream into g (once). Capture
    // needs a stream other than the legacy default one; where it fails the
    // launches run directly. No memset is recorded (the graphs hold kernels
    // only).
    template <class F>
    bool replay(Graph& g, F&& body) {
        if (!P.useGraphs || P.profiling) return false;
        if (!g.tried) {
            g.tried = true;
            hipStream_t st = P.stream;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
            hipGraph_t graph = nullptr;
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            if (hipStreamEndCapture(st, &graph) != hipSuccess || !graph) {
                (void)hipGetLastError();
                return false;
            }
            const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                g.exec = nullptr;
                return false;
            }
        }
        if (!g.exec) return false;
        CQ_HIP(hipGraphLaunch(g.exec, P.stream));
        return true;
    }

    void ensure_factor_schedule() {
        const int ft = P.frontType;
        if (!fbuilt[ft]) {
            build_factor_schedule(farenas[ft], fsched[ft], block(), pivoted());
            P.ensure_dep_events(nextEvent);
            fbuilt[ft] = true;
        }
    }

    void factor(const V* vals) {
        hipStream_t st = P.stream;
        ensure_factor_schedule();
        CQ_HIP(hipMemsetAsync(fpool.p, 0, fpool.bytes, st));
        // host-staged fronts replace their assembled values: converted to
        // the working type on the host, copied in after the assembly
        std::vector<std::vector<T>> staged;
        auto body = [&] {
            launch_assemble<T, V>(vals, P.dAsmDst.as<int64_t>(), P.nnz, fpool.as<T>(), P.conjugate,
                                  st);
            for (const auto& kv : P.stagedFronts) {
                const FrontLayout& f = P.fronts[kv.first];
                const size_t mn = (size_t)f.m * f.n;
                const V* src = reinterpret_cast<const V*>(kv.second.data());
                staged.emplace_back(mn);
                for (size_t e = 0; e < mn; ++e) staged.back()[e] = cvt<T>(src[e]);
                CQ_HIP(hipMemcpyAsync(L(kv.first), staged.back().data(), mn * sizeof(T),
                                      hipMemcpyHostToDevice, st));
            }
            run(fsched[P.frontType], farenas[P.frontType], nullptr, 0);
        };
        if (!P.stagedFronts.empty() || !replay(fgraph[P.frontType], body)) body();
        // the copies read the host buffers until they complete
        CQ_HIP(hipStreamSynchronize(st));
        P.stagedFronts.clear();
    }

    // x: V, N x k (ld ldx); nodal or original ordering; solved in place
    void solve(V* x, int64_t ldx, int k, bool nodal) {
        SolveSched& S = schedule(k);
        hipStream_t st = P.stream;
        if (k > 65535) throw Error(1, "at most 65535 right-hand sides per solve");
        const int* map = P.dMap.as<int>();
        if (nodal) {
            if (!identity.p) {
                std::vector<int> id(P.N);
                std::iota(id.begin(), id.end(), 0);
                upload(identity, id);
            }
            map = identity.as<int>();
        }
        launch_pull<T, V>(x, ldx, map, S.x1, P.N, k, st);
        solve_core(S, k);
        launch_push<T, V>(S.x1, map, x, ldx, P.N, k, st);
        CQ_HIP(hipGetLastError());
    }

    // The allocation-free solve: the same sweeps with every buffer in the
    // scratch array (no device memory is allocated or freed; the first call
    // for a right-hand-side count builds its launch schedule on the host).
    void solve_noalloc(V* x, int64_t ldx, int k, char* base, size_t cap) {
        if (k > 65535) throw Error(1, "at most 65535 right-hand sides per solve");
        SolveSched& S = noalloc_schedule(k, base, cap);
        hipStream_t st = P.stream;
        // the scratch array holds the last call's data: this call's
        // descriptors go to its head, its status word is cleared
        CQ_HIP(hipMemcpyAsync(base, S.arena.host.data(), S.arena.host.size(),
                              hipMemcpyHostToDevice, st));
        launch_fill_int(S.status, 4, 0, st);
        launch_pull<T, V>(x, ldx, P.dMap.as<int>(), S.x1, P.N, k, st);
        solve_core(S, k);
        launch_push<T, V>(S.x1, P.dMap.as<int>(), x, ldx, P.N, k, st);
        CQ_HIP(hipGetLastError());
        int h[3] = {0, 0, 0};
        CQ_HIP(hipMemcpyAsync(h, S.status, sizeof(h), hipMemcpyDeviceToHost, st));
        CQ_HIP(hipStreamSynchronize(st));
        if (h[0] != 0)
            throw Error(2, "Solve: zero or non-finite pivot at nodal index " + std::to_string(h[2]));
    }

    // forward, diagonal and backward sweeps on the nodal buffer S.x1 (graph
    // replay only for the plan's own schedules)
    void solve_core(SolveSched& S, int k) {
        auto body = [&] {
            run(S.fwd, S.arena, &S, k);
            // block fronts hold inv(ATL), so there is no separate diagonal
            // solve (solve.hpp:48-53)
            if (pivoted() && !block())
                launch_quasi_diag_solve<T, V>(S.x2, dpool.as<V>(), subpool.as<V>(), P.N, k,
                                              P.conjugate, S.status, P.stream);
            else if (!block())
                launch_diag_solve<T, V>(S.x2, dpool.as<V>(), P.N, k, S.status, P.stream);
            run(S.bwd, S.arena, &S, k);
        };
        if (S.arena.base || !replay(S.graph, body)) body();
    }

    // The autotuning pass (see TuneTable): the current front type's
    // factorization of vals and, with rhs > 0, the forward and backward
    // sweeps of a solve with rhs right-hand sides, every grouped GEMM timed
    // in each admissible tile shape. Records the fastest shape per bucket,
    // gives every built schedule of this plan the tuned shapes and drops the
    // recorded graphs. The fronts hold no factorization afterwards.
    void autotune(const V* vals, int rhs, int* nBuckets, int* nChanged) {
        hipStream_t st = P.stream;
        ensure_factor_schedule();
        std::vector<TuneSample> samples;
        tuning = &samples;
        tuneStep = 0;
        try {
            CQ_HIP(hipMemsetAsync(fpool.p, 0, fpool.bytes, st));
            launch_assemble<T, V>(vals, P.dAsmDst.as<int64_t>(), P.nnz, fpool.as<T>(), P.conjugate,
                                  st);
            run(fsched[P.frontType], farenas[P.frontType], nullptr, 0);
            if (rhs > 0) {
                SolveSched& S = schedule(rhs);
                run(S.fwd, S.arena, &S, rhs);
                run(S.bwd, S.arena, &S, rhs);
            }
        } catch (...) {
            tuning = nullptr;
            throw;
        }
        tuning = nullptr;
        CQ_HIP(hipStreamSynchronize(st));
        // per launch: milliseconds of each tile shape
        constexpr int NV = NUM_GEMM_VARIANTS;
        using Times = std::array<double, NV>;
        Times unmeasured;
        unmeasured.fill(-1.0);
        std::map<int, std::pair<int, Times>> perStep;   // step -> (bucket, ms)
        for (const TuneSample& x : samples) {
            float ms = 0.f;
            CQ_HIP(hipEventElapsedTime(&ms, x.e0, x.e1));
            auto it = perStep.find(x.step);
            if (it == perStep.end())
                it = perStep.emplace(x.step, std::make_pair(x.bucket, unmeasured)).first;
            it->second.second[variant_slot(x.variant)] = ms;
        }
        P.reset_events();
        // per bucket: total of each shape over the launches that admit every
        // shape the bucket's launches have in common
        std::map<int, Times> tot;
        std::map<int, std::array<bool, NV>> common;
        for (const auto& kv : perStep) {
            const int b = kv.second.first;
            const auto& t = kv.second.second;
            auto ci = common.find(b);
            if (ci == common.end()) {
                std::array<bool, NV> c;
                for (int i = 0; i < NV; ++i) c[i] = t[i] >= 0;
                common[b] = c;
                tot[b].fill(0.0);
            } else {
                for (int i = 0; i < NV; ++i) ci->second[i] = ci->second[i] && t[i] >= 0;
            }
            for (int i = 0; i < NV; ++i)
                if (t[i] >= 0) tot[b][i] += t[i];
        }
        std::map<int, int> before;   // a shape each bucket used before the pass
        auto note = [&](const std::vector<Step>& steps) {
            for (const Step& s : steps)
                if (s.kind == S_GEMM && s.bucket >= 0) before.emplace(s.bucket, s.variant);
        };
        for (const auto& fs : fsched) note(fs);
        for (auto& kv : solves) {
            note(kv.second->fwd);
            note(kv.second->bwd);
        }
        int changed = 0;
        {
            TuneTable& tt = tune_table();
            std::lock_guard<std::mutex> lock(tt.mu);
            for (const auto& kv : tot) {
                const int b = kv.first;
                int best = -1;
                for (int i = 0; i < NV; ++i)
                    if (common[b][i] && (best < 0 || kv.second[i] < kv.second[best])) best = i;
                if (best < 0) continue;
                tt.choice[b] = GEMM_VARIANTS[best];
                for (int i = 0; i < NV; ++i) tt.ms[b][i] = common[b][i] ? kv.second[i] : -1.0;
                auto bi = before.find(b);
                if (bi != before.end() && bi->second != GEMM_VARIANTS[best]) ++changed;
            }
            tune_save_locked(tt);
        }
        apply_tuning();
        *nBuckets = (int)tot.size();
        *nChanged = changed;
    }

    // Every built schedule of this plan takes the tuned tile shapes.
    void apply_tuning() {
        auto apply = [](std::vector<Step>& steps) {
            for (Step& s : steps) {
                if (s.kind != S_GEMM) continue;
                const int v = tuned_variant(s.bucket);
                if (v >= 0 && step_has_variant(s, v)) set_variant(s, v);
            }
        };
        for (auto& fs : fsched) apply(fs);
        for (auto& kv : solves) {
            apply(kv.second->fwd);
            apply(kv.second->bwd);
            drop(kv.second->graph);
        }
        for (auto& kv : noallocSolves) {
            apply(kv.second->fwd);
            apply(kv.second->bwd);
        }
        for (Graph& g : fgraph) drop(g);
    }
    static void drop(Graph& g) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.tried = false;
    }

    V* scratch_cols(int k) {
        if (rbufCols < k) {
            rbuf.alloc(std::max<int64_t>((int64_t)P.N * k, 1) * sizeof(V));
            rbufCols = k;
        }
        return rbuf.as<V>();
    }

    // D's subdiagonal and the pivot permutation of a front of a pivoted
    // factorization (SymmFront::subdiag, SymmFront::piv)
    void get_pivots(int s, void* hsub, int* hpiv) {
        const FrontLayout& f = P.fronts[s];
        if (f.n == 0) return;
        CQ_HIP(hipStreamSynchronize(P.stream));
        if (hsub) CQ_HIP(hipMemcpy(hsub, Sub(s), f.n * sizeof(V), hipMemcpyDeviceToHost));
        if (hpiv) CQ_HIP(hipMemcpy(hpiv, Piv(s), f.n * sizeof(int), hipMemcpyDeviceToHost));
    }

    void get_front(int s, void* hL, void* hdiag) {
        const FrontLayout& f = P.fronts[s];
        const int64_t mn = (int64_t)f.m * f.n;
        if (hL && mn > 0) {
            DevBuf tmp(mn * sizeof(V));
            launch_widen<T, V>(L(s), tmp.as<V>(), mn, P.stream);
            CQ_HIP(hipStreamSynchronize(P.stream));
            CQ_HIP(hipMemcpy(hL, tmp.p, mn * sizeof(V), hipMemcpyDeviceToHost));
        }
        // block fronts keep no diagonal (ldl/local.hpp:106-107)
        if (hdiag && f.n > 0 && !block()) {
            CQ_HIP(hipStreamSynchronize(P.stream));
            CQ_HIP(hipMemcpy(hdiag, D(s), f.n * sizeof(V), hipMemcpyDeviceToHost));
        }
    }
};

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
Plan::Plan(std::shared_ptr<const Symbolic> symIn, const int64_t* rowOff, const int* cols,
           bool cpx, bool conj, hipStream_t st)
    : sym(std::move(symIn)), complex(cpx), conjugate(conj), stream(st) {
    const Symbolic& S = *sym;
    N = S.N;
    if (const char* e = std::getenv("CLIQG_LOOKAHEAD")) lookahead = std::atoi(e) != 0;
    if (const char* e = std::getenv("CLIQG_SOLVE_BLOCK")) {
        const int v = std::atoi(e);
        if (v < NB || v % NB) throw Error(1, "CLIQG_SOLVE_BLOCK must be a positive multiple of 64");
        sb = v;
    }
    numFronts = (int)S.nodes.size();
    nnz = rowOff[N] - rowOff[0];
    hostRowOff.assign(rowOff, rowOff + N + 1);
    for (auto& r : hostRowOff) r -= rowOff[0];
    const int* colp = cols + rowOff[0];

    fronts.resize(numFronts);
    int64_t Lp = 0, Up = 0, Wp = 0, Ip = 0, IWp = 0;
    std::vector<int> relAll, lowAll;
    int maxh = 0;
    for (int s = 0; s < numFronts; ++s) {
        const SymNode& nd = S.nodes[s];
        FrontLayout& f = fronts[s];
        f.n = nd.size;
        f.u = (int)nd.lowerStruct.size();
        f.m = f.n + f.u;
        f.off = nd.off;
        f.parent = nd.parent;
        f.nchild = (int)nd.children.size();
        f.child[0] = f.nchild > 0 ? nd.children[0] : -1;
        f.child[1] = f.nchild > 1 ? nd.children[1] : -1;
        f.height = 0;
        for (int c = 0; c < f.nchild; ++c) f.height = std::max(f.height, fronts[f.child[c]].height + 1);
        maxh = std::max(maxh, f.height);
        f.slot = -1;
        f.L = Lp; Lp += (int64_t)f.m * f.n;
        f.U = 0;
        f.W = Wp; Wp += f.u;
        f.Linv = Ip; Ip += (int64_t)((f.n + NB - 1) / NB) * NB * NB;
        f.LinvW = IWp;
        for (int j0 = 0; j0 < f.n; j0 += sb) {
            const int64_t w = std::min(sb, f.n - j0);
            IWp += w * w;
        }
        f.low = (int64_t)lowAll.size();
        lowAll.insert(lowAll.end(), nd.lowerStruct.begin(), nd.lowerStruct.end());
    }
    for (int s = 0; s < numFronts; ++s) {
        FrontLayout& f = fronts[s];
        f.rel = (int64_t)relAll.size();
        if (f.parent >= 0) {
            const SymNode& p = S.nodes[f.parent];
            f.slot = (p.children.size() > 0 && p.children[0] == s) ? 0 : 1;
            const std::vector<int>& r = f.slot == 0 ? p.leftRelInds : p.rightRelInds;
            if ((int)r.size() != f.u) throw Error(1, "relative indices do not match the update size");
            // the extend-adds split a child's update at its first index past
            // the parent's own columns, which needs increasing relative indices
            for (size_t t = 1; t < r.size(); ++t)
                if (r[t] <= r[t - 1]) throw Error(1, "relative indices are not increasing");
            relAll.insert(relAll.end(), r.begin(), r.end());
        } else if (f.u != 0) {
            throw Error(1, "a root front has a nonempty lower structure");
        }
    }
    factorEntries = Lp;
    wRows = Wp;
    linvEntries = Ip;
    linvWEntries = IWp;
    numLevels = numFronts ? maxh + 1 : 0;
    levels.assign(numLevels, {});
    for (int s = 0; s < numFronts; ++s) levels[fronts[s].height].push_back(s);
    // update matrices live from their front's level (extend-add target) to
    // their parent's level (extend-add source); storage is reused after that
    {
        IntervalAlloc ua;
        for (int h = 0; h < numLevels; ++h) {
            for (int s : levels[h]) fronts[s].U = ua.alloc((int64_t)fronts[s].u * fronts[s].u);
            for (int s : levels[h])
                for (int c = 0; c < fronts[s].nchild; ++c) {
                    const FrontLayout& fc = fronts[fronts[s].child[c]];
                    ua.release(fc.U, (int64_t)fc.u * fc.u);
                }
        }
        updateEntries = ua.top;
        (void)Up;
    }

    // value -> front scatter map (dist_symm_front_tree/impl.hpp:255-296)
    std::vector<int64_t> dst(nnz, -1);
    for (int s = 0; s < numFronts; ++s) {
        const SymNode& nd = S.nodes[s];
        const FrontLayout& f = fronts[s];
        for (int t = 0; t < f.n; ++t) {
            const int i = nd.inds[t];
            for (int64_t e = hostRowOff[i]; e < hostRowOff[i + 1]; ++e) {
                const int c = colp[e];
                if (c < 0 || c >= N) throw Error(1, "column index out of range");
                const int target = S.map[c];
                if (target < f.off + t) continue;
                int row;
                if (target < f.off + f.n) {
                    row = target - f.off;
                } else {
                    auto it = std::lower_bound(nd.origLowerStruct.begin(), nd.origLowerStruct.end(),
                                               target);
                    if (it == nd.origLowerStruct.end() || *it != target)
                        throw Error(1, "matrix pattern is not covered by the symbolic analysis");
                    row = nd.origLowerRelInds[it - nd.origLowerStruct.begin()];
                }
                dst[e] = f.L + row + (int64_t)t * f.m;
            }
        }
    }

    CQ_HIP(hipGetDevice(&device));
    if (!stream) {
        CQ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        ownStream = true;
    }
    CQ_HIP(hipEventCreate(&ev0_));
    CQ_HIP(hipEventCreate(&ev1_));
    CQ_HIP(hipEventCreateWithFlags(&evCaller_, hipEventDisableTiming));
    CQ_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    upload(dRowOff, hostRowOff);
    std::vector<int> colv(colp, colp + nnz);
    upload(dCols, colv);
    upload(dAsmDst, dst);
    upload(dRel, relAll);
    relHost = std::move(relAll);
    upload(dLow, lowAll);
    upload(dMap, S.map);
    dStatus.alloc(4 * sizeof(int));
    dNorms.alloc(4096 * sizeof(unsigned long long));
    dVals.alloc(std::max<int64_t>(nnz, 1) * (complex ? 16 : 8));
}

Plan::~Plan() {
    num.reset();
    for (hipEvent_t e : evPool_) (void)hipEventDestroy(e);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (evCaller_) (void)hipEventDestroy(evCaller_);
    for (hipEvent_t e : depEvents) (void)hipEventDestroy(e);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (ownStream && stream) (void)hipStreamDestroy(stream);
}

void Plan::ensure_dep_events(int n) {
    while ((int)depEvents.size() < n) {
        hipEvent_t e;
        CQ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        depEvents.push_back(e);
    }
}

void Plan::after_caller() {
    if (!ownStream) return;   // the caller's stream orders the work itself
    CQ_HIP(hipEventRecord(evCaller_, nullptr));
    CQ_HIP(hipStreamWaitEvent(stream, evCaller_, 0));
}

size_t Plan::device_bytes() const {
    size_t b = dRowOff.bytes + dCols.bytes + dAsmDst.bytes + dRel.bytes + dLow.bytes +
               dMap.bytes + dStatus.bytes + dNorms.bytes + dVals.bytes;
    if (num) b += num->device_bytes();
    return b;
}

hipEvent_t Plan::next_event() {
    if (evNext_ == evPool_.size()) {
        hipEvent_t e;
        CQ_HIP(hipEventCreate(&e));
        evPool_.push_back(e);
    }
    return evPool_[evNext_++];
}

void Plan::resolve_profile() {
    // CLIQG_PROFILE_TRACE=<file>: one line per profiled launch, appended
    static const char* tracePath = std::getenv("CLIQG_PROFILE_TRACE");
    FILE* trace = tracePath ? std::fopen(tracePath, "a") : nullptr;
    for (const Pending& p : pending) {
        float ms = 0.f, t0 = 0.f;
        CQ_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        if (trace) {
            // start time from the first profiled launch of this call
            CQ_HIP(hipEventElapsedTime(&t0, pending.front().e0, p.e0));
            std::fprintf(trace, "%s %.4f %.6e %.6e %d %d %d %d %.4f\n", p.name, ms, p.flops,
                         p.bytes, p.ndesc, p.nwork, p.variant, p.stream, t0);
        }
        ProfileEntry& e = profile[p.name];
        e.launches += 1;
        e.ms += ms;
        e.flops += p.flops;
        e.bytes += p.bytes;
    }
    if (trace) std::fclose(trace);
    pending.clear();
    evNext_ = 0;
}

int Plan::check_status(const char* what) {
    int st[3] = {0, 0, 0};
    CQ_HIP(hipMemcpyAsync(st, dStatus.p, sizeof(st), hipMemcpyDeviceToHost, stream));
    CQ_HIP(hipStreamSynchronize(stream));
    if (profiling) resolve_profile();
    if (st[0] != 0)
        throw Error(2, std::string(what) + ": zero or non-finite pivot in front " +
                           std::to_string(st[1]) + " at column " + std::to_string(st[2]));
    return 0;
}

template <class F>
static void dispatch(bool complex, int precision, F&& f) {
    if (complex) {
        if (precision == 0) f((cdouble*)nullptr); else f((cfloat*)nullptr);
    } else {
        if (precision == 0) f((double*)nullptr); else f((float*)nullptr);
    }
}

void Plan::factor(const void* d_vals, int prec, int ftype) {
    after_caller();
    if (prec != 0 && prec != 1) throw Error(1, "unknown precision");
    if (ftype < 0 || ftype > 3) throw Error(1, "unknown front type");
    if (!num || prec != precision) {
        num.reset();
        precision = prec;
        dispatch(complex, prec, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            num = std::make_unique<Numeric<T>>(*this);
        });
    }
    frontType = ftype;
    factored = false;
    const size_t vb = complex ? 16 : 8;
    CQ_HIP(hipMemcpyAsync(dVals.p, d_vals, nnz * vb, hipMemcpyDeviceToDevice, stream));
    CQ_HIP(hipMemsetAsync(dStatus.p, 0, dStatus.bytes, stream));
    CQ_HIP(hipMemsetAsync(dNorms.p, 0, sizeof(unsigned long long), stream));
    if (complex)
        launch_row_abs_sum_max<cdouble>(N, dRowOff.as<int64_t>(), dVals.as<cdouble>(),
                                        dNorms.as<unsigned long long>(), stream);
    else
        launch_row_abs_sum_max<double>(N, dRowOff.as<int64_t>(), dVals.as<double>(),
                                       dNorms.as<unsigned long long>(), stream);
    CQ_HIP(hipEventRecord(ev0_, stream));
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        static_cast<Numeric<T>*>(num.get())->factor(dVals.as<V>());
    });
    CQ_HIP(hipEventRecord(ev1_, stream));
    unsigned long long na = 0;
    CQ_HIP(hipMemcpyAsync(&na, dNorms.p, sizeof(na), hipMemcpyDeviceToHost, stream));
    check_status("LDL");
    float ms = 0.f;
    CQ_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    lastFactorMs_ = ms;
    double nd;
    std::memcpy(&nd, &na, sizeof(nd));
    normA = nd;
    factored = true;
}

void Plan::solve(void* d_x, int64_t ldx, int k, bool nodal) {
    after_caller();
    if (!factored) throw Error(6, "the plan has not been factored");
    if (k <= 0) return;
    if (ldx < N) throw Error(1, "leading dimension smaller than N");
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        static_cast<Numeric<T>*>(num.get())->solve(static_cast<V*>(d_x), ldx, k, nodal);
    });
    check_status("Solve");
}

void Plan::solve_noalloc(void* d_x, int64_t ldx, int k) {
    if (!factored) throw Error(6, "the plan has not been factored");
    if (k <= 0) return;
    if (ldx < N) throw Error(1, "leading dimension smaller than N");
    int dev = -1;
    CQ_HIP(hipGetDevice(&dev));
    if (dev != device)
        throw Error(1, "the allocation-free solve must run with the plan's device current");
    char* base = scratch_address(dev);   // first call on this device: the lookup
    after_caller();
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        static_cast<Numeric<T>*>(num.get())->solve_noalloc(static_cast<V*>(d_x), ldx, k, base,
                                                           NOALLOC_SCRATCH_BYTES);
    });
}

void Plan::multiply(double alpha, const void* d_x, int64_t ldx, double beta, void* d_y,
                    int64_t ldy, int k) {
    after_caller();
    if (!factored) throw Error(6, "the plan holds no matrix values yet");
    if (complex)
        launch_spmv<cdouble>(N, dRowOff.as<int64_t>(), dCols.as<int>(), dVals.as<cdouble>(), alpha,
                             (const cdouble*)d_x, ldx, beta, (cdouble*)d_y, ldy, k, stream);
    else
        launch_spmv<double>(N, dRowOff.as<int64_t>(), dCols.as<int>(), dVals.as<double>(), alpha,
                            (const double*)d_x, ldx, beta, (double*)d_y, ldy, k, stream);
    CQ_HIP(hipGetLastError());
    if (ownStream) {
        // the caller's later work on the default stream reads y
        CQ_HIP(hipEventRecord(evCaller_, stream));
        CQ_HIP(hipStreamWaitEvent(nullptr, evCaller_, 0));
    }
}

void Plan::refine(const void* d_b, int64_t ldb, void* d_x, int64_t ldx, int k, double tol,
                  int maxIters, int* itersOut, int* convOut, double* etaOut) {
    after_caller();
    if (!factored) throw Error(6, "the plan has not been factored");
    if (k <= 0) return;
    if (k > 1000) throw Error(1, "at most 1000 right-hand sides per refinement call");
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        auto* nm = static_cast<Numeric<T>*>(num.get());
        const V* b = static_cast<const V*>(d_b);
        V* x = static_cast<V*>(d_x);
        unsigned long long* nrm = dNorms.as<unsigned long long>();
        // nrm[0] = ||A||, nrm[1..k] = |b|, [k+1..2k] = |x|, [2k+1..3k] = |r|
        CQ_HIP(hipMemsetAsync(nrm + 1, 0, sizeof(unsigned long long) * 3 * k, stream));
        if (k > 65535) throw Error(1, "at most 65535 right-hand sides per solve");
        auto& S = nm->schedule(k);
        T* xnb = S.x1;
        // x0 = solve(b): b goes straight into the solve's nodal buffer
        launch_col_abs_max<V>(b, ldb, N, k, nrm + 1, stream);
        launch_pull<T, V>(b, ldb, dMap.as<int>(), xnb, N, k, stream);
        nm->solve_core(S, k);
        launch_push<T, V>(xnb, dMap.as<int>(), x, ldx, N, k, stream);
        launch_col_abs_max<V>(x, ldx, N, k, nrm + 1 + k, stream);
        std::vector<unsigned long long> h(3 * k + 1);
        int it = 0;
        bool conv = false;
        double eta = 0.0;
        for (;;) {
            // r = b - A x with |r| per column in one pass, written in nodal
            // order and working precision straight into the solve's buffer;
            // |x| came with x
            launch_residual_nodal<T, V>(N, dRowOff.as<int64_t>(), dCols.as<int>(), dVals.as<V>(), x,
                                        ldx, b, ldb, k, nrm + 1 + 2 * k, xnb, dMap.as<int>(),
                                        stream);
            CQ_HIP(hipMemcpyAsync(h.data(), nrm, sizeof(unsigned long long) * (3 * k + 1),
                                  hipMemcpyDeviceToHost, stream));
            CQ_HIP(hipStreamSynchronize(stream));
            eta = 0.0;
            std::vector<int> act;   // columns still above the tolerance
            for (int j = 0; j < k; ++j) {
                double bn, xn, rn;
                std::memcpy(&bn, &h[1 + j], 8);
                std::memcpy(&xn, &h[1 + k + j], 8);
                std::memcpy(&rn, &h[1 + 2 * k + j], 8);
                const double den = normA * xn + bn;
                double e = den > 0 ? rn / den : 0.0;
                if (!std::isfinite(e)) e = INFINITY;
                eta = std::max(eta, e);
                if (!(e <= tol)) act.push_back(j);
            }
            if (eta <= tol) { conv = true; break; }
            if (it >= maxIters) break;
            // Converged columns keep their x: once at least 16 columns have
            // converged, the correction solve takes only the others, gathered
            // into a narrower solve (width rounded up to a multiple of 16, so
            // few distinct schedules are built; padding columns are zero).
            const int kc = (int)act.size();
            const int kr = std::min(k, (kc + 15) / 16 * 16);
            if (kr < k) {
                auto& S2 = nm->schedule(kr);
                for (int c = 0; c < kc; ++c)
                    CQ_HIP(hipMemcpyAsync(S2.x1 + (int64_t)c * N, xnb + (int64_t)act[c] * N,
                                          sizeof(T) * N, hipMemcpyDeviceToDevice, stream));
                if (kr > kc)
                    CQ_HIP(hipMemsetAsync(S2.x1 + (int64_t)kc * N, 0, sizeof(T) * N * (kr - kc),
                                          stream));
                nm->solve_core(S2, kr);
                CQ_HIP(hipMemsetAsync(nrm + 1 + 2 * k, 0, sizeof(unsigned long long) * k, stream));
                for (int c = 0; c < kc; ++c) {
                    CQ_HIP(hipMemsetAsync(nrm + 1 + k + act[c], 0, sizeof(unsigned long long),
                                          stream));
                    launch_push_add_abs<T, V>(S2.x1 + (int64_t)c * N, dMap.as<int>(),
                                              x + (int64_t)act[c] * ldx, ldx, N, 1,
                                              nrm + 1 + k + act[c], stream);
                }
            } else {
                nm->solve_core(S, k);
                CQ_HIP(hipMemsetAsync(nrm + 1 + k, 0, sizeof(unsigned long long) * 2 * k, stream));
                launch_push_add_abs<T, V>(xnb, dMap.as<int>(), x, ldx, N, k, nrm + 1 + k, stream);
            }
            ++it;
        }
        *itersOut = it;
        *convOut = conv ? 1 : 0;
        *etaOut = eta;
    });
    check_status("refine");
}

void Plan::adjoint(const void* d_g, int64_t ldg, const void* d_x, int64_t ldx, int k,
                   void* d_lambda, int64_t ldl, void* d_dA) {
    after_caller();
    if (!factored) throw Error(6, "the plan has not been factored");
    if (k <= 0) return;
    if (precision == 1) {
        // FP32 factor: lambda is refined in FP64 like the forward solution
        int it = 0, conv = 0;
        double eta = 0.0;
        refine(d_g, ldg, d_lambda, ldl, k, 1e-12, 30, &it, &conv, &eta);
        if (!conv)
            throw Error(3, "adjoint refinement stopped at backward error " + sci(eta));
    }
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        auto* nm = static_cast<Numeric<T>*>(num.get());
        V* lam = static_cast<V*>(d_lambda);
        if (precision == 0) {
            launch_copy_cols<V>(lam, ldl, static_cast<const V*>(d_g), ldg, N, k, stream);
            nm->solve(lam, ldl, k, false);
        }
        if (d_dA)
            launch_adjoint_dA<V>(N, dRowOff.as<int64_t>(), dCols.as<int>(), lam, ldl,
                                 static_cast<const V*>(d_x), ldx, k, conjugate,
                                 static_cast<V*>(d_dA), stream);
    });
    check_status("adjoint");
}

void Plan::set_front(int s, const void* hL) {
    if (s < 0 || s >= numFronts) throw Error(1, "front index out of range");
    const size_t mn = (size_t)fronts[s].m * fronts[s].n;
    if (mn == 0) return;
    if (!hL) throw Error(1, "null front values");
    const size_t vb = complex ? 16 : 8;
    std::vector<char>& buf = stagedFronts[s];
    buf.resize(mn * vb);
    std::memcpy(buf.data(), hL, mn * vb);
    // only the lower trapezoid is a front's data: its strict upper triangle
    // is kept zero, as the assembly leaves it
    const int m = fronts[s].m, n = fronts[s].n;
    for (int j = 1; j < n; ++j) std::memset(buf.data() + (size_t)j * m * vb, 0, (size_t)j * vb);
}

void Plan::autotune(const void* d_vals, int prec, int ftype, int rhs, int* nBuckets,
                    int* nChanged) {
    if (rhs < 0 || rhs > 65535) throw Error(1, "bad right-hand-side count for tuning");
    factor(d_vals, prec, ftype);
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        using V = typename Scalar<T>::wide;
        static_cast<Numeric<T>*>(num.get())->autotune(dVals.as<V>(), rhs, nBuckets, nChanged);
    });
    factored = false;
    factor(d_vals, prec, ftype);
}

std::string tune_table_text() {
    TuneTable& t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    tune_load_locked(t);
    std::string out;
    char line[256];
    for (size_t b = 0; b < t.keys.size(); ++b) {
        if (t.choice[b] < 0) continue;
        std::snprintf(line, sizeof(line), "%s %d", t.keys[b].c_str(), t.choice[b]);
        out += line;
        for (int i = 0; i < NUM_GEMM_VARIANTS; ++i) {
            std::snprintf(line, sizeof(line), " %.4f", t.ms[b][i]);
            out += line;
        }
        out += "\n";
    }
    return out;
}

void Plan::get_pivots(int s, void* hsub, int* hpiv) {
    if (!factored) throw Error(6, "the plan has not been factored");
    if (frontType != 2) throw Error(6, "the factorization is not pivoted");
    if (s < 0 || s >= numFronts) throw Error(1, "front index out of range");
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        static_cast<Numeric<T>*>(num.get())->get_pivots(s, hsub, hpiv);
    });
}

void Plan::get_front(int s, void* hL, void* hdiag) {
    if (!factored) throw Error(6, "the plan has not been factored");
    if (s < 0 || s >= numFronts) throw Error(1, "front index out of range");
    dispatch(complex, precision, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        static_cast<Numeric<T>*>(num.get())->get_front(s, hL, hdiag);
    });
}

}  // namespace cliqg
