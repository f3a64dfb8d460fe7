// Device kernels of libcliqg for gfx950 (CDNA4). See kernels.hpp.
#include "kernels.hpp"

#include <cstdlib>
#include <type_traits>

namespace cliqg {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <class R> struct Acc4;
template <> struct Acc4<double> { using type = d4v; };
template <> struct Acc4<float> { using type = f4v; };

// 16x16x4 MFMA: lane l supplies A(l%16, l/16) and B(l/16, l%16). Result
// element r of lane l is C(4*(l/16) + r, l%16) for FP32 but C(l/16 + 4*r, l%16)
// for FP64 (measured on gfx950 with an in-bounds single-wave probe).
__device__ __forceinline__ d4v mfma16(double a, double b, d4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma16(float a, float b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> __device__ __forceinline__ T make_t(typename Scalar<T>::real r,
                                                       typename Scalar<T>::real i);
template <> __device__ __forceinline__ float make_t<float>(float r, float) { return r; }
template <> __device__ __forceinline__ double make_t<double>(double r, double) { return r; }
template <> __device__ __forceinline__ cfloat make_t<cfloat>(float r, float i) { return {r, i}; }
template <> __device__ __forceinline__ cdouble make_t<cdouble>(double r, double i) { return {r, i}; }

// 1 / a with one real division
__device__ __forceinline__ double pivot_inverse(double a) { return 1.0 / a; }
__device__ __forceinline__ cdouble pivot_inverse(cdouble a) {
    const double s = 1.0 / (a.x * a.x + a.y * a.y);
    return {a.x * s, -a.y * s};
}

__device__ __forceinline__ double real_part_only(double a) { return a; }
__device__ __forceinline__ cdouble real_part_only(cdouble a) { return {a.x, 0.0}; }

__device__ __forceinline__ int find_segment(const int* __restrict__ start, int n, int g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Load through the global address space. Operand pointers come from
// descriptor arrays, so the compiler cannot prove where they point and would
// emit flat loads, which also count against the LDS wait counter: every LDS
// wait in the GEMM's MFMA loop would then wait for the prefetch too.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float gload(const float* p) {
    return *(const __attribute__((address_space(1))) float*)(p);
}
__device__ __forceinline__ double gload(const double* p) {
    return *(const __attribute__((address_space(1))) double*)(p);
}
// complex entries are loaded as two-element vectors (a struct copy through
// a global pointer would still become a flat load)
__device__ __forceinline__ cfloat gload(const cfloat* p) {
    const f2v v = *(const __attribute__((address_space(1))) f2v*)(p);
    return cfloat{v.x, v.y};
}
__device__ __forceinline__ cdouble gload(const cdouble* p) {
    const d2v v = *(const __attribute__((address_space(1))) d2v*)(p);
    return cdouble{v.x, v.y};
}

__device__ __forceinline__ void gstore(float* p, float v) {
    *(__attribute__((address_space(1))) float*)(p) = v;
}
__device__ __forceinline__ void gstore(double* p, double v) {
    *(__attribute__((address_space(1))) double*)(p) = v;
}
__device__ __forceinline__ void gstore(cfloat* p, cfloat v) {
    *(__attribute__((address_space(1))) f2v*)(p) = f2v{v.x, v.y};
}
__device__ __forceinline__ void gstore(cdouble* p, cdouble v) {
    *(__attribute__((address_space(1))) d2v*)(p) = d2v{v.x, v.y};
}

// Select by value, component by component: a conditional between two
// lvalues of a struct type selects their addresses and sends the operands to
// scratch memory.
__device__ __forceinline__ float pick(bool c, float a, float b) { return c ? a : b; }
__device__ __forceinline__ double pick(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ cfloat pick(bool c, cfloat a, cfloat b) {
    return cfloat{c ? a.x : b.x, c ? a.y : b.y};
}
__device__ __forceinline__ cdouble pick(bool c, cdouble a, cdouble b) {
    return cdouble{c ? a.x : b.x, c ? a.y : b.y};
}

__device__ __forceinline__ void record_status(int* status, int code, int front, int col) {
    if (atomicCAS(status, 0, code) == 0) {
        status[1] = front;
        status[2] = col;
    }
}

constexpr int LPAD = 16;

// Three-multiplication (Gauss) complex products in the complex FP32 GEMM.
#ifndef CLIQG_GEMM_3M
#define CLIQG_GEMM_3M 1
#endif
constexpr bool GEMM_3M = CLIQG_GEMM_3M != 0;

// number of column-major lower-trapezoid tiles before tile column tj
__host__ __device__ __forceinline__ long long lower_prefix(long long tj, long long tilesM) {
    return tj * tilesM - tj * (tj - 1) / 2;
}

// Workgroup w of a launch runs on XCD w % 8, and each XCD has its own L2:
// give XCD x the x-th contiguous eighth of the launch's tiles, so the tiles
// an XCD holds at one time are neighbours that share operand panels.
__device__ __forceinline__ int xcd_remap(int w, int n) {
    const int q = n / 8, rem = n % 8, x = w % 8, l = w / 8;
    return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + l;
}

// Tiles of one descriptor in column groups of GROUP tile columns, row-major
// inside a group: consecutive tiles sweep GROUP columns of one tile row,
// then the next row. Lower launches keep only tiles with ti >= tj.
constexpr int GROUP = 8;
__device__ __forceinline__ void group_tile(int t, int tm, int tn, bool lower, int& ti, int& tj) {
    if (!lower) {
        const int g = t / (tm * GROUP), c0 = g * GROUP;
        const int w = min(GROUP, tn - c0), tl = t - g * tm * GROUP;
        ti = tl / w;
        tj = c0 + tl % w;
        return;
    }
    const int c = min(tm, tn);
    int c0 = 0;
    for (;;) {
        const int w = min(GROUP, c - c0);
        const int tri = w * (w + 1) / 2, cnt = tri + (tm - c0 - w) * w;
        if (t < cnt) {
            if (t < tri) {
                int r = (int)((sqrtf(8.0f * t + 1.0f) - 1.0f) * 0.5f);
                while (r > 0 && r * (r + 1) / 2 > t) --r;
                while ((r + 1) * (r + 2) / 2 <= t) ++r;
                ti = c0 + r;
                tj = c0 + (t - r * (r + 1) / 2);
            } else {
                const int u = t - tri;
                ti = c0 + w + u / w;
                tj = c0 + u % w;
            }
            return;
        }
        t -= cnt;
        c0 += GROUP;
    }
}

}  // namespace

int gemm_tile_m(int variant) {
    if (variant == TILE_STREAM) return 64 * STREAM_ROW_BLOCKS;
    return (variant == TILE_32 || variant == TILE_32_K64)
               ? 32
               : ((variant == TILE_128 || variant == TILE_128_8W) ? 128 : 64);
}
// a streaming tile takes all N <= 64 columns
int gemm_tile_n(int variant) { return variant == TILE_STREAM ? 64 : gemm_tile_m(variant); }

int gemm_tiles(int M, int N, int flags, int variant) {
    if (M <= 0 || N <= 0) return 0;
    const int TM = gemm_tile_m(variant), TN = gemm_tile_n(variant);
    const int tm = (M + TM - 1) / TM, tn = (N + TN - 1) / TN;
    if ((flags & G_LOWER) && TM != TN)
        throw Error(1, "lower-masked GEMM launches need square tiles");
    if (flags & G_LOWER) {
        const int c = tn < tm ? tn : tm;
        return (int)lower_prefix(c, tm);
    }
    return tm * tn;
}

// ---------------------------------------------------------------------------
// Grouped GEMM: one TM x TN output tile per workgroup of WM x WN waves;
// each wave owns a (TM/WM) x (TN/WN) part of 16x16 MFMA blocks. K is
// consumed in chunks of 16 through two LDS buffers (one barrier per chunk);
// the next chunk is prefetched into registers while the current one feeds
// the MFMAs. Complex products run on planar (re, im) tiles: three real MFMAs
// per block for complex FP32 (below), four for complex FP64.
// ---------------------------------------------------------------------------
template <class T, int TM, int TN, int ROLE, int TK, int WM = 2, int WN = 2>
__global__ __launch_bounds__(64 * WM * WN)
__attribute__((amdgpu_waves_per_eu((TM * TN / (WM * WN) >= 2048) ? 2
                                   : (sizeof(T) <= 8 ? 4 : 2), 8))) void k_gemm(const GemmDesc<T>* __restrict__ descs,
                                                       const int* __restrict__ tile_start,
                                                       int ndesc) {
    using R = typename Scalar<T>::real;
    using W = typename Scalar<T>::wide;
    using A4 = typename Acc4<R>::type;
    constexpr bool CPX = Scalar<T>::complex;
    constexpr int NP = CPX ? 2 : 1;
    constexpr int NT = 64 * WM * WN;                       // threads
    constexpr int WTM = TM / WM, WTN = TN / WN;            // wave tile
    constexpr int BM = WTM / 16, BN = WTN / 16;            // 16x16 blocks per wave
    constexpr int LA = TM * TK / NT, LB = TN * TK / NT;
    constexpr bool F64 = sizeof(R) == 8;
    constexpr bool SCALED = ROLE == 2;   // every descriptor of the launch has scaleK
    constexpr bool PRIO = !(ROLE == 3 && !F64);   // raised priority over the MFMA loop

    __shared__ R As[2][NP][TK][TM + LPAD];
    __shared__ R Bs[2][NP][TK][TN + LPAD];
    // The pivots of a chunk (scaleK) live in column TM of As, which is padding
    // the fragments never read: a separate array would push the complex tiles
    // past a quarter (complex float) or half (complex double) of the 160 KB
    // LDS and cost one resident block per CU.

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int di = find_segment(tile_start, ndesc, tile);
    const GemmDesc<T> d = descs[di];
    const int t = tile - tile_start[di];
    int ti, tj;
    group_tile(t, (d.M + TM - 1) / TM, (d.N + TN - 1) / TN, d.flags & G_LOWER, ti, tj);
    const int row0 = ti * TM, col0 = tj * TN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave % WM, wj = wave / WM;
    const bool transA = d.flags & G_TRANSA, transB = d.flags & G_TRANSB;
    const bool conjA = d.flags & G_CONJA, conjB = d.flags & G_CONJB;
    const W* scaleK = reinterpret_cast<const W*>(d.scaleK);

    // Complex FP32 products use three real MFMAs per block (Gauss):
    // P1 = sum aR bR, P2 = sum aI bI, P3 = sum (aR + aI)(bR + bI);
    // Re = P1 - P2, Im = P3 - P1 - P2. Complex FP64 keeps four.
    constexpr bool M3 = CPX && !F64 && GEMM_3M && BM * BN <= 8;   // registers: 64x32 wave tiles at most
    A4 accR[BM][BN], accI[BM][BN], accS[M3 ? BM : 1][M3 ? BN : 1];
#pragma unroll
    for (int a = 0; a < BM; ++a)
#pragma unroll
        for (int b = 0; b < BN; ++b) {
            accR[a][b] = A4{0, 0, 0, 0};
            accI[a][b] = A4{0, 0, 0, 0};
            if constexpr (M3) accS[a][b] = A4{0, 0, 0, 0};
        }

    // Fixed per-thread load geometry: element r of the A chunk is
    // (iA + r*dIA, kA + r*dKA) of the TM x TK tile, element r of the B chunk
    // (kB + r*dKB, jB + r*dJB) of the TK x TN tile; the global addresses of a
    // thread's elements differ by multiples of lda (ldb), and a chunk advances
    // them by a fixed stride.
    static_assert(NT % TM == 0 && NT % TN == 0 && NT % TK == 0, "load geometry");
    int iA, kA, dIA, dKA, jB, kB, dJB, dKB;
    if (!transA) { iA = tid % TM; kA = tid / TM; dIA = 0; dKA = NT / TM; }
    else { kA = tid % TK; iA = tid / TK; dIA = NT / TK; dKA = 0; }
    if (transB) { jB = tid % TN; kB = tid / TN; dJB = 0; dKB = NT / TN; }
    else { kB = tid % TK; jB = tid / TK; dJB = NT / TK; dKB = 0; }
    const int64_t stepA = transA ? TK : (int64_t)TK * d.lda;
    const int64_t stepB = transB ? (int64_t)TK * d.ldb : TK;
    unsigned rowOK = 0, colOK = 0;   // bit r: element r's row (column) is inside the matrix
#pragma unroll
    for (int r = 0; r < LA; ++r) rowOK |= (row0 + iA + r * dIA < d.M) ? 1u << r : 0u;
#pragma unroll
    for (int r = 0; r < LB; ++r) colOK |= (col0 + jB + r * dJB < d.N) ? 1u << r : 0u;
    const int ldsA = kA * (TM + LPAD) + iA, ldsAStep = dKA * (TM + LPAD) + dIA;
    const int ldsB = kB * (TN + LPAD) + jB, ldsBStep = dKB * (TN + LPAD) + dJB;

    T ra[LA], rb[LB];
    // element r of a thread sits r*offA elements after its first one (one of
    // dIA, dKA is zero); pA is the lane's pointer at the current chunk
    const int64_t offA = (int64_t)(transA ? dIA : dKA) * d.lda;
    const int64_t offB = (int64_t)(transB ? dKB : dJB) * d.ldb;
    const T* pA = nullptr;
    const T* pB = nullptr;
    auto start_at = [&](int k0) {
        pA = transA ? d.A + (k0 + kA) + (int64_t)(row0 + iA) * d.lda
                    : d.A + (row0 + iA) + (int64_t)(k0 + kA) * d.lda;
        pB = transB ? d.B + (col0 + jB) + (int64_t)(k0 + kB) * d.ldb
                    : d.B + (k0 + kB) + (int64_t)(col0 + jB) * d.ldb;
    };
    // Loads never branch per element: an element outside the matrix reads the
    // operand's first entry and is replaced by zero when it is stored to LDS,
    // so all of a chunk's loads are in flight together and the only wait is
    // before the store (a divergent load would wait inside its branch). Tiles
    // inside the matrix take a uniform fast path with no masks at all: every
    // VALU instruction here is issue time the MFMAs do not get.
    W rdw = one<W>();
    unsigned okA = 0, okB = 0;
    bool okD = false, fastA = false, fastB = false;
    const bool innerA = row0 + TM <= d.M, innerB = col0 + TN <= d.N;
    auto load = [&](int k0) {
        const bool fullK = k0 + TK <= d.K;
        if constexpr (SCALED) {
            okD = tid < TK && k0 + tid < d.K;
            rdw = gload(scaleK + (okD ? k0 + tid : 0));
        }
        fastA = innerA && fullK;
        fastB = innerB && fullK;
        if (fastA) {
#pragma unroll
            for (int r = 0; r < LA; ++r) ra[r] = gload(pA + r * offA);
        } else {
            okA = 0;
#pragma unroll
            for (int r = 0; r < LA; ++r) {
                const bool ok = ((rowOK >> r) & 1u) && (fullK || k0 + kA + r * dKA < d.K);
                okA |= ok ? 1u << r : 0u;
                ra[r] = gload(ok ? pA + r * offA : d.A);
            }
        }
        if (fastB) {
#pragma unroll
            for (int r = 0; r < LB; ++r) rb[r] = gload(pB + r * offB);
        } else {
            okB = 0;
#pragma unroll
            for (int r = 0; r < LB; ++r) {
                const bool ok = ((colOK >> r) & 1u) && (fullK || k0 + kB + r * dKB < d.K);
                okB |= ok ? 1u << r : 0u;
                rb[r] = gload(ok ? pB + r * offB : d.B);
            }
        }
        pA += stepA;
        pB += stepB;
    };
    // stage n elements v[] at p0[off + r*step] (re) and p1[...] (im), with the
    // mask and conjugation decided once per chunk (uniform branches)
    auto stage = [&](const T* v, int n, unsigned ok, bool fast, bool conj, R* p0, R* p1, int off,
                     int step) {
        if (fast && !conj) {
#pragma unroll
            for (int r = 0; r < (LA > LB ? LA : LB); ++r)
                if (r < n) {
                    p0[off + r * step] = re(v[r]);
                    if constexpr (CPX) p1[off + r * step] = im(v[r]);
                }
        } else {
#pragma unroll
            for (int r = 0; r < (LA > LB ? LA : LB); ++r)
                if (r < n) {
                    T x = (fast || ((ok >> r) & 1u)) ? v[r] : zero<T>();
                    if (conj) x = conj_(x);
                    p0[off + r * step] = re(x);
                    if constexpr (CPX) p1[off + r * step] = im(x);
                }
        }
    };
    auto store = [&](auto bufc) {
        constexpr int buf = decltype(bufc)::value;
        if constexpr (SCALED)
            if (tid < TK) {
                T dv = zero<T>();
                if (okD) dv = (d.flags & G_INVK) ? cvt<T>(one<W>() / rdw) : cvt<T>(rdw);
                As[buf][0][tid][TM] = re(dv);
                if constexpr (CPX) As[buf][NP - 1][tid][TM] = im(dv);
            }
        stage(ra, LA, okA, fastA, conjA, &As[buf][0][0][0], &As[buf][NP - 1][0][0], ldsA, ldsAStep);
        stage(rb, LB, okB, fastB, conjB, &Bs[buf][0][0][0], &Bs[buf][NP - 1][0][0], ldsB, ldsBStep);
    };

    // K range of this row tile: triangular op(A) bounds it by the tile's rows
    int kbeg = 0, kend = d.K;
    if (d.flags & G_KLO) kbeg = min((row0 / TK) * TK, d.K);
    if (d.flags & G_KHI) kend = min(d.K, row0 + TM);
    const int nk = kend > kbeg ? (kend - kbeg + TK - 1) / TK : 0;
    start_at(kbeg);
    if (nk > 0) {
        load(kbeg);
        store(std::integral_constant<int, 0>{});
        __syncthreads();
    }
    // One K chunk from LDS buffer BUF (a compile-time constant: the loop below
    // runs chunks in pairs, so LDS addresses fold into instruction offsets).
    auto chunk = [&](int c, auto bufc) {
        constexpr int buf = decltype(bufc)::value;
        if (c + 1 < nk) load(kbeg + (c + 1) * TK);
        // raised wave priority while feeding the MFMAs (measured on the
        // 120^3 complex factorizations: FP64 updates 1.5 % faster with it,
        // FP32 updates 1 % faster without); the TRSMs, on the factorization's
        // critical path beside the lookahead updates, one level higher
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(ROLE == 0 ? 2 : 1);
#pragma unroll
        for (int kk = 0; kk < TK / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            R aR[BM], aI[BM], bR[BN], bI[BN];
#pragma unroll
            for (int q = 0; q < BM; ++q) {
                aR[q] = As[buf][0][kr][wi * WTM + q * 16 + (lane & 15)];
                if constexpr (CPX) aI[q] = As[buf][NP - 1][kr][wi * WTM + q * 16 + (lane & 15)];
            }
            if constexpr (SCALED) {
                // op(A)(i,k) * D(k): the lane's k is kr for all of its A fragments
                const R dR = As[buf][0][kr][TM];
                const R dI = CPX ? As[buf][NP - 1][kr][TM] : R(0);
#pragma unroll
                for (int q = 0; q < BM; ++q) {
                    if constexpr (CPX) {
                        const R xr = aR[q] * dR - aI[q] * dI;
                        const R xi = aR[q] * dI + aI[q] * dR;
                        aR[q] = xr;
                        aI[q] = xi;
                    } else {
                        aR[q] = aR[q] * dR;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < BN; ++q) {
                bR[q] = Bs[buf][0][kr][wj * WTN + q * 16 + (lane & 15)];
                if constexpr (CPX) bI[q] = Bs[buf][NP - 1][kr][wj * WTN + q * 16 + (lane & 15)];
            }
            if constexpr (M3) {
                R sA[BM], sB[BN];
#pragma unroll
                for (int a = 0; a < BM; ++a) sA[a] = aR[a] + aI[a];
#pragma unroll
                for (int b = 0; b < BN; ++b) sB[b] = bR[b] + bI[b];
#pragma unroll
                for (int a = 0; a < BM; ++a)
#pragma unroll
                    for (int b = 0; b < BN; ++b) {
                        accR[a][b] = mfma16(aR[a], bR[b], accR[a][b]);   // P1
                        accI[a][b] = mfma16(aI[a], bI[b], accI[a][b]);   // P2
                        accS[a][b] = mfma16(sA[a], sB[b], accS[a][b]);   // P3
                    }
            } else {
#pragma unroll
                for (int a = 0; a < BM; ++a)
#pragma unroll
                    for (int b = 0; b < BN; ++b) {
                        accR[a][b] = mfma16(aR[a], bR[b], accR[a][b]);
                        if constexpr (CPX) {
                            accR[a][b] = mfma16(-aI[a], bI[b], accR[a][b]);
                            accI[a][b] = mfma16(aR[a], bI[b], accI[a][b]);
                            accI[a][b] = mfma16(aI[a], bR[b], accI[a][b]);
                        }
                    }
            }
        }
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(ROLE == 0 ? 1 : 0);
        if (c + 1 < nk) store(std::integral_constant<int, buf ^ 1>{});
        __syncthreads();
    };
    for (int c = 0; c < nk; c += 2) {
        chunk(c, std::integral_constant<int, 0>{});
        if (c + 1 < nk) chunk(c + 1, std::integral_constant<int, 1>{});
    }

    const W* scaleN = reinterpret_cast<const W*>(d.scaleN);
    const bool lower = d.flags & G_LOWER, accum = d.flags & G_ACCUM, neg = d.flags & G_NEG;
    const bool symm = lower && (d.flags & G_SYMM);
    // a lane's output columns do not depend on r: their pivots and the
    // pivots' reciprocals are formed once per 16-column block
    W sN[BN], rN[BN];
    if (scaleN) {
#pragma unroll
        for (int b = 0; b < BN; ++b) {
            const int gj = col0 + wj * WTN + b * 16 + (lane & 15);
            sN[b] = gload(scaleN + (gj < d.N ? gj : 0));
            rN[b] = pivot_inverse(sN[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < BM; ++a)
#pragma unroll
        for (int b = 0; b < BN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rin = F64 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                const int gi = row0 + wi * WTM + a * 16 + rin;
                const int gj = col0 + wj * WTN + b * 16 + (lane & 15);
                if (gi >= d.M || gj >= d.N || (lower && gi < gj)) continue;
                T v;
                if constexpr (M3)
                    v = make_t<T>(accR[a][b][r] - accI[a][b][r],
                                  accS[a][b][r] - accR[a][b][r] - accI[a][b][r]);
                else
                    v = make_t<T>(accR[a][b][r], CPX ? accI[a][b][r] : R(0));
                if (scaleN) {
                    v = cvt<T>(cvt<W>(v) * rN[b]);
                    // W = L D from the rounded L, exactly as a pivot-scaled
                    // update would form it, so updates from W stay consistent
                    // with the stored factor
                    if (d.C2) gstore(d.C2 + gi + (int64_t)gj * d.ldc, v * cvt<T>(sN[b]));
                }
                if (neg) v = -v;
                T* p = d.C + gi + (int64_t)gj * d.ldc;
                if (accum) v = gload(p) + v;
                if (symm) {
                    // Hermitian output (conj mirror) has a real diagonal
                    if (conjA && gi == gj) v = make_t<T>(re(v), R(0));
                    if (gi > gj) gstore(d.C + gj + (int64_t)gi * d.ldc, conjA ? conj_(v) : v);
                }
                gstore(p, v);
            }
}

// eight consecutive FP32-type entries through 16-byte global loads
__device__ __forceinline__ void gload8(const float* p, float* dst) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(1))) f4* q = (const __attribute__((address_space(1))) f4*)(p);
    const f4 a = q[0], b = q[1];
    dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
    dst[4] = b.x; dst[5] = b.y; dst[6] = b.z; dst[7] = b.w;
}
__device__ __forceinline__ void gload8(const cfloat* p, cfloat* dst) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(1))) f4* q = (const __attribute__((address_space(1))) f4*)(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const f4 a = q[i];
        dst[2 * i] = cfloat{a.x, a.y};
        dst[2 * i + 1] = cfloat{a.z, a.w};
    }
}

// ---------------------------------------------------------------------------
// Streaming product of the solves (tile shape TILE_STREAM, FP32 working
// types): C (M x N, N <= 64) (+)= +-op(A) B, where A is the factor (read once
// per solve) and B a block of right-hand sides. A workgroup owns 64 rows of C
// and all of its columns; wave w the rows 16w..16w+15. A goes from global
// memory straight into the MFMA operand registers (no LDS staging, no
// barrier on its path); only the K x N block of B passes through LDS, where
// the four waves share it. K runs in chunks of 32, permuted inside a chunk
// so that a lane's eight elements of op(A) are consecutive in k: MFMA block
// s (k-step of 4) takes k = k0 + 8 g + s from lane group g = lane / 16, for
// A and B alike, which leaves the sum unchanged. For op(A) = A^T (the
// backward sweep) those eight elements are contiguous in memory. The next
// chunk of A and B is loaded while the current one feeds the MFMAs.
// Complex products use three real MFMAs per block (Gauss), B kept in LDS as
// planes re, im and re + im.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_gemm_stream(
    const GemmDesc<T>* __restrict__ descs, const int* __restrict__ tile_start, int ndesc) {
    static_assert(sizeof(typename Scalar<T>::real) == 4, "FP32 working types only");
    constexpr bool CPX = Scalar<T>::complex;
    constexpr int NPL = CPX ? 3 : 1;   // planes of B: re (, im, re + im)
    constexpr int KC = 32;             // chunk depth
    constexpr int KS = KC + 4;         // LDS row stride (floats) of a column of B
    constexpr int RB = STREAM_ROW_BLOCKS;   // 16-row blocks per wave
    __shared__ float Bs[2][NPL][64][KS];

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int di = find_segment(tile_start, ndesc, tile);
    const GemmDesc<T> d = descs[di];
    const int row0 = (tile - tile_start[di]) * (64 * RB);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, li = lane & 15;
    const int ncb = (d.N + 15) >> 4;   // 16-column blocks of C
    const bool transA = d.flags & G_TRANSA, conjA = d.flags & G_CONJA;
    // K range of this row tile: a triangular op(A) (the solves' inverted
    // blocks, whose other triangle is stored as zeros) bounds it by the
    // tile's rows; K is the range's end, k0 runs from kbeg
    int kbeg = 0, K = d.K;
    if (d.flags & G_KLO) kbeg = min((row0 / KC) * KC, K);
    if (d.flags & G_KHI) K = min(K, row0 + 64 * RB);

    // this lane's rows of op(A): op(A)(r, k) = A[r + k lda], or A[k + r lda]
    const int wrow0 = row0 + wave * 16 * RB;
    bool rowOK[RB];
    const T* pa[RB];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const int arow = wrow0 + 16 * b + li;
        rowOK[b] = arow < d.M;
        pa[b] = d.A + (rowOK[b] ? (transA ? (int64_t)arow * d.lda : (int64_t)arow) : 0);
    }
    const int64_t astep = transA ? 1 : d.lda;
    // B: thread t loads k = 8 (t & 3) .. +7 of column t >> 2
    const int bj = tid >> 2, bq = tid & 3;
    const bool colOK = bj < d.N;
    const T* pb = d.B + (colOK ? (int64_t)bj * d.ldb : 0);

    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 acc[4][RB][NPL];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int p = 0; p < NPL; ++p) acc[c][b][p] = f4{0.f, 0.f, 0.f, 0.f};

    T ra[RB][8], rn[RB][8], rb[8];
    // Loads never branch per element: an element outside the matrix reads
    // the operand's first entry and becomes zero afterwards.
    auto loadA = [&](int k0, T (*dst)[8]) {
        const int kb = k0 + 8 * g;
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            if (rowOK[b] && kb + 8 <= K) {
                if (transA) {
                    gload8(pa[b] + kb, dst[b]);   // eight consecutive entries of a column of A
                } else {
#pragma unroll
                    for (int s = 0; s < 8; ++s) dst[b][s] = gload(pa[b] + (int64_t)(kb + s) * d.lda);
                }
                continue;
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const bool ok = rowOK[b] && kb + s < K;
                dst[b][s] = gload(pa[b] + (ok ? (int64_t)(kb + s) * astep : 0));
            }
#pragma unroll
            for (int s = 0; s < 8; ++s)
                if (!(rowOK[b] && kb + s < K)) dst[b][s] = zero<T>();
        }
    };
    auto loadB = [&](int k0) {
        const int kb = k0 + 8 * bq;
        if (colOK && kb + 8 <= K) {
            gload8(pb + kb, rb);
            return;
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const bool ok = colOK && kb + s < K;
            rb[s] = gload(pb + (ok ? kb + s : 0));
        }
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (!(colOK && kb + s < K)) rb[s] = zero<T>();
    };
    auto storeB = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            Bs[buf][0][bj][8 * bq + s] = re(rb[s]);
            if constexpr (CPX) {
                Bs[buf][1][bj][8 * bq + s] = im(rb[s]);
                Bs[buf][2][bj][8 * bq + s] = re(rb[s]) + im(rb[s]);
            }
        }
    };

    if (kbeg < K) {
        loadA(kbeg, ra);
        loadB(kbeg);
        storeB(0);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < K; k0 += KC) {
        const bool more = k0 + KC < K;
        if (more) {
            loadA(k0 + KC, rn);
            loadB(k0 + KC);
        }
        // s outer, column blocks inside: 4 x RB x 3 independent accumulators
        // between dependent MFMAs, few B values live at a time
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            float aR[RB], aI[RB];
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                aR[b] = re(ra[b][s]);
                if constexpr (CPX) aI[b] = conjA ? -im(ra[b][s]) : im(ra[b][s]);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (c >= ncb) break;
                const float bR = Bs[buf][0][16 * c + li][8 * g + s];
                if constexpr (CPX) {
                    const float bI = Bs[buf][1][16 * c + li][8 * g + s];
                    const float bS = Bs[buf][2][16 * c + li][8 * g + s];
#pragma unroll
                    for (int b = 0; b < RB; ++b) {
                        acc[c][b][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[b], bR, acc[c][b][0], 0, 0, 0);
                        acc[c][b][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(aI[b], bI, acc[c][b][1], 0, 0, 0);
                        acc[c][b][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[b] + aI[b], bS, acc[c][b][2], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int b = 0; b < RB; ++b)
                        acc[c][b][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(aR[b], bR, acc[c][b][0], 0, 0, 0);
                }
            }
        }
        if (more) {
            storeB(buf ^ 1);
#pragma unroll
            for (int b = 0; b < RB; ++b)
#pragma unroll
                for (int s = 0; s < 8; ++s) ra[b][s] = rn[b][s];
        }
        __syncthreads();
        buf ^= 1;
    }
    // element r of lane l of a 16x16 block is C(4 (l / 16) + r, l % 16)
    const bool neg = d.flags & G_NEG, accum = d.flags & G_ACCUM;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (c >= ncb) break;
        const int col = 16 * c + li;
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wrow0 + 16 * b + 4 * g + r;
                if (row >= d.M || col >= d.N) continue;
                T v;
                if constexpr (CPX) {
                    const float p1 = acc[c][b][0][r], p2 = acc[c][b][1][r], p3 = acc[c][b][2][r];
                    v = make_t<T>(p1 - p2, p3 - p1 - p2);
                } else {
                    v = make_t<T>(acc[c][b][0][r], 0.f);
                }
                if (neg) v = -v;
                T* pc = d.C + row + (int64_t)col * d.ldc;
                if (accum) v = *pc + v;
                *pc = v;
            }
    }
}

template <class T, int TM, int TN, int KT = 16, int WM = 2, int WN = 2>
static void launch_gemm_shape(const GemmDesc<T>* d, const int* ts, int nd, int nt, int role,
                              hipStream_t s) {
    const dim3 block(64 * WM * WN);
    if (role == 0)
        hipLaunchKernelGGL((k_gemm<T, TM, TN, 0, KT, WM, WN>), dim3(nt), block, 0, s, d, ts, nd);
    else if (role == 2)
        hipLaunchKernelGGL((k_gemm<T, TM, TN, 2, KT, WM, WN>), dim3(nt), block, 0, s, d, ts, nd);
    else if (role == 3)
        hipLaunchKernelGGL((k_gemm<T, TM, TN, 3, KT, WM, WN>), dim3(nt), block, 0, s, d, ts, nd);
    else
        hipLaunchKernelGGL((k_gemm<T, TM, TN, 1, KT, WM, WN>), dim3(nt), block, 0, s, d, ts, nd);
}

template <class T>
void launch_gemm(const GemmDesc<T>* d_descs, const int* d_tile_start, int ndesc, int ntiles,
                 int variant, int role, hipStream_t s) {
    if (ntiles <= 0 || ndesc <= 0) return;
    if (variant == TILE_128) {
        if constexpr (gemm_variant_allowed<T>(TILE_128))
            launch_gemm_shape<T, 128, 128, 16, 2, 2>(d_descs, d_tile_start, ndesc, ntiles, role, s);
        else
            throw Error(1, "128x128 GEMM tiles are built for real types only");
    } else if (variant == TILE_128_8W) {
        if constexpr (gemm_variant_allowed<T>(TILE_128_8W))
            launch_gemm_shape<T, 128, 128, 16, 2, 4>(d_descs, d_tile_start, ndesc, ntiles, role, s);
        else
            throw Error(1, "eight-wave 128x128 GEMM tiles are built for complex FP32 only");
    } else if (variant == TILE_STREAM) {
        if constexpr (gemm_variant_allowed<T>(TILE_STREAM))
            hipLaunchKernelGGL((k_gemm_stream<T>), dim3(ntiles), dim3(256), 0, s, d_descs,
                               d_tile_start, ndesc);
        else
            throw Error(1, "streaming GEMM tiles are built for FP32 working types only");
    } else if (variant == TILE_32) {
        launch_gemm_shape<T, 32, 32>(d_descs, d_tile_start, ndesc, ntiles, role, s);
    } else if (variant == TILE_32_K64) {
        if constexpr (gemm_variant_allowed<T>(TILE_32_K64))
            launch_gemm_shape<T, 32, 32, 64>(d_descs, d_tile_start, ndesc, ntiles, role, s);
        else
            throw Error(1, "64-deep 32x32 GEMM tiles are not built for complex FP64");
    } else {
        launch_gemm_shape<T, 64, 64>(d_descs, d_tile_start, ndesc, ntiles, role, s);
    }
}

// ---------------------------------------------------------------------------
// Clearing of update matrices (one ZERO_CHUNK-entry chunk per workgroup)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_zero_segments(const ZeroSeg<T>* __restrict__ segs,
                                                       const int* __restrict__ chunk_start,
                                                       int nseg) {
    const int g = blockIdx.x;
    const int si = find_segment(chunk_start, nseg, g);
    const ZeroSeg<T> sg = segs[si];
    const int64_t base = (int64_t)(g - chunk_start[si]) * ZERO_CHUNK;
    const int64_t end = min((int64_t)ZERO_CHUNK, sg.n - base);
    T* p = sg.p + base;
    for (int64_t i = threadIdx.x; i < end; i += blockDim.x) gstore(p + i, zero<T>());
}

// one wave per column, four per workgroup
template <class T>
__global__ __launch_bounds__(256) void k_zero_lower(const ZeroLowerDesc<T>* __restrict__ descs,
                                                    const int* __restrict__ col_start, int nd,
                                                    int ncols) {
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= ncols) return;
    const int c = find_segment(col_start, nd, g);
    const ZeroLowerDesc<T> d = descs[c];
    const int j = g - col_start[c];
    T* col = d.p + (int64_t)j * d.ld;
    for (int i = j + (threadIdx.x & 63); i < d.rows; i += 64) gstore(col + i, zero<T>());
}

template <class T>
void launch_zero_lower(const ZeroLowerDesc<T>* d_descs, const int* d_col_start, int ndesc,
                       int ncols, hipStream_t s) {
    if (ncols <= 0 || ndesc <= 0) return;
    hipLaunchKernelGGL(k_zero_lower<T>, dim3((ncols + 3) / 4), dim3(256), 0, s, d_descs,
                       d_col_start, ndesc, ncols);
}

template <class T>
void launch_zero_segments(const ZeroSeg<T>* d_segs, const int* d_chunk_start, int nseg,
                          int nchunks, hipStream_t s) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(k_zero_segments<T>, dim3(nchunks), dim3(256), 0, s, d_segs, d_chunk_start,
                       nseg);
}

template <class T>
__global__ __launch_bounds__(256) void k_copy2d(const CopyDesc<T>* __restrict__ descs,
                                                const int* __restrict__ chunk_start, int ndesc) {
    const int g = blockIdx.x;
    const int di = find_segment(chunk_start, ndesc, g);
    const CopyDesc<T> d = descs[di];
    const int64_t total = (int64_t)d.rows * d.cols;
    const int64_t base = (int64_t)(g - chunk_start[di]) * 1024;
    // all four loads before the stores (source and destination may share a
    // pool, so the compiler would otherwise wait a round trip per entry)
    T v[4];
    int64_t at[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t e = base + threadIdx.x + 256 * q;
        const bool ok = e < total;
        const int64_t r = ok ? e % d.rows : 0, c = ok ? e / d.rows : 0;
        at[q] = ok ? r + c * d.ldd : -1;
        v[q] = d.src[r + c * d.lds];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (at[q] >= 0) d.dst[at[q]] = v[q];
}

template <class T>
void launch_copy2d(const CopyDesc<T>* d_descs, const int* d_chunk_start, int ndesc, int nchunks,
                   hipStream_t s) {
    if (nchunks <= 0 || ndesc <= 0) return;
    hipLaunchKernelGGL(k_copy2d<T>, dim3(nchunks), dim3(256), 0, s, d_descs, d_chunk_start, ndesc);
}

// One 1024-element chunk of one split product per workgroup; each element
// sums its slices in order, so results do not depend on scheduling.
template <class T>
__global__ __launch_bounds__(256) void k_reduce_slices(const ReduceDesc<T>* __restrict__ descs,
                                                       const int* __restrict__ chunk_start,
                                                       int ndesc) {
    const int g = blockIdx.x;
    const int di = find_segment(chunk_start, ndesc, g);
    const ReduceDesc<T> d = descs[di];
    const int64_t total = (int64_t)d.rows * d.cols;
    const int64_t base = (int64_t)(g - chunk_start[di]) * 1024;
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) {
        const int64_t e = base + i;
        if (e >= total) break;
        T v = gload(d.ws + e);
        for (int sl = 1; sl < d.slices; ++sl) v = v + gload(d.ws + sl * total + e);
        if (d.flags & G_NEG) v = -v;
        T* p = d.C + e % d.rows + (e / d.rows) * d.ldc;
        if (d.flags & G_ACCUM) v = gload(p) + v;
        gstore(p, v);
    }
}

template <class T>
void launch_reduce_slices(const ReduceDesc<T>* d_descs, const int* d_chunk_start, int ndesc,
                          int nchunks, hipStream_t s) {
    if (nchunks <= 0 || ndesc <= 0) return;
    hipLaunchKernelGGL(k_reduce_slices<T>, dim3(nchunks), dim3(256), 0, s, d_descs, d_chunk_start,
                       ndesc);
}

// ---------------------------------------------------------------------------
// Assembly of A into the fronts (scatter through a precomputed map)
// ---------------------------------------------------------------------------
template <class T, class V>
__global__ void k_assemble(const V* __restrict__ vals, const int64_t* __restrict__ dst,
                           int64_t nnz, T* __restrict__ pool, bool conj) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = dst[e];
        if (t >= 0) pool[t] = cvt<T>(cj(vals[e], conj));
    }
}

template <class T, class V>
void launch_assemble(const V* vals, const int64_t* dst, int64_t nnz, T* pool, bool conj,
                     hipStream_t s) {
    if (nnz <= 0) return;
    int64_t blocks = (nnz + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((k_assemble<T, V>), dim3((unsigned)blocks), dim3(256), 0, s, vals, dst,
                       nnz, pool, conj);
}

// ---------------------------------------------------------------------------
// The solves' inverted diagonal block of a small front (n <= 2 NB) in one
// workgroup: LW(0:n0, 0:n0) = X0, LW(0:n0, n0:n) = 0, LW(n0:n, n0:n) = X1 and
// LW(n0:n, 0:n0) = -X1 (L10 X0), the two 64 x 64 products with both operands
// in LDS and a 4 x 4 register tile per thread. One launch per level replaces
// the clearing, the two block copies and the two grouped products of the
// general path (wide blocks of larger fronts keep that path).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_wide_inverse_small(const WideInvDesc<T>* __restrict__ descs) {
    // rows padded to a multiple of four entries: a thread's four consecutive
    // entries are one 16-byte aligned LDS read (or two, complex)
    __shared__ __attribute__((aligned(16))) T Sa[NB][NB + 4];   // Sa[k][i] = A(i, k)
    __shared__ __attribute__((aligned(16))) T Sb[NB][NB + 4];   // Sb[k][j] = B(k, j)
    constexpr int PT = NB * NB / 256;   // entries of a 64 x 64 block per thread
    const WideInvDesc<T> d = descs[blockIdx.x];
    const int n = d.n, tid = threadIdx.x;
    const int n0 = min(n, NB), b1 = n - n0;
    // Every copy issues all of a thread's loads before its first store (a
    // load followed by its store per iteration would wait a global round
    // trip per entry). Element t of a thread is (i, j) = (e % 64, e / 64)
    // with e = tid + 256 t: lanes run down a column.
    T v[PT], w[PT];
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
        v[t] = gload(d.X0 + i + j * NB);
    }
    if (b1 > 0) {
#pragma unroll
        for (int t = 0; t < PT; ++t) {
            const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
            w[t] = gload(d.X1 + i + j * NB);
        }
    }
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
        if (i < n0 && j < n0) d.LW[i + (int64_t)j * n] = v[t];
    }
    if (b1 <= 0) return;
    T l[PT];
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
        const bool ok = i < b1;
        l[t] = gload(d.L10 + (ok ? i + (int64_t)j * d.ldl : 0));
    }
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
        // zero block above X1, X1 itself
        if (j < b1) d.LW[i + (int64_t)(NB + j) * n] = zero<T>();
        if (i < b1 && j < b1) d.LW[(NB + i) + (int64_t)(NB + j) * n] = w[t];
        // T1 = L10 X0: Sa[k][i] = L10(i, k) (rows >= b1 read as zero), Sb[k][j] = X0(k, j)
        Sa[j][i] = (i < b1) ? l[t] : zero<T>();
        Sb[i][j] = v[t];
    }
    __syncthreads();
    const int ti = 4 * (tid & 15), tj = 4 * (tid >> 4);
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero<T>();
    for (int k2 = 0; k2 < NB; ++k2) {
        T av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = Sa[k2][ti + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = Sb[k2][tj + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
    // X10 = -X1 T1: Sa[k][i] = X1(i, k), Sb[k][j] = T1(k, j)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Sb[ti + a][tj + b] = acc[a][b];
#pragma unroll
    for (int t = 0; t < PT; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), j = e / NB;
        Sa[j][i] = w[t];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero<T>();
    for (int k2 = 0; k2 < NB; ++k2) {
        T av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = Sa[k2][ti + a];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = Sb[k2][tj + b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (ti + a < b1) d.LW[(NB + ti + a) + (int64_t)(tj + b) * n] = -acc[a][b];
}

template <class T>
void launch_wide_inverse_small(const WideInvDesc<T>* d_descs, int ndesc, hipStream_t s) {
    if (ndesc <= 0) return;
    hipLaunchKernelGGL(k_wide_inverse_small<T>, dim3(ndesc), dim3(256), 0, s, d_descs);
}

// ---------------------------------------------------------------------------
// Extend-add: one wave per pair of child update columns (j0 + p and
// j0 + nc - 1 - p, so every wave adds about the same number of entries),
// four waves per workgroup. Children of one slot have distinct parents, so no
// two waves touch the same entry. Each lane issues the loads of four rows
// before their stores (the pools may alias, so the compiler would not).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void extend_column(const ExtendDesc<T>& d, int j, int lane) {
    const int J = d.rel[j];
    const T* col = d.Uc + (int64_t)j * d.uc;
    T* dst;
    if (J < d.np) dst = d.Lp + (int64_t)J * d.mp;
    else dst = d.Up + (int64_t)(J - d.np) * d.up - d.np;
    for (int i = j + lane; i < d.uc; i += 256) {
        int r[4];
        T v[4], w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + 64 * q < d.uc) {
                r[q] = d.rel[i + 64 * q];
                v[q] = col[i + 64 * q];
            }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + 64 * q < d.uc) w[q] = dst[r[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + 64 * q < d.uc) dst[r[q]] = w[q] + v[q];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_extend_add(const ExtendDesc<T>* __restrict__ descs,
                                                    const int* __restrict__ pair_start, int nd,
                                                    int npairs) {
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= npairs) return;
    const int lane = threadIdx.x & 63;
    const int c = find_segment(pair_start, nd, g);
    const ExtendDesc<T> d = descs[c];
    const int p = g - pair_start[c];
    const int j = d.j0 + p, j2 = d.j0 + d.nc - 1 - p;
    extend_column(d, j, lane);
    if (j2 != j) extend_column(d, j2, lane);
}

template <class T>
void launch_extend_add(const ExtendDesc<T>* d_descs, const int* d_col_start, int ndesc,
                       int ncols, hipStream_t s) {
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_extend_add<T>, dim3((ncols + 3) / 4), dim3(256), 0, s, d_descs,
                       d_col_start, ndesc, ncols);
}

// ---------------------------------------------------------------------------
// Fused panel step of the LDL solves' triangular sweeps over the inverted
// SB-wide diagonal blocks (FrontLowerForwardSolve / FrontLowerBackwardSolve,
// lower_solve/local_front.hpp:35-82): one launch per panel of a level instead
// of two dependent products. Every workgroup forms, for 16 right-hand sides,
//     forward:  Y = Linv_pp X(p),         backward: Y = op(Linv_pp)^T X2(p)
// (Linv_pp lower triangular: only the triangle's K range is read); the
// panel's own workgroups write Y as the panel's rows of the output, the band
// workgroups apply to 64 rows r of the band
//     forward:  X(r) -= L(r, p) Y,        backward: X2(r) -= op(L(p, r))^T Y.
// Y is recomputed by each workgroup (no dependency inside the launch); the
// host keeps the band to the panels of one super-panel and updates the rows
// past it with one grouped product per super-panel.
// 512 threads: in the Y product wave w owns rows 16w..16w+15; in the band
// product wave w owns rows 16(w & 3).. of the 64 and half (w >> 2) of K, the
// two halves summed through LDS. Complex FP32 uses three real MFMAs per
// block as the grouped GEMM does.
// ---------------------------------------------------------------------------
template <class T, bool BWD, bool CONJ>
__global__ __launch_bounds__(512) void k_panel_step(const PanelStepDesc<T>* __restrict__ descs,
                                                    const int* __restrict__ task_start, int nd) {
    using R = typename Scalar<T>::real;
    using A4 = typename Acc4<R>::type;
    constexpr bool CPX = Scalar<T>::complex;
    constexpr bool F64 = sizeof(R) == 8;
    constexpr bool M3 = CPX && !F64 && GEMM_3M;
    constexpr int NP = CPX ? 2 : 1;
    constexpr int NA = M3 ? 3 : (CPX ? 2 : 1);   // accumulators per block
    constexpr int PB = PANEL_MAX;
    constexpr int K2 = PB / 8;                   // 4-deep steps of one K half of the band product
    __shared__ R Bs[NP][PB][16];     // the panel's rows of the input
    __shared__ R Ys[NP][PB][16];     // Y
    __shared__ R Red[4][NA][64][4];  // K-half partial sums of the band product

    const int t = blockIdx.x;
    const int di = find_segment(task_start, nd, t);
    const PanelStepDesc<T> d = descs[di];
    const int cg = (d.k + 15) / 16;
    const int loc = t - task_start[di];
    const bool owner = loc < cg;
    const int c16 = owner ? loc : (loc - cg) % cg;
    const int rb = owner ? 0 : (loc - cg) / cg;
    const int c0 = 16 * c16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sbw = d.sbw;
    auto opc = [](T v) { return CONJ ? conj_(v) : v; };
    // fragment layout of the 16x16x4 MFMAs: lane holds A(l % 16, k0 + l / 16),
    // B(k0 + l / 16, l % 16) and, of the 16x16 result, rows rin(r), column l % 16
    const int fr = lane & 15, fk = lane >> 4;
    auto rin = [&](int r) { return F64 ? fk + 4 * r : 4 * fk + r; };

    // Loads that do not depend on Y are all issued first, none behind a
    // branch (a value merged after a branch waits for its load there; an
    // out-of-range element reads the block's first entry and is replaced by
    // zero when used): the panel's rows of the input, the band fragments
    // op(L)(r, k) of this wave, and the band rows of X this wave updates.
    T xb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int idx = tid + 512 * e, kr = idx >> 4, j = idx & 15;
        const bool ok = kr < sbw && c0 + j < d.k;
        xb[e] = gload(d.Xin + (ok ? (d.w0 + kr) + (int64_t)(c0 + j) * d.ldx : 0));
    }
    const int r0 = d.b0 + 64 * rb;
    const int half = wave >> 2, blk = wave & 3;
    const int r2 = r0 + 16 * blk + fr;
    const int kb2 = 64 * half, ke2 = min(sbw, kb2 + 64);
    T a2[K2];
    unsigned ok2 = 0;
#pragma unroll
    for (int q = 0; q < K2; ++q) {
        const int kA = kb2 + 4 * q + fk;
        const bool ok = !owner && r2 < d.b1 && kA < ke2;
        const int64_t off = BWD ? kA + (int64_t)r2 * d.ldl : r2 + (int64_t)kA * d.ldl;
        a2[q] = gload(d.Lp + (ok ? off : 0));
        ok2 |= ok ? 1u << q : 0u;
    }
    T xo[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = r0 + 16 * blk + rin(q);
        const bool ok = !owner && half == 0 && row < d.b1 && c0 + fr < d.k;
        xo[q] = gload(d.Xin + (ok ? row + (int64_t)(c0 + fr) * d.ldx : 0));
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int idx = tid + 512 * e, kr = idx >> 4, j = idx & 15;
        const T x = pick(kr < sbw && c0 + j < d.k, xb[e], zero<T>());
        Bs[0][kr][j] = re(x);
        if constexpr (CPX) Bs[NP - 1][kr][j] = im(x);
    }
    __syncthreads();

    auto mac = [&](A4 (&acc)[3], T a, int kr, R (*B)[PB][16]) {
        const R aR = re(a), aI = CPX ? im(a) : R(0);
        const R bR = B[0][kr][fr], bI = CPX ? B[NP - 1][kr][fr] : R(0);
        if constexpr (M3) {
            acc[0] = mfma16(aR, bR, acc[0]);
            acc[1] = mfma16(aI, bI, acc[1]);
            acc[2] = mfma16(aR + aI, bR + bI, acc[2]);
        } else {
            acc[0] = mfma16(aR, bR, acc[0]);
            if constexpr (CPX) {
                acc[0] = mfma16(-aI, bI, acc[0]);
                acc[1] = mfma16(aR, bI, acc[1]);
                acc[1] = mfma16(aI, bR, acc[1]);
            }
        }
    };
    auto value = [&](const A4 (&acc)[3], int r) -> T {
        if constexpr (M3) return make_t<T>(acc[0][r] - acc[1][r], acc[2][r] - acc[0][r] - acc[1][r]);
        else return make_t<T>(acc[0][r], CPX ? acc[1][r] : R(0));
    };

    // Y: wave w, rows 16w..16w+15, K over the triangle in 16-deep chunks,
    // the next chunk's fragments loaded while the current one is multiplied
    {
        A4 acc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[a] = A4{0, 0, 0, 0};
        const int i1 = 16 * wave + fr;
        const int kb = BWD ? 16 * wave : 0;
        const int ke = 16 * wave < sbw ? (BWD ? sbw : min(sbw, 16 * wave + 16)) : kb;
        auto fetch = [&](T (&av)[4], int k0) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kA = k0 + 4 * q + fk;
                const bool ok = i1 < sbw && kA < ke;
                const int64_t off = BWD ? kA + (int64_t)i1 * sbw : i1 + (int64_t)kA * sbw;
                const T v = gload(d.LW + (ok ? off : 0));
                av[q] = pick(ok, BWD ? opc(v) : v, zero<T>());
            }
        };
        T cur[4], nxt[4];
        if (kb < ke) fetch(cur, kb);
        for (int k0 = kb; k0 < ke; k0 += 16) {
            if (k0 + 16 < ke) fetch(nxt, k0 + 16);
#pragma unroll
            for (int q = 0; q < 4; ++q) mac(acc, cur[q], k0 + 4 * q + fk, Bs);
#pragma unroll
            for (int q = 0; q < 4; ++q) cur[q] = nxt[q];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * wave + rin(r);
            const T y = value(acc, r);
            Ys[0][row][fr] = re(y);
            if constexpr (CPX) Ys[NP - 1][row][fr] = im(y);
            if (owner && row < sbw && c0 + fr < d.k)
                gstore(d.Xout + (d.w0 + row) + (int64_t)(c0 + fr) * d.ldx, y);
        }
    }
    if (owner) return;
    __syncthreads();

    // band rows r0 .. r0 + 63 (clipped to b1): X(r) -= op(L) Y
    A4 acc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = A4{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < K2; ++q) {
        const int k4 = kb2 + 4 * q;
        if (k4 < ke2) {   // uniform across the wave
            const T a = pick((ok2 >> q) & 1u, BWD ? opc(a2[q]) : a2[q], zero<T>());
            mac(acc, a, k4 + fk, Ys);
        }
    }
    if (half == 1) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) Red[blk][a][lane][q] = acc[a][q];
    }
    __syncthreads();
    if (half == 1) return;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[a][q] += Red[blk][a][lane][q];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = r0 + 16 * blk + rin(q);
        if (row >= d.b1 || c0 + fr >= d.k) continue;
        gstore(d.Xin + row + (int64_t)(c0 + fr) * d.ldx, xo[q] - value(acc, q));
    }
}

template <class T>
void launch_panel_step(const PanelStepDesc<T>* d, const int* task_start, int nd, int ntasks,
                       bool backward, bool conj, hipStream_t s) {
    if (nd <= 0 || ntasks <= 0) return;
    if (!backward)
        hipLaunchKernelGGL((k_panel_step<T, false, false>), dim3(ntasks), dim3(512), 0, s, d,
                           task_start, nd);
    else if (conj)
        hipLaunchKernelGGL((k_panel_step<T, true, true>), dim3(ntasks), dim3(512), 0, s, d,
                           task_start, nd);
    else
        hipLaunchKernelGGL((k_panel_step<T, true, false>), dim3(ntasks), dim3(512), 0, s, d,
                           task_start, nd);
}

// ---------------------------------------------------------------------------
// Diagonal block: unpivoted LDL^T of a b x b block (b <= NB) in FP64 (field
// type) and the inverse of its unit lower factor. One workgroup of 256
// threads per front. Thread (ti, tj) = (tid % 16, tid / 16) keeps the
// elements M[ti + 16 r][tj + 16 q], r, q < 4, of a 64 x 64 register-resident
// matrix (a cyclic layout: the columns already finished at step j are whole
// slots q with 16 q + 15 <= j, which every wave skips). The lower triangle
// holds the Schur-updated block (its column j is the unscaled L column at
// step j), the strict upper triangle the transpose of the inverse being
// built, li(c, i) at M[i][c]. Step j broadcasts column j of M through LDS
// (double buffered, one barrier per step):
//     a(i,c)  -= (a(i,j) / d_j) op(a(c,j))      lower, i >= c > j
//     li(c,i) -= u_i (a(c,j) / d_j)             upper, i <= j < c,
// with u_i = li(j,i) = M[i][j] for i < j and u_j = 1. Slots r > q hold only
// lower elements and r < q only upper ones; the slot r == q is lower in the
// threads with ti >= tj.
// ---------------------------------------------------------------------------

// Each workgroup is latency-bound (one barrier per pivot step), so a level
// of many small fronts needs many resident workgroups: the real
// instantiations (about 66 registers) may run eight waves per SIMD, the
// complex ones (about 124) four.
template <class T, bool HERM>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(1, Scalar<T>::complex ? 4 : 8))) void k_diag_ldl(
    const DiagDesc<T>* __restrict__ descs, int* status) {
    using W = typename Scalar<T>::wide;
    __shared__ W col[2][NB];
    __shared__ W dinv[NB];
    __shared__ W dval[NB];

    // the diagonal blocks are the factorization's critical path while the
    // lookahead's updates run beside them on the same CUs
    __builtin_amdgcn_s_setprio(3);
    const DiagDesc<T> d = descs[blockIdx.x];
    const int b = d.b, tid = threadIdx.x;
    const int ti = tid & 15, tj = tid >> 4;
    W M[4][4];
    // all sixteen loads in flight at once (an element outside the lower
    // triangle reads A11(0,0) and is replaced by zero)
    T raw[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ti + 16 * r, c = tj + 16 * q;
            const bool in = i < b && c < b && i >= c;
            raw[r][q] = gload(d.A11 + (in ? i + (int64_t)c * d.lda : 0));
        }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ti + 16 * r, c = tj + 16 * q;
            M[r][q] = (i < b && c < b && i >= c) ? cvt<W>(raw[r][q]) : W{};
        }
    const bool lowdiag = ti >= tj;
    W* Dout = reinterpret_cast<W*>(d.D);
    for (int j = 0; j < b; ++j) {
        const int buf = j & 1, J = j >> 4;
        if (tj == (j & 15)) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q == J) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) col[buf][ti + 16 * r] = M[r][q];
                }
        }
        __syncthreads();
        W piv = col[buf][j];
        if (HERM) piv = real_part_only(piv);
        const W ip = pivot_inverse(piv);
        // pivots stay in LDS during the loop: a global store here would make
        // every barrier wait for it (the barrier's fence covers global memory)
        if (tid == 0) {
            dval[j] = piv;
            dinv[j] = ip;
        }
        // row multipliers: x_i = a(i,j)/d_j (lower), u_i (upper, zero for i > j)
        W x[4], u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = ti + 16 * r;
            const W a = col[buf][i];
            x[r] = a * ip;
            u[r] = (i < j) ? a : ((i == j) ? one<W>() : W{});
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (16 * q + 15 <= j) continue;  // every column of this slot is finished
            const int c = tj + 16 * q;
            const bool cact = c > j;
            const W cv = col[buf][c];
            const W g = cact ? (HERM ? conj_(cv) : cv) : W{};
            const W xx = cact ? cv * ip : W{};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (r > q) {
                    M[r][q] -= x[r] * g;
                } else if (r < q) {
                    if (16 * r <= j) M[r][q] -= u[r] * xx;
                } else {
                    M[r][q] -= pick(lowdiag, x[r], u[r]) * pick(lowdiag, g, xx);
                }
            }
        }
    }
    __syncthreads();
    for (int j = tid; j < b; j += blockDim.x) {
        Dout[j] = dval[j];
        if (is_zero_or_nonfinite(dval[j])) record_status(status, 2, d.front, d.col0 + j);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ti + 16 * r, c = tj + 16 * q;
            // L11 (unit diagonal) back into the front
            if (i < b && c < b && i >= c)
                d.A11[i + (int64_t)c * d.lda] = (i == c) ? one<T>() : cvt<T>(M[r][q] * dinv[c]);
            // inverse of the unit lower factor, li(c,i) = M[i][c] for i < c.
            // Each Linv entry has one writer: the owner of M[i][i] writes the
            // diagonal, the owner of M[i][c] (i < c) writes (c,i) and (i,c).
            if (i == c) {
                d.Linv[i + c * NB] = (i < b) ? one<T>() : zero<T>();
            } else if (i < c) {
                d.Linv[c + i * NB] = (c < b) ? cvt<T>(M[r][q]) : zero<T>();
                d.Linv[i + c * NB] = zero<T>();
            }
        }
}

// ---------------------------------------------------------------------------
// Diagonal block, sub-panel form: the same outputs as k_diag_ldl (unpivoted
// LDL^T of a b x b block in FP64 field type, L back into the front, D, the
// inverse of the unit lower factor) computed with 16-column sub-panels. The
// block lives in LDS (column-major, ld NB + 1, field type). Per sub-panel p:
//   1. one wave factors the (64 - 16p) x 16 panel with one row per lane: the
//      pivot and the panel's column j reach every lane by cross-lane reads
//      (v_readlane), so the 16 pivot steps need no barrier and no LDS round
//      trip; L (scaled) goes back into the block, W = L D of the rows below
//      the sub-panel into a side buffer;
//   2. all four waves apply the rank-16 update A22 -= L21 op(W21)^T to the
//      trailing lower triangle in 2 x 2 register tiles.
// The inverse X = L^{-1} is built by blocks of 16: each diagonal block by
// the panel wave along with its panel (the same row eliminations applied to
// the identity), the blocks below by distance from the diagonal,
// X_qp = -X_qq (sum_{p<=r<q} L_qr X_rp). X's
// strict lower triangle is kept transposed in the block's free upper
// triangle. Rows and columns >= b are padded with the identity.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rdlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ cdouble rdlane(cdouble v, int l) {
    return cdouble{rdlane(v.x, l), rdlane(v.y, l)};
}

// Four workgroups per CU (40 KB of LDS for real blocks; the complex
// instantiations, 78 KB, are not launched).
template <class T, bool HERM>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(Scalar<T>::complex ? 2 : 4, 8))) void k_diag_ldl_sp(const DiagDesc<T>* __restrict__ descs,
                                                     int* status) {
    using W = typename Scalar<T>::wide;
    constexpr int LDA = NB + 1;
    constexpr int PW = 16;            // sub-panel width
    constexpr int WR = NB - PW;       // rows of W kept per sub-panel column (rows >= 16)
    __shared__ W A[NB * LDA];
    __shared__ W Wp[PW * WR];         // W = L D below the sub-panel; products of the inverse
    __shared__ W dval[NB];

    __builtin_amdgcn_s_setprio(3);
    const DiagDesc<T> d = descs[blockIdx.x];
    const int b = d.b, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    {
        // lower triangle of the block, identity outside b x b; sixteen loads
        // in flight per thread (rows along the lanes: coalesced columns)
        T raw[NB * NB / 256];
#pragma unroll
        for (int t = 0; t < NB * NB / 256; ++t) {
            const int e = tid + 256 * t, i = e & (NB - 1), c = e / NB;
            const bool in = i < b && c < b && i >= c;
            raw[t] = gload(d.A11 + (in ? i + (int64_t)c * d.lda : 0));
        }
#pragma unroll
        for (int t = 0; t < NB * NB / 256; ++t) {
            const int e = tid + 256 * t, i = e & (NB - 1), c = e / NB;
            W v{};
            if (i >= c) {
                if (i < b && c < b) v = cvt<W>(raw[t]);
                else if (i == c) v = one<W>();
            }
            A[c * LDA + i] = v;
        }
    }
    __syncthreads();
    const int np = (b + PW - 1) / PW;
    for (int p = 0; p < np; ++p) {
        const int c0 = PW * p;
        if (wv == 0) {
            const int i = c0 + lane;             // this lane's row
            const bool act = i < NB;
            // r: this row's panel entries; xr: for the sub-block's rows
            // (lanes < 16), the row of X_pp = L_pp^{-1}, built by the same
            // eliminations (row_i -= L(i, j) row_j for i > j)
            W r[PW], xr[PW];
#pragma unroll
            for (int q = 0; q < PW; ++q) {
                r[q] = act ? A[(c0 + q) * LDA + (act ? i : 0)] : W{};
                xr[q] = (lane == q) ? one<W>() : W{};
            }
#pragma unroll
            for (int j = 0; j < PW; ++j) {
                W piv = rdlane(r[j], j);
                if (HERM) piv = real_part_only(piv);
                const W ip = pivot_inverse(piv);
                const W x = r[j] * ip;           // L(i, c0 + j) for i > c0 + j
#pragma unroll
                for (int c = j + 1; c < PW; ++c) {
                    const W a = rdlane(r[j], c);  // unscaled A(c0 + c, c0 + j)
                    r[c] -= x * (HERM ? conj_(a) : a);
                }
                const W xl = (lane > j) ? x : W{};
#pragma unroll
                for (int c = 0; c <= j; ++c) xr[c] -= xl * rdlane(xr[c], j);
                if (act && i > c0 + j) {
                    A[(c0 + j) * LDA + i] = x;
                    if (i >= PW) Wp[j * WR + (i - PW)] = r[j];
                }
                if (lane == 0) dval[c0 + j] = piv;
            }
            // X_pp's strict lower triangle, transposed into the block's upper one
#pragma unroll
            for (int c = 0; c < PW; ++c)
                if (c < lane && lane < PW) A[(c0 + lane) * LDA + c0 + c] = xr[c];
        }
        __syncthreads();
        // trailing update of rows and columns s0..63 (lower triangle) in 2 x 2 tiles
        const int s0 = c0 + PW, h = (NB - s0) / 2;
        for (int t = tid; t < h * (h + 1) / 2; t += 256) {
            int ti = (int)((__fsqrt_rn(8.f * t + 1.f) - 1.f) * 0.5f);
            while (ti * (ti + 1) / 2 > t) --ti;
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            const int tc = t - ti * (ti + 1) / 2;
            const int i0 = s0 + 2 * ti, k0 = s0 + 2 * tc;
            W a00 = A[k0 * LDA + i0], a10 = A[k0 * LDA + i0 + 1];
            W a01 = A[(k0 + 1) * LDA + i0], a11 = A[(k0 + 1) * LDA + i0 + 1];
#pragma unroll 4
            for (int j = 0; j < PW; ++j) {
                const W l0 = A[(c0 + j) * LDA + i0], l1 = A[(c0 + j) * LDA + i0 + 1];
                W w0 = Wp[j * WR + (k0 - PW)], w1 = Wp[j * WR + (k0 + 1 - PW)];
                if (HERM) {
                    w0 = conj_(w0);
                    w1 = conj_(w1);
                }
                a00 -= l0 * w0;
                a10 -= l1 * w0;
                a01 -= l0 * w1;
                a11 -= l1 * w1;
            }
            A[k0 * LDA + i0] = a00;
            A[k0 * LDA + i0 + 1] = a10;
            if (ti != tc) A[(k0 + 1) * LDA + i0] = a01;
            A[(k0 + 1) * LDA + i0 + 1] = a11;
        }
        __syncthreads();
    }
    // (the diagonal blocks of X = L^{-1} were built by the panel wave)
    // blocks below the diagonal by distance: T = sum_r L_qr X_rp, X_qp = -X_qq T
    for (int dist = 1; dist < np; ++dist) {
        const int ei = tid & (PW - 1), ec = tid >> 4;
        for (int p = 0; p + dist < np; ++p) {
            const int q = p + dist;
            W s{};
            for (int rr = p; rr < q; ++rr) {
#pragma unroll
                for (int k = 0; k < PW; ++k) {
                    const int row = PW * rr + k, col = PW * p + ec;
                    // X(row, col): unit diagonal, zero above, transposed storage below
                    W xv = (row > col) ? A[row * LDA + col] : ((row == col) ? one<W>() : W{});
                    s += A[row * LDA + PW * q + ei] * xv;   // L(16q + ei, row)
                }
            }
            Wp[p * PW * PW + ec * PW + ei] = s;
        }
        __syncthreads();
        for (int p = 0; p + dist < np; ++p) {
            const int q = p + dist;
            W s = Wp[p * PW * PW + ec * PW + ei];
#pragma unroll
            for (int k = 0; k < PW; ++k)
                if (k < ei) s += A[(PW * q + ei) * LDA + PW * q + k] * Wp[p * PW * PW + ec * PW + k];
            A[(PW * q + ei) * LDA + PW * p + ec] = -s;
        }
        __syncthreads();
    }
    W* Dout = reinterpret_cast<W*>(d.D);
    for (int j = tid; j < b; j += 256) {
        Dout[j] = dval[j];
        if (is_zero_or_nonfinite(dval[j])) record_status(status, 2, d.front, d.col0 + j);
    }
#pragma unroll 1
    for (int t = 0; t < NB * NB / 256; ++t) {
        const int e = tid + 256 * t, i = e & (NB - 1), c = e / NB;
        if (i < b && c < b && i >= c)
            d.A11[i + (int64_t)c * d.lda] = (i == c) ? one<T>() : cvt<T>(A[c * LDA + i]);
        T li = zero<T>();
        if (i == c) li = (i < b) ? one<T>() : zero<T>();
        else if (i > c && i < b) li = cvt<T>(A[i * LDA + c]);
        d.Linv[i + c * NB] = li;
    }
}

template <class T>
void launch_diag_ldl(const DiagDesc<T>* d_descs, int ndesc, bool herm, int* d_status,
                     hipStream_t s) {
    if (ndesc <= 0) return;
    // Real blocks take the sub-panel kernel (one 64-block: 38 us against 47;
    // CLIQG_DIAG_KERNEL=0 selects the register-cyclic one). Complex blocks
    // keep the register-cyclic kernel: the sub-panel one needs 78 KB of LDS
    // and then waits, on CUs full of the lookahead's GEMM tiles, for a CU to
    // free that much (120^3 PML factor: 126 ms of diagonal blocks against 48).
    if constexpr (!Scalar<T>::complex) {
        static const bool sp = [] {
            const char* e = std::getenv("CLIQG_DIAG_KERNEL");
            return !(e && std::atoi(e) == 0);
        }();
        if (sp) {
            if (herm)
                hipLaunchKernelGGL((k_diag_ldl_sp<T, true>), dim3(ndesc), dim3(256), 0, s,
                                   d_descs, d_status);
            else
                hipLaunchKernelGGL((k_diag_ldl_sp<T, false>), dim3(ndesc), dim3(256), 0, s,
                                   d_descs, d_status);
            return;
        }
    }
    if (herm)
        hipLaunchKernelGGL((k_diag_ldl<T, true>), dim3(ndesc), dim3(256), 0, s, d_descs, d_status);
    else
        hipLaunchKernelGGL((k_diag_ldl<T, false>), dim3(ndesc), dim3(256), 0, s, d_descs, d_status);
}

// ---------------------------------------------------------------------------
// Solve helpers. Column j of an N x k block is blockIdx.y (k <= 65535 is
// checked by the host), rows are strided over blockIdx.x: no 64-bit index
// division per element.
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}
static inline dim3 grid_cols(int N, int k) {
    unsigned gx = grid_for(N);
    if (gx > 512) gx = 512;
    return dim3(gx, (unsigned)k);
}

// xn(map[i], j) = x(i, j): reads of x (the wide type) coalesced, map[i] the
// nodal index of original row i
template <class T, class V>
__global__ void k_pull(const V* __restrict__ x, int64_t ldx, const int* __restrict__ map,
                       T* __restrict__ xn, int N) {
    const int j = blockIdx.y;
    const V* xc = x + (int64_t)j * ldx;
    T* yc = xn + (int64_t)j * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        yc[map[i]] = cvt<T>(xc[i]);
}
// x(i, j) = xn(map[i], j): writes of x coalesced
template <class T, class V>
__global__ void k_push(const T* __restrict__ xn, const int* __restrict__ map, V* __restrict__ x,
                       int64_t ldx, int N) {
    const int j = blockIdx.y;
    const T* xc = xn + (int64_t)j * N;
    V* yc = x + (int64_t)j * ldx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        yc[i] = cvt<V>(xc[map[i]]);
}

template <class T, class V>
void launch_pull(const V* x, int64_t ldx, const int* map, T* xn, int N, int k, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL((k_pull<T, V>), grid_cols(N, k), dim3(256), 0, s, x, ldx, map, xn, N);
}
template <class T, class V>
void launch_push(const T* xn, const int* map, V* x, int64_t ldx, int N, int k, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL((k_push<T, V>), grid_cols(N, k), dim3(256), 0, s, xn, map, x, ldx, N);
}

// Right-hand-side columns a gather / solve extend-add thread handles: all of
// a thread's loads are issued before its stores (the destination may alias
// the source pool, so the compiler could not reorder them itself).
constexpr int RHS_CHUNK = 8;

template <class T>
__global__ void k_gather(const SolveSeg* __restrict__ segs, const int* __restrict__ row_start,
                         int nseg, int nrows, T* __restrict__ wpool, const T* __restrict__ xn,
                         int N, int k, const int* __restrict__ lowpool) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows) return;
    const int j0 = blockIdx.y * RHS_CHUNK;
    const int si = find_segment(row_start, nseg, g);
    const SolveSeg sg = segs[si];
    const int i = g - row_start[si];
    T* w = wpool + sg.w * k + i + (int64_t)j0 * sg.u;
    const T* src = xn + lowpool[sg.low + i] + (int64_t)j0 * N;
    T v[RHS_CHUNK];
#pragma unroll
    for (int q = 0; q < RHS_CHUNK; ++q)
        if (j0 + q < k) v[q] = src[(int64_t)q * N];
#pragma unroll
    for (int q = 0; q < RHS_CHUNK; ++q)
        if (j0 + q < k) w[(int64_t)q * sg.u] = v[q];
}
template <class T>
void launch_gather(const SolveSeg* segs, const int* row_start, int nseg, int nrows, T* wpool,
                   const T* xn, int N, int k, const int* lowpool, hipStream_t s) {
    if (nrows <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_gather<T>, dim3((nrows + 255) / 256, (k + RHS_CHUNK - 1) / RHS_CHUNK),
                       dim3(256), 0, s, segs, row_start, nseg, nrows, wpool, xn, N, k, lowpool);
}

template <class T>
__global__ void k_solve_extend(const SolveExtend* __restrict__ ds, const int* __restrict__ row_start,
                               int nd, int nrows, T* __restrict__ wpool, T* __restrict__ xn, int N,
                               const int* __restrict__ relpool, int k) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows) return;
    const int j0 = blockIdx.y * RHS_CHUNK;
    const int c = find_segment(row_start, nd, g);
    const SolveExtend d = ds[c];
    const int i = g - row_start[c] + d.i0;
    const T* wc = wpool + d.wc * k + i + (int64_t)j0 * d.uc;
    const int I = relpool[d.rel + i];
    T* dst;
    int64_t ld;
    if (I < d.np) {
        dst = xn + d.offp + I;
        ld = N;
    } else {
        dst = wpool + d.wp * k + (I - d.np);
        ld = d.up;
    }
    dst += (int64_t)j0 * ld;
    T a[RHS_CHUNK], b[RHS_CHUNK];
#pragma unroll
    for (int q = 0; q < RHS_CHUNK; ++q)
        if (j0 + q < k) {
            a[q] = wc[(int64_t)q * d.uc];
            b[q] = dst[(int64_t)q * ld];
        }
#pragma unroll
    for (int q = 0; q < RHS_CHUNK; ++q)
        if (j0 + q < k) dst[(int64_t)q * ld] = b[q] + a[q];
}
template <class T>
void launch_solve_extend(const SolveExtend* d, const int* row_start, int nd, int nrows,
                         T* wpool, T* xn, int N, const int* relpool, int k, hipStream_t s) {
    if (nrows <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_solve_extend<T>,
                       dim3((nrows + 255) / 256, (k + RHS_CHUNK - 1) / RHS_CHUNK), dim3(256), 0, s,
                       d, row_start, nd, nrows, wpool, xn, N, relpool, k);
}

// One thread per row: the pivot's reciprocal once, then its k columns
template <class T, class V>
__global__ void k_diag_solve(T* __restrict__ xn, const V* __restrict__ D, int N, int k,
                             int* status) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const V dv = D[p];
        if (is_zero_or_nonfinite(dv)) record_status(status, 2, -1, p);
        const V r = one<V>() / dv;
        T* xc = xn + p;
        for (int j = 0; j < k; ++j) xc[(int64_t)j * N] = cvt<T>(cvt<V>(xc[(int64_t)j * N]) * r);
    }
}
template <class T, class V>
void launch_diag_solve(T* xn, const V* D, int N, int k, int* d_status, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL((k_diag_solve<T, V>), dim3(grid_for(N)), dim3(256), 0, s, xn, D, N, k,
                       d_status);
}

// ---------------------------------------------------------------------------
// FP64 sparse kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// y := alpha A x + beta b over CSR rows (b = y when null); with colmax,
// colmax[j] = max(colmax[j], max_i |y(i, j)|). One thread per (row, 8
// right-hand sides): the row's indices and values are read once per chunk
// of columns, and for rows of at most 8 entries all of the thread's loads
// are issued before its products.
constexpr int SPMV_COLS = 8, SPMV_ROWNZ = 8;
// With xn, y is not written: row i of the result goes to xn(map[i], j) in the
// working type (the nodal right-hand side of a refinement solve), ld N.
template <class V, class T>
__global__ __launch_bounds__(256) void k_spmv(int N, const int64_t* __restrict__ rowoff,
                                              const int* __restrict__ cols,
                                              const V* __restrict__ vals, double alpha,
                                              const V* __restrict__ x, int64_t ldx, double beta,
                                              const V* __restrict__ b, int64_t ldb,
                                              V* __restrict__ y, int64_t ldy, int k,
                                              unsigned long long* colmax, T* __restrict__ xn,
                                              const int* __restrict__ map) {
    __shared__ double red[SPMV_COLS][4];
    // one-dimensional grid of (row block, column chunk) pairs, remapped so
    // each XCD runs a contiguous range of row blocks of one column chunk:
    // the stencil's neighbours in x then come from that XCD's L2
    const int rb = (N + 255) / 256;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int i = (t % rb) * 256 + threadIdx.x;
    const int j0 = (t / rb) * SPMV_COLS;
    const int nj = min(SPMV_COLS, k - j0);
    double m[SPMV_COLS];
#pragma unroll
    for (int q = 0; q < SPMV_COLS; ++q) m[q] = 0.0;
    if (i < N) {
        V acc[SPMV_COLS];
#pragma unroll
        for (int q = 0; q < SPMV_COLS; ++q) acc[q] = V{};
        const int64_t e0 = rowoff[i], e1 = rowoff[i + 1];
        const int ne = (int)(e1 - e0);
        if (ne <= SPMV_ROWNZ) {
            int c[SPMV_ROWNZ];
            V a[SPMV_ROWNZ];
#pragma unroll
            for (int u = 0; u < SPMV_ROWNZ; ++u)
                if (u < ne) {
                    c[u] = cols[e0 + u];
                    a[u] = vals[e0 + u];
                }
#pragma unroll
            for (int q = 0; q < SPMV_COLS; ++q)
                if (q < nj) {
                    const V* xc = x + (int64_t)(j0 + q) * ldx;
                    V sum = V{};
#pragma unroll
                    for (int u = 0; u < SPMV_ROWNZ; ++u)
                        if (u < ne) sum = sum + a[u] * xc[c[u]];
                    acc[q] = sum;
                }
        } else {
            for (int64_t e = e0; e < e1; ++e) {
                const V a = vals[e];
                const V* xc = x + cols[e] + (int64_t)j0 * ldx;
#pragma unroll
                for (int q = 0; q < SPMV_COLS; ++q)
                    if (q < nj) acc[q] = acc[q] + a * xc[q * ldx];
            }
        }
#pragma unroll
        for (int q = 0; q < SPMV_COLS; ++q) {
            if (q >= nj) break;
            V* py = y + i + (int64_t)(j0 + q) * ldy;
            V out = scale(acc[q], alpha);
            if (beta != 0.0) out = out + scale(b ? b[i + (int64_t)(j0 + q) * ldb] : *py, beta);
            if (xn) xn[map[i] + (int64_t)(j0 + q) * N] = cvt<T>(out);
            else *py = out;
            m[q] = abs1(out);
        }
    }
    if (colmax) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < SPMV_COLS; ++q) {
            double v = m[q];
            for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
            if (lane == 0) red[q][wave] = v;
        }
        __syncthreads();
        if (threadIdx.x < nj) {
            const int q = threadIdx.x;
            const double v = fmax(fmax(red[q][0], red[q][1]), fmax(red[q][2], red[q][3]));
            atomic_max_nonneg(colmax + j0 + q, v);
        }
    }
}
template <class V, class T>
static void launch_spmv_any(int N, const int64_t* rowoff, const int* cols, const V* vals,
                            double alpha, const V* x, int64_t ldx, double beta, const V* b,
                            int64_t ldb, V* y, int64_t ldy, int k, unsigned long long* colmax,
                            T* xn, const int* map, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    const int64_t blocks = (int64_t)((N + 255) / 256) * ((k + SPMV_COLS - 1) / SPMV_COLS);
    if (blocks > 0x7fffffff) throw Error(1, "SpMV grid too large");
    hipLaunchKernelGGL((k_spmv<V, T>), dim3((unsigned)blocks), dim3(256), 0, s, N, rowoff, cols,
                       vals, alpha, x, ldx, beta, b, ldb, y, ldy, k, colmax, xn, map);
}
template <class V>
void launch_spmv_ex(int N, const int64_t* rowoff, const int* cols, const V* vals, double alpha,
                    const V* x, int64_t ldx, double beta, const V* b, int64_t ldb, V* y,
                    int64_t ldy, int k, unsigned long long* colmax, hipStream_t s) {
    launch_spmv_any<V, V>(N, rowoff, cols, vals, alpha, x, ldx, beta, b, ldb, y, ldy, k, colmax,
                          nullptr, nullptr, s);
}
template <class T, class V>
void launch_residual_nodal(int N, const int64_t* rowoff, const int* cols, const V* vals,
                           const V* x, int64_t ldx, const V* b, int64_t ldb, int k,
                           unsigned long long* colmax, T* xn, const int* map, hipStream_t s) {
    launch_spmv_any<V, T>(N, rowoff, cols, vals, -1.0, x, ldx, 1.0, b, ldb, nullptr, 0, k, colmax,
                          xn, map, s);
}

// x(i, j) += xn(map[i], j) with colmax[j] = max(colmax[j], max_i |x(i, j)|)
template <class T, class V>
__global__ __launch_bounds__(256) void k_push_add_abs(const T* __restrict__ xn,
                                                      const int* __restrict__ map,
                                                      V* __restrict__ x, int64_t ldx, int N,
                                                      unsigned long long* colmax) {
    __shared__ double red[4];
    const T* xc = xn + (int64_t)blockIdx.y * N;
    V* yc = x + (int64_t)blockIdx.y * ldx;
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const V v = yc[i] + cvt<V>(xc[map[i]]);
        yc[i] = v;
        m = fmax(m, abs1(v));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        atomic_max_nonneg(colmax + blockIdx.y, fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}
template <class T, class V>
void launch_push_add_abs(const T* xn, const int* map, V* x, int64_t ldx, int N, int k,
                         unsigned long long* colmax, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL((k_push_add_abs<T, V>), grid_cols(N, k), dim3(256), 0, s, xn, map, x, ldx,
                       N, colmax);
}
template <class V>
void launch_spmv(int N, const int64_t* rowoff, const int* cols, const V* vals, double alpha,
                 const V* x, int64_t ldx, double beta, V* y, int64_t ldy, int k, hipStream_t s) {
    launch_spmv_ex<V>(N, rowoff, cols, vals, alpha, x, ldx, beta, nullptr, 0, y, ldy, k, nullptr,
                      s);
}

template <class V>
__global__ void k_row_abs_sum_max(int N, const int64_t* __restrict__ rowoff,
                                  const V* __restrict__ vals, unsigned long long* out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int64_t e = rowoff[i]; e < rowoff[i + 1]; ++e) s += abs1(vals[e]);
        m = fmax(m, s);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(out, red[0]);
}
template <class V>
void launch_row_abs_sum_max(int N, const int64_t* rowoff, const V* vals, unsigned long long* out,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_row_abs_sum_max<V>, dim3(grid_for(N) > 1024 ? 1024 : grid_for(N)),
                       dim3(256), 0, s, N, rowoff, vals, out);
}

template <class V>
__global__ void k_col_abs_max(const V* __restrict__ x, int64_t ldx, int N,
                              unsigned long long* out) {
    __shared__ double red[256];
    const int j = blockIdx.y;
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        m = fmax(m, abs1(x[i + j * ldx]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(out + j, red[0]);
}
template <class V>
void launch_col_abs_max(const V* x, int64_t ldx, int N, int k, unsigned long long* out,
                        hipStream_t s) {
    unsigned gx = grid_for(N);
    if (gx > 256) gx = 256;
    hipLaunchKernelGGL(k_col_abs_max<V>, dim3(gx, k), dim3(256), 0, s, x, ldx, N, out);
}

template <class V>
__global__ void k_copy_cols(V* __restrict__ y, int64_t ldy, const V* __restrict__ x, int64_t ldx,
                            int N) {
    V* yc = y + (int64_t)blockIdx.y * ldy;
    const V* xc = x + (int64_t)blockIdx.y * ldx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        yc[i] = xc[i];
}
template <class V>
void launch_copy_cols(V* y, int64_t ldy, const V* x, int64_t ldx, int N, int k, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_copy_cols<V>, grid_cols(N, k), dim3(256), 0, s, y, ldy, x, ldx, N);
}

template <class V>
__global__ void k_adjoint_dA(int N, const int64_t* __restrict__ rowoff, const int* __restrict__ cols,
                             const V* __restrict__ lam, int64_t ldl, const V* __restrict__ x,
                             int64_t ldx, int k, bool conj, V* __restrict__ dA) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        for (int64_t e = rowoff[i]; e < rowoff[i + 1]; ++e) {
            const int c = cols[e];
            V acc = V{};
            for (int r = 0; r < k; ++r) acc = acc + lam[i + r * ldl] * cj(x[c + r * ldx], conj);
            dA[e] = -acc;
        }
    }
}
template <class V>
void launch_adjoint_dA(int N, const int64_t* rowoff, const int* cols, const V* lambda,
                       int64_t ldl, const V* x, int64_t ldx, int k, bool conj, V* dA,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_dA<V>, dim3(grid_for(N)), dim3(256), 0, s, N, rowoff, cols, lambda,
                       ldl, x, ldx, k, conj, dA);
}

template <class T, class V>
__global__ void k_widen(const T* __restrict__ src, V* __restrict__ dst, int64_t n) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x)
        dst[g] = cvt<V>(src[g]);
}
template <class T, class V>
void launch_widen(const T* src, V* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_widen<T, V>), dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
}

template <class V>
__global__ void k_permute_rows(const V* __restrict__ src, int64_t lds, V* __restrict__ dst,
                               int64_t ldd, const int* __restrict__ perm, int N, bool scatter) {
    const V* sc = src + (int64_t)blockIdx.y * lds;
    V* dc = dst + (int64_t)blockIdx.y * ldd;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        if (scatter) dc[perm[p]] = sc[p];
        else dc[p] = sc[perm[p]];
    }
}
template <class V>
void launch_permute_rows(const V* src, int64_t lds, V* dst, int64_t ldd, const int* perm, int N,
                         int k, bool scatter, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_permute_rows<V>, grid_cols(N, k), dim3(256), 0, s, src, lds, dst, ldd,
                       perm, N, scatter);
}

template <class V>
__global__ __launch_bounds__(256) void k_col_sumsq(const V* __restrict__ x, int64_t ldx, int N,
                                                   double* out) {
    __shared__ double red[256];
    const V* xc = x + (int64_t)blockIdx.x * ldx;
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += 256) acc += sq_mod(xc[i]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
template <class V>
void launch_col_sumsq(const V* x, int64_t ldx, int N, int k, double* out, hipStream_t s) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_col_sumsq<V>, dim3(k), dim3(256), 0, s, x, ldx, N, out);
}

__global__ void k_fill_int(int* __restrict__ p, int n, int value) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = value;
}
void launch_fill_int(int* p, int n, int value, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n)), dim3(256), 0, s, p, n, value);
}

// ---------------------------------------------------------------------------
// explicit instantiations
// ---------------------------------------------------------------------------
#define CQ_INST_T(T)                                                                           \
    template void launch_zero_segments<T>(const ZeroSeg<T>*, const int*, int, int, hipStream_t); \
    template void launch_copy2d<T>(const CopyDesc<T>*, const int*, int, int, hipStream_t);       \
    template void launch_reduce_slices<T>(const ReduceDesc<T>*, const int*, int, int, hipStream_t); \
    template void launch_zero_lower<T>(const ZeroLowerDesc<T>*, const int*, int, int, hipStream_t); \
    template void launch_gemm<T>(const GemmDesc<T>*, const int*, int, int, int, int, hipStream_t);        \
    template void launch_extend_add<T>(const ExtendDesc<T>*, const int*, int, int, hipStream_t); \
    template void launch_diag_ldl<T>(const DiagDesc<T>*, int, bool, int*, hipStream_t);         \
    template void launch_wide_inverse_small<T>(const WideInvDesc<T>*, int, hipStream_t);        \
    template void launch_gather<T>(const SolveSeg*, const int*, int, int, T*, const T*, int,    \
                                   int, const int*, hipStream_t);                               \
    template void launch_solve_extend<T>(const SolveExtend*, const int*, int, int, T*, T*, int, \
                                         const int*, int, hipStream_t);                         \
    template void launch_panel_step<T>(const PanelStepDesc<T>*, const int*, int, int, bool, bool, \
                                       hipStream_t);

CQ_INST_T(float)
CQ_INST_T(double)
CQ_INST_T(cfloat)
CQ_INST_T(cdouble)

#define CQ_INST_TV(T, V)                                                                        \
    template void launch_assemble<T, V>(const V*, const int64_t*, int64_t, T*, bool, hipStream_t); \
    template void launch_pull<T, V>(const V*, int64_t, const int*, T*, int, int, hipStream_t);   \
    template void launch_push<T, V>(const T*, const int*, V*, int64_t, int, int, hipStream_t);   \
    template void launch_diag_solve<T, V>(T*, const V*, int, int, int*, hipStream_t);           \
    template void launch_widen<T, V>(const T*, V*, int64_t, hipStream_t);                        \
    template void launch_residual_nodal<T, V>(int, const int64_t*, const int*, const V*,         \
                                              const V*, int64_t, const V*, int64_t, int,         \
                                              unsigned long long*, T*, const int*, hipStream_t); \
    template void launch_push_add_abs<T, V>(const T*, const int*, V*, int64_t, int, int,         \
                                            unsigned long long*, hipStream_t);

CQ_INST_TV(float, double)
CQ_INST_TV(double, double)
CQ_INST_TV(cfloat, cdouble)
CQ_INST_TV(cdouble, cdouble)

#define CQ_INST_V(V)                                                                            \
    template void launch_spmv<V>(int, const int64_t*, const int*, const V*, double, const V*,   \
                                 int64_t, double, V*, int64_t, int, hipStream_t);              \
    template void launch_spmv_ex<V>(int, const int64_t*, const int*, const V*, double,          \
                                    const V*, int64_t, double, const V*, int64_t, V*, int64_t,  \
                                    int, unsigned long long*, hipStream_t);                    \
    template void launch_row_abs_sum_max<V>(int, const int64_t*, const V*, unsigned long long*, \
                                            hipStream_t);                                      \
    template void launch_col_abs_max<V>(const V*, int64_t, int, int, unsigned long long*,       \
                                        hipStream_t);                                          \
    template void launch_copy_cols<V>(V*, int64_t, const V*, int64_t, int, int, hipStream_t);   \
    template void launch_permute_rows<V>(const V*, int64_t, V*, int64_t, const int*, int, int,  \
                                         bool, hipStream_t);                                    \
    template void launch_col_sumsq<V>(const V*, int64_t, int, int, double*, hipStream_t);      \
    template void launch_adjoint_dA<V>(int, const int64_t*, const int*, const V*, int64_t,      \
                                       const V*, int64_t, int, bool, V*, hipStream_t);

CQ_INST_V(double)
CQ_INST_V(cdouble)

}  // namespace cliqg
