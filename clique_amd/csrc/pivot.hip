// Device kernels of intra-front Bunch-Kaufman pivoting (gfx950): the
// pivoted LDL^T of each front's ATL, the quasi-diagonal scaling of ABL, and
// the row permutations and quasi-diagonal solve of the pivoted solves.
//
// Reference: FrontLDLIntraPiv (include/clique/numeric/ldl/local_front.hpp:
// 73-94), which calls El::LDL(ATL, dSub, p, conjugate, BUNCH_KAUFMAN_A),
// El::PermuteCols(ABL, p), a unit TRSM, El::QuasiDiagonalSolve and Trrk;
// QuasiDiagonalSolve in the diagonal solve (diagonal_solve.hpp:58-60);
// PermuteRows / InversePermuteRows around the front solves
// (lower_solve/local_front.hpp:53-62, 84-93).
//
// The pivoted LDL^T of ATL is right-looking and blocked by panels of nbp
// columns with delayed updates inside the panel: the panel's columns are
// brought up to date one at a time from the panel's own L and W = L D (a
// matrix-vector product over the panel), the Bunch-Kaufman "A" rule picks a
// 1x1 or 2x2 pivot from the updated column k and, when needed, the updated
// column r of the largest off-diagonal entry, and the symmetric interchange
// is applied to the trailing matrix, to the rows of every finished L column
// (so L is in standard form: (P A P^T)(i, j) = A(piv[i], piv[j])) and to the
// panel's W rows. After each panel the trailing lower triangle of ATL gets
// one grouped MFMA update with K = nbp + 1 (the launch after this kernel).
#include "kernels.hpp"

namespace cliqg {

namespace {

constexpr int BK_THREADS = 512;
constexpr int BK_WAVES = BK_THREADS / 64;

__device__ __forceinline__ bool nonzero_v(double a) { return a != 0.0; }
__device__ __forceinline__ bool nonzero_v(cdouble a) { return a.x != 0.0 || a.y != 0.0; }

__device__ __forceinline__ double recip(double a) { return 1.0 / a; }
__device__ __forceinline__ cdouble recip(cdouble a) { return cdouble{1.0, 0.0} / a; }

__device__ __forceinline__ double real_only(double a) { return a; }
__device__ __forceinline__ cdouble real_only(cdouble a) { return {a.x, 0.0}; }

__device__ __forceinline__ void status_once(int* status, int code, int front, int col) {
    if (atomicCAS(status, 0, code) == 0) {
        status[1] = front;
        status[2] = col;
    }
}

__device__ __forceinline__ int segment_of(const int* __restrict__ start, int n, int g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Block-wide maximum of (v, idx): the larger value wins, equal values go to
// the smaller index (the first maximum of a sequential scan). Every thread
// returns the result. Called by all threads of the block.
__device__ __forceinline__ void block_argmax(double& v, int& idx, double* sv, int* si) {
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(v, off, 64);
        const int oi = __shfl_down(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sv[w] = v;
        si[w] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bv = sv[0];
        int bi = si[0];
        for (int q = 1; q < BK_WAVES; ++q)
            if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) {
                bv = sv[q];
                bi = si[q];
            }
        sv[BK_WAVES] = bv;
        si[BK_WAVES] = bi;
    }
    __syncthreads();
    v = sv[BK_WAVES];
    idx = si[BK_WAVES];
    __syncthreads();
}

// One panel of the pivoted LDL^T of ATL for every front of a launch (one
// workgroup per front). On entry columns [0, k0) are finished and the
// trailing matrix is up to date except for column k0 when the previous panel
// ended on a 1x1 pivot (its delayed update is applied first) or, when it
// ended on a 2x2 pivot that took column k0 along, W(:, k0) is cleared so the
// next trailing update does not apply that column twice. On exit column e =
// k0 + nbp of W is cleared unless a 2x2 pivot took it (the trailing update
// runs over columns k0..e).
template <class T, bool HERM>
__global__ __launch_bounds__(BK_THREADS) void k_bk_panel(const BkDesc<T>* __restrict__ descs,
                                                         int nbp, int* status) {
    using V = typename Scalar<T>::wide;
    __shared__ V wrow[BK_MAX_PANEL + 1];
    __shared__ double sv[BK_WAVES + 1];
    __shared__ int si[BK_WAVES + 1];
    __shared__ double sabs[2];

    const BkDesc<T> d = descs[blockIdx.x];
    const int n = d.n, k0 = d.k0, tid = threadIdx.x;
    const int e = min(k0 + nbp, n);
    T* __restrict__ A = d.A;
    T* __restrict__ W = d.W;
    const int64_t ld = d.ld;
    V* Dv = reinterpret_cast<V*>(d.D);
    V* Sv = reinterpret_cast<V*>(d.sub);
    int* piv = d.piv;
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;

    if (k0 == 0) {
        for (int i = tid; i < n; i += BK_THREADS) {
            piv[i] = i;
            Sv[i] = V{};
        }
        __syncthreads();
    }
    int k = k0, ks = k0;
    if (k0 > 0) {
        if (nonzero_v(Sv[k0 - 1])) {
            // a 2x2 pivot of the previous panel finished column k0
            for (int i = k0 + tid; i < n; i += BK_THREADS) W[i + (int64_t)k0 * ld] = T{};
            k = ks = k0 + 1;
            __syncthreads();
        } else {
            // column k0 still lacks the previous panel's update
            const int p0 = k0 - nbp, np = nbp;
            if (tid < np) wrow[tid] = cj(cvt<V>(W[k0 + (int64_t)(p0 + tid) * ld]), HERM);
            __syncthreads();
            for (int i = k0 + tid; i < n; i += BK_THREADS) {
                V acc = cvt<V>(A[i + (int64_t)k0 * ld]);
                for (int q = 0; q < np; ++q) acc -= cvt<V>(A[i + (int64_t)(p0 + q) * ld]) * wrow[q];
                A[i + (int64_t)k0 * ld] = cvt<T>(acc);
            }
            __syncthreads();
        }
    }
    while (k < e) {
        const int nw = k - ks;   // delayed updates from panel columns [ks, k)
        // (1) W(k:n, k) = A(k:n, k) - L(k:n, ks:k) op(W(k, ks:k))^T
        if (tid < nw) wrow[tid] = cj(cvt<V>(W[k + (int64_t)(ks + tid) * ld]), HERM);
        __syncthreads();
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += BK_THREADS) {
            V acc = cvt<V>(A[i + (int64_t)k * ld]);
            for (int q = 0; q < nw; ++q) acc -= cvt<V>(A[i + (int64_t)(ks + q) * ld]) * wrow[q];
            const T t = cvt<T>(acc);
            W[i + (int64_t)k * ld] = t;
            const double a = abs1(t);
            if (i == k) {
                sabs[0] = a;
            } else if (a > best) {
                best = a;
                bi = i;
            }
        }
        block_argmax(best, bi, sv, si);
        const double akk = sabs[0];
        const double lam = best > 0.0 ? best : 0.0;
        const int r = bi;
        int kstep = 1, kp = k;
        if (akk == 0.0 && lam == 0.0) {
            if (tid == 0) status_once(status, 2, d.front, k);
        } else if (akk < alpha * lam) {
            // (2) the updated column r into W(k:n, k+1): rows above r come
            // from row r of the lower triangle (op-transposed)
            if (tid < nw) wrow[tid] = cj(cvt<V>(W[r + (int64_t)(ks + tid) * ld]), HERM);
            __syncthreads();
            double sig = -1.0;
            int dummy = 0;
            for (int i = k + tid; i < n; i += BK_THREADS) {
                V v = (i < r) ? cj(cvt<V>(A[r + (int64_t)i * ld]), HERM) : cvt<V>(A[i + (int64_t)r * ld]);
                for (int q = 0; q < nw; ++q) v -= cvt<V>(A[i + (int64_t)(ks + q) * ld]) * wrow[q];
                const T t = cvt<T>(v);
                W[i + (int64_t)(k + 1) * ld] = t;
                const double a = abs1(t);
                if (i == r) sabs[1] = a;
                else sig = fmax(sig, a);
            }
            block_argmax(sig, dummy, sv, si);
            const double arr = sabs[1];
            if (akk * sig >= alpha * lam * lam) {
                kstep = 1;
                kp = k;
            } else if (arr >= alpha * sig) {
                kstep = 1;
                kp = r;
            } else {
                kstep = 2;
                kp = r;
            }
        }
        const int kk = k + kstep - 1;
        if (kstep == 1 && kp != k) {
            // 1x1 pivot on r: its column becomes column k
            for (int i = k + tid; i < n; i += BK_THREADS)
                W[i + (int64_t)k * ld] = W[i + (int64_t)(k + 1) * ld];
            __syncthreads();
        }
        if (kp != kk) {
            // (3) symmetric interchange of kk and kp
            for (int c = ks + tid; c <= kk; c += BK_THREADS) {
                const T a = W[kk + (int64_t)c * ld], b = W[kp + (int64_t)c * ld];
                W[kk + (int64_t)c * ld] = b;
                W[kp + (int64_t)c * ld] = a;
            }
            for (int c = tid; c < k; c += BK_THREADS) {
                const T a = A[kk + (int64_t)c * ld], b = A[kp + (int64_t)c * ld];
                A[kk + (int64_t)c * ld] = b;
                A[kp + (int64_t)c * ld] = a;
            }
            if (tid == 0) {
                A[kp + (int64_t)kp * ld] = A[kk + (int64_t)kk * ld];
                const int t = piv[kk];
                piv[kk] = piv[kp];
                piv[kp] = t;
            }
            for (int c = kk + 1 + tid; c < kp; c += BK_THREADS)
                A[kp + (int64_t)c * ld] = cj(A[c + (int64_t)kk * ld], HERM);
            for (int i = kp + 1 + tid; i < n; i += BK_THREADS)
                A[i + (int64_t)kp * ld] = A[i + (int64_t)kk * ld];
            __syncthreads();
        }
        // (4) L columns from W, and D
        if (kstep == 1) {
            V dk = cvt<V>(W[k + (int64_t)k * ld]);
            if (HERM) dk = real_only(dk);
            const V inv = recip(dk);
            for (int i = k + 1 + tid; i < n; i += BK_THREADS)
                A[i + (int64_t)k * ld] = cvt<T>(cvt<V>(W[i + (int64_t)k * ld]) * inv);
            if (tid == 0) {
                A[k + (int64_t)k * ld] = one<T>();
                Dv[k] = dk;
                Sv[k] = V{};
                if (is_zero_or_nonfinite(dk)) status_once(status, 2, d.front, k);
            }
        } else {
            V d11 = cvt<V>(W[k + (int64_t)k * ld]);
            V d22 = cvt<V>(W[(k + 1) + (int64_t)(k + 1) * ld]);
            const V d21 = cvt<V>(W[(k + 1) + (int64_t)k * ld]);
            if (HERM) {
                d11 = real_only(d11);
                d22 = real_only(d22);
            }
            const V d12 = cj(d21, HERM);
            const V det = d11 * d22 - d12 * d21;
            const V idet = recip(det);
            for (int i = k + 2 + tid; i < n; i += BK_THREADS) {
                const V w1 = cvt<V>(W[i + (int64_t)k * ld]), w2 = cvt<V>(W[i + (int64_t)(k + 1) * ld]);
                A[i + (int64_t)k * ld] = cvt<T>((w1 * d22 - w2 * d21) * idet);
                A[i + (int64_t)(k + 1) * ld] = cvt<T>((w2 * d11 - w1 * d12) * idet);
            }
            if (tid == 0) {
                A[k + (int64_t)k * ld] = one<T>();
                A[(k + 1) + (int64_t)(k + 1) * ld] = one<T>();
                A[(k + 1) + (int64_t)k * ld] = T{};
                Dv[k] = d11;
                Dv[k + 1] = d22;
                Sv[k] = d21;
                Sv[k + 1] = V{};
                if (is_zero_or_nonfinite(det)) status_once(status, 2, d.front, k);
            }
        }
        __syncthreads();
        k += kstep;
    }
    if (k == e && e < n)
        for (int i = e + tid; i < n; i += BK_THREADS) W[i + (int64_t)e * ld] = T{};
}

// Inverse of the unit lower triangle of a b x b diagonal block (b <= NB),
// written NB x NB (ld NB) with zeros outside the b x b lower triangle and
// the unit diagonal: X = L^{-1} column by column, X(i, j) =
// -sum_{j <= p < i} L(i, p) X(p, j), one thread per column, in FP64.
template <class T>
__global__ __launch_bounds__(NB) void k_unit_lower_inverse(const TriInvDesc<T>* __restrict__ descs) {
    using V = typename Scalar<T>::wide;
    __shared__ V Ls[NB * (NB + 1)];   // L(i, p) at Ls[i * (NB + 1) + p]
    __shared__ V Xs[NB * (NB + 1)];   // X(p, j) at Xs[j * (NB + 1) + p]
    const TriInvDesc<T> d = descs[blockIdx.x];
    const int j = threadIdx.x, b = d.b;
    for (int p = 0; p < NB; ++p) {
        // column p of the block, coalesced over the rows j
        Ls[j * (NB + 1) + p] = (j < b && p < j) ? cvt<V>(d.L[j + (int64_t)p * d.ld]) : V{};
        Xs[j * (NB + 1) + p] = (p == j && j < b) ? one<V>() : V{};
    }
    __syncthreads();
    V* xc = Xs + j * (NB + 1);
    for (int i = 1; i < b; ++i) {
        V acc{};
        const V* li = Ls + i * (NB + 1);
        for (int p = 0; p < i; ++p) acc += li[p] * xc[p];
        if (j < i) xc[i] = -acc;
    }
    __syncthreads();
    // Linv(t, c) = X(t, c): thread t writes row t of every column
    for (int c = 0; c < NB; ++c) d.Linv[j + c * NB] = cvt<T>(Xs[c * (NB + 1) + j]);
}

// dst(i, c) = src(i, perm[c]): one workgroup per (column, 256-row chunk)
template <class T>
__global__ __launch_bounds__(256) void k_gather_cols(const GatherColsDesc<T>* __restrict__ descs,
                                                     const int* __restrict__ start, int ndesc) {
    const int g = blockIdx.x;
    const int di = segment_of(start, ndesc, g);
    const GatherColsDesc<T> d = descs[di];
    const int chunks = (d.rows + 255) / 256;
    const int u = g - start[di];
    const int c = u / chunks, i = (u % chunks) * 256 + threadIdx.x;
    if (i < d.rows) d.dst[i + (int64_t)c * d.ldd] = d.src[i + (int64_t)d.perm[c] * d.lds];
}

// L(i, c) = (W D^{-1})(i, c) with D block diagonal of 1x1 and 2x2 pivots
// (a 2x2 pivot at (c, c+1) has sub[c] != 0): El::QuasiDiagonalSolve from the
// right (ldl/local_front.hpp:91). One workgroup per (column, 256-row chunk).
template <class T, bool HERM>
__global__ __launch_bounds__(256) void k_quasi_scale(const QuasiDesc<T>* __restrict__ descs,
                                                     const int* __restrict__ start, int ndesc) {
    using V = typename Scalar<T>::wide;
    const int g = blockIdx.x;
    const int di = segment_of(start, ndesc, g);
    const QuasiDesc<T> d = descs[di];
    const int chunks = (d.rows + 255) / 256;
    const int u = g - start[di];
    const int c = u / chunks, i = (u % chunks) * 256 + threadIdx.x;
    if (i >= d.rows) return;
    const V* D = reinterpret_cast<const V*>(d.D);
    const V* S = reinterpret_cast<const V*>(d.sub);
    const bool first = nonzero_v(S[c]);
    const bool second = c > 0 && nonzero_v(S[c - 1]);
    V out;
    if (first || second) {
        const int c0 = first ? c : c - 1;
        const V d11 = D[c0], d22 = D[c0 + 1], d21 = S[c0], d12 = cj(d21, HERM);
        const V idet = recip(d11 * d22 - d12 * d21);
        const V w1 = cvt<V>(d.W[i + (int64_t)c0 * d.ld]), w2 = cvt<V>(d.W[i + (int64_t)(c0 + 1) * d.ld]);
        out = first ? (w1 * d22 - w2 * d21) * idet : (w2 * d11 - w1 * d12) * idet;
    } else {
        out = cvt<V>(d.W[i + (int64_t)c * d.ld]) * recip(D[c]);
    }
    d.L[i + (int64_t)c * d.ld] = cvt<T>(out);
}

// Row permutation of each front's rows of the nodal right-hand sides:
// gather dst(off + i, j) = src(off + piv[off + i], j) (PermuteRows before the
// forward solve) or scatter dst(off + piv[off + i], j) = src(off + i, j)
// (InversePermuteRows after the backward solve).
template <class T>
__global__ void k_permute_front_rows(const PermSeg* __restrict__ segs, const int* __restrict__ row_start,
                                     int nseg, int nrows, const T* __restrict__ src, T* __restrict__ dst,
                                     int N, int k, const int* __restrict__ piv, bool scatter) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nrows) return;
    const int si = segment_of(row_start, nseg, g);
    const PermSeg sg = segs[si];
    const int i = g - row_start[si];
    const int p = sg.off + piv[sg.off + i];
    const int from = scatter ? sg.off + i : p, to = scatter ? p : sg.off + i;
    const int j0 = blockIdx.y * 8;
    for (int q = 0; q < 8 && j0 + q < k; ++q)
        dst[to + (int64_t)(j0 + q) * N] = src[from + (int64_t)(j0 + q) * N];
}

// X := D^{-1} X with D block diagonal (El::QuasiDiagonalSolve from the left,
// diagonal_solve.hpp:58-60); the thread of a 2x2 pivot's first row solves
// both rows.
template <class T, class V, bool HERM>
__global__ void k_quasi_diag_solve(T* __restrict__ xn, const V* __restrict__ D,
                                   const V* __restrict__ S, int N, int k, int* status) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        if (p > 0 && nonzero_v(S[p - 1])) continue;   // second row of a 2x2 pivot
        if (nonzero_v(S[p])) {
            const V d11 = D[p], d22 = D[p + 1], d21 = S[p], d12 = cj(d21, HERM);
            const V det = d11 * d22 - d12 * d21;
            if (is_zero_or_nonfinite(det)) status_once(status, 2, -1, p);
            const V idet = recip(det);
            T* x1 = xn + p;
            T* x2 = xn + p + 1;
            for (int j = 0; j < k; ++j) {
                const V a = cvt<V>(x1[(int64_t)j * N]), b = cvt<V>(x2[(int64_t)j * N]);
                x1[(int64_t)j * N] = cvt<T>((d22 * a - d12 * b) * idet);
                x2[(int64_t)j * N] = cvt<T>((d11 * b - d21 * a) * idet);
            }
        } else {
            const V dv = D[p];
            if (is_zero_or_nonfinite(dv)) status_once(status, 2, -1, p);
            const V r = recip(dv);
            T* xc = xn + p;
            for (int j = 0; j < k; ++j) xc[(int64_t)j * N] = cvt<T>(cvt<V>(xc[(int64_t)j * N]) * r);
        }
    }
}

// Z(i, c) = (D^{-1} X)(i, c): the block-LDL inverse's middle factor with
// 2x2 pivots (Trdtrmm with the quasi-diagonal D, ldl/local_front_block.hpp:53)
template <class T, bool HERM>
__global__ __launch_bounds__(256) void k_quasi_rows(const QuasiRowsDesc<T>* __restrict__ descs,
                                                    const int* __restrict__ start, int ndesc) {
    using V = typename Scalar<T>::wide;
    const int g = blockIdx.x;
    const int di = segment_of(start, ndesc, g);
    const QuasiRowsDesc<T> d = descs[di];
    const int chunks = (d.n + 255) / 256;
    const int u = g - start[di];
    const int c = u / chunks, i = (u % chunks) * 256 + threadIdx.x;
    if (i >= d.n) return;
    const V* D = reinterpret_cast<const V*>(d.D);
    const V* S = reinterpret_cast<const V*>(d.sub);
    const bool first = nonzero_v(S[i]);
    const bool second = i > 0 && nonzero_v(S[i - 1]);
    V out;
    if (first || second) {
        const int r0 = first ? i : i - 1;
        const V d11 = D[r0], d22 = D[r0 + 1], d21 = S[r0], d12 = cj(d21, HERM);
        const V idet = recip(d11 * d22 - d12 * d21);
        const V a = cvt<V>(d.X[r0 + (int64_t)c * d.ldx]), b = cvt<V>(d.X[r0 + 1 + (int64_t)c * d.ldx]);
        out = first ? (d22 * a - d12 * b) * idet : (d11 * b - d21 * a) * idet;
    } else {
        out = cvt<V>(d.X[i + (int64_t)c * d.ldx]) * recip(D[i]);
    }
    d.Z[i + (int64_t)c * d.ldz] = cvt<T>(out);
}

// A(piv[i], piv[j]) = Y(i, j)
template <class T>
__global__ __launch_bounds__(256) void k_permute_sym(const PermSymDesc<T>* __restrict__ descs,
                                                     const int* __restrict__ start, int ndesc) {
    const int g = blockIdx.x;
    const int di = segment_of(start, ndesc, g);
    const PermSymDesc<T> d = descs[di];
    const int chunks = (d.n + 255) / 256;
    const int u = g - start[di];
    const int j = u / chunks, i = (u % chunks) * 256 + threadIdx.x;
    if (i >= d.n) return;
    d.A[d.piv[i] + (int64_t)d.piv[j] * d.lda] = d.Y[i + (int64_t)j * d.ldy];
}

inline unsigned blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

template <class T>
void launch_bk_panel(const BkDesc<T>* d_descs, int ndesc, int nbp, bool herm, int* d_status,
                     hipStream_t s) {
    if (ndesc <= 0) return;
    if (herm)
        hipLaunchKernelGGL((k_bk_panel<T, true>), dim3(ndesc), dim3(BK_THREADS), 0, s, d_descs, nbp,
                           d_status);
    else
        hipLaunchKernelGGL((k_bk_panel<T, false>), dim3(ndesc), dim3(BK_THREADS), 0, s, d_descs, nbp,
                           d_status);
}

template <class T>
void launch_unit_lower_inverse(const TriInvDesc<T>* d_descs, int ndesc, hipStream_t s) {
    if (ndesc <= 0) return;
    hipLaunchKernelGGL(k_unit_lower_inverse<T>, dim3(ndesc), dim3(NB), 0, s, d_descs);
}

template <class T>
void launch_gather_cols(const GatherColsDesc<T>* d_descs, const int* d_start, int ndesc,
                        int nblocks, hipStream_t s) {
    if (ndesc <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_gather_cols<T>, dim3(nblocks), dim3(256), 0, s, d_descs, d_start, ndesc);
}

template <class T>
void launch_quasi_scale(const QuasiDesc<T>* d_descs, const int* d_start, int ndesc, int nblocks,
                        bool herm, hipStream_t s) {
    if (ndesc <= 0 || nblocks <= 0) return;
    if (herm)
        hipLaunchKernelGGL((k_quasi_scale<T, true>), dim3(nblocks), dim3(256), 0, s, d_descs,
                           d_start, ndesc);
    else
        hipLaunchKernelGGL((k_quasi_scale<T, false>), dim3(nblocks), dim3(256), 0, s, d_descs,
                           d_start, ndesc);
}

template <class T>
void launch_permute_front_rows(const PermSeg* d_segs, const int* d_row_start, int nseg, int nrows,
                               const T* src, T* dst, int N, int k, const int* d_piv, bool scatter,
                               hipStream_t s) {
    if (nrows <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_permute_front_rows<T>, dim3((nrows + 255) / 256, (k + 7) / 8), dim3(256), 0,
                       s, d_segs, d_row_start, nseg, nrows, src, dst, N, k, d_piv, scatter);
}

template <class T, class V>
void launch_quasi_diag_solve(T* xn, const V* D, const V* sub, int N, int k, bool herm,
                             int* d_status, hipStream_t s) {
    if (N <= 0 || k <= 0) return;
    if (herm)
        hipLaunchKernelGGL((k_quasi_diag_solve<T, V, true>), dim3(blocks_for(N)), dim3(256), 0, s, xn,
                           D, sub, N, k, d_status);
    else
        hipLaunchKernelGGL((k_quasi_diag_solve<T, V, false>), dim3(blocks_for(N)), dim3(256), 0, s,
                           xn, D, sub, N, k, d_status);
}

template <class T>
void launch_quasi_rows(const QuasiRowsDesc<T>* d_descs, const int* d_start, int ndesc,
                       int nblocks, bool herm, hipStream_t s) {
    if (ndesc <= 0 || nblocks <= 0) return;
    if (herm)
        hipLaunchKernelGGL((k_quasi_rows<T, true>), dim3(nblocks), dim3(256), 0, s, d_descs, d_start,
                           ndesc);
    else
        hipLaunchKernelGGL((k_quasi_rows<T, false>), dim3(nblocks), dim3(256), 0, s, d_descs,
                           d_start, ndesc);
}

template <class T>
void launch_permute_sym(const PermSymDesc<T>* d_descs, const int* d_start, int ndesc, int nblocks,
                        hipStream_t s) {
    if (ndesc <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_permute_sym<T>, dim3(nblocks), dim3(256), 0, s, d_descs, d_start, ndesc);
}

#define CQ_INST_PIV(T, V)                                                                        \
    template void launch_bk_panel<T>(const BkDesc<T>*, int, int, bool, int*, hipStream_t);       \
    template void launch_unit_lower_inverse<T>(const TriInvDesc<T>*, int, hipStream_t);          \
    template void launch_gather_cols<T>(const GatherColsDesc<T>*, const int*, int, int,          \
                                        hipStream_t);                                            \
    template void launch_quasi_scale<T>(const QuasiDesc<T>*, const int*, int, int, bool,         \
                                        hipStream_t);                                            \
    template void launch_permute_front_rows<T>(const PermSeg*, const int*, int, int, const T*,   \
                                               T*, int, int, const int*, bool, hipStream_t);     \
    template void launch_quasi_diag_solve<T, V>(T*, const V*, const V*, int, int, bool, int*,    \
                                                hipStream_t);                                    \
    template void launch_quasi_rows<T>(const QuasiRowsDesc<T>*, const int*, int, int, bool,     \
                                       hipStream_t);                                             \
    template void launch_permute_sym<T>(const PermSymDesc<T>*, const int*, int, int, hipStream_t);

CQ_INST_PIV(float, double)
CQ_INST_PIV(double, double)
CQ_INST_PIV(cfloat, cdouble)
CQ_INST_PIV(cdouble, cdouble)

}  // namespace cliqg
