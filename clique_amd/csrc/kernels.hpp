// Device kernels of libcliqg (gfx950) and their launchers.
//
// All work of one tree level is issued as a handful of grouped launches
// whose work lists (descriptor arrays) are built once per plan, so the
// factorization and the solves are fixed launch sequences.
#pragma once
#include "common.hpp"

namespace cliqg {

// Width of a diagonal block (panel) of the partial LDL^T.
constexpr int NB = 64;
// Tile shapes of the grouped GEMM (every launch uses one of them; the value
// is the `variant` of a launch and appears in profile traces).
enum GemmTile : int {
    TILE_64 = 0,      // 64x64 tiles, four waves of 32x32 (the default)
    TILE_128 = 4,     // 128x128 tiles, four waves of 64x64 (real types, large launches)
    TILE_32 = 5,      // 32x32 tiles, four waves of 16x16 (latency-bound launches)
    TILE_128_8W = 6,  // 128x128 tiles, eight waves of 64x32 (complex FP32)
    TILE_32_K64 = 7,  // 32x32 tiles with 64-deep K chunks (latency-bound launches
                      // with small K: one global round trip per 64 of K, not per 16)
    TILE_STREAM = 8,  // solve products with N <= 64 (FP32 types): 64 rows by all
                      // columns, op(A) streamed from global memory into the MFMAs
};
int gemm_tile_m(int variant);
int gemm_tile_n(int variant);
// The shapes built for each working type: 128x128 tiles of four waves for
// the real types (a complex 64x64 wave tile would not keep its
// three-multiplication products in registers), of eight waves for complex
// FP32 (three multiplications per block on 64x32 wave tiles); complex FP64
// keeps 64x64 tiles (a 128x128 complex FP64 tile would fill the LDS).
template <class T>
constexpr bool gemm_variant_allowed(int variant) {
    if (variant == TILE_128) return !Scalar<T>::complex;
    if (variant == TILE_128_8W) return Scalar<T>::complex && sizeof(T) == 8;
    // 64-deep chunks of complex FP64 would need 192 KB of LDS
    if (variant == TILE_32_K64) return !(Scalar<T>::complex && sizeof(T) == 16);
    if (variant == TILE_STREAM) return sizeof(typename Scalar<T>::real) == 4;
    return variant == TILE_64 || variant == TILE_32;
}
// Every tile shape, in the order the autotuner tries them
constexpr int NUM_GEMM_VARIANTS = 6;
constexpr int GEMM_VARIANTS[NUM_GEMM_VARIANTS] = {TILE_32, TILE_64, TILE_128, TILE_128_8W,
                                                  TILE_32_K64, TILE_STREAM};

enum GemmFlags : int {
    G_TRANSA = 1,   // op(A)(i,k) = A[k + i*lda]   (else A[i + k*lda])
    G_TRANSB = 2,   // op(B)(k,j) = B[j + k*ldb]   (else B[k + j*ldb])
    G_CONJA = 4,
    G_CONJB = 8,
    G_LOWER = 16,   // compute/write only C(i,j) with i >= j
    G_ACCUM = 32,   // C := C + alpha*op(A)op(B)  (else C := alpha*op(A)op(B))
    G_NEG = 64,     // alpha = -1 (else +1)
    G_INVK = 128,   // scale op(A)'s column k by 1/scaleK[k] instead of scaleK[k]
    G_KLO = 256,    // op(A)(i,k) = 0 for k < i: a row tile starts K at its first row
    G_KHI = 512,    // op(A)(i,k) = 0 for k > i: a row tile stops K after its last row
    G_SYMM = 1024,  // with G_LOWER: also write C(j,i) = C(i,j) (conjugated with G_CONJA)
};
// 16-row blocks per wave of a streaming tile (four waves: 64 x this rows)
constexpr int STREAM_ROW_BLOCKS = 1;   // two measured slower on the 120^3 PML solves (28.9 ms vs 24.7)
// Flags a streaming tile accepts (a plain or transposed A, a triangular K
// range, no output masks)
constexpr int STREAM_FLAGS = G_TRANSA | G_CONJA | G_ACCUM | G_NEG | G_KLO | G_KHI;

// C(MxN) (+)= +-op(A)(MxK) diag(scaleK) op(B)(KxN) [diag(scaleN)^{-1}], where
// op(A) is A, A^T, conj(A) or A^H per the flags (scaleK is never conjugated).
template <class T>
struct GemmDesc {
    const T* A;
    const T* B;
    T* C;
    const void* scaleK;   // FP64 field-type pivots multiplying op(A)'s column k (or null)
    const void* scaleN;   // FP64 field-type pivots dividing output column j (or null)
    int64_t lda, ldb, ldc;
    int M, N, K;
    int flags;
    T* C2 = nullptr;      // with scaleN: optional second output (ld ldc), C * diag(scaleN)
};

// Dense copy dst(0:rows, 0:cols) := src(0:rows, 0:cols), both column-major.
template <class T>
struct CopyDesc {
    const T* src;
    T* dst;
    int64_t lds, ldd;
    int rows, cols;
};

// Sum of the K slices of a split product: C = [C +] (-)sum_s W_s, summed in
// slice order (slice s at ws + s*rows*cols, ld rows); flags G_ACCUM, G_NEG.
template <class T>
struct ReduceDesc {
    const T* ws;
    T* C;
    int64_t ldc;
    int rows, cols, slices;
    int flags;
};

// Child-to-parent extend-add of one child's update matrix
// (reference ldl/local.hpp:60-102).
template <class T>
struct ExtendDesc {
    const T* Uc;      // uc x uc, ld uc, lower triangle
    const int* rel;   // uc relative indices into the parent's front
    T* Lp;            // parent frontL (mp x np, ld mp)
    T* Up;            // parent update matrix (up x up, ld up)
    int uc, np, mp, up;
    int j0, nc;       // child columns j0..j0+nc-1 (col_start counts column pairs)
};

// One front of a diagonal-block factorization step.
template <class T>
struct DiagDesc {
    T* A11;          // frontL + k + k*m
    void* D;         // FP64 pivots of this block (field type)
    T* Linv;         // NB x NB inverse of the unit lower block (ld NB)
    int64_t lda;
    int b;
    int front;       // for error reports
    int col0;        // k, for error reports
};

// Per-front segment of a solve operation: the update part W_s (u x k,
// ld u) of the front's solve workspace and its lowerStruct. The front's own
// rows live directly in the nodal right-hand sides X (ld N).
struct SolveSeg {
    int64_t w;       // offset of W_s in the workspace pool (per right-hand side)
    int64_t low;     // offset of lowerStruct in the index pool
    int u, pad;
};

// Child update rows added into the parent (lower_solve/local.hpp:83-104):
// relative indices below np address the parent's rows of X (at offp), the
// others the parent's W.
struct SolveExtend {
    int64_t wc;      // child workspace offset (per right-hand side)
    int64_t wp;      // parent workspace offset (per right-hand side)
    int64_t rel;     // offset of the child's relative indices
    int uc, np, offp, up;
    int i0;          // first child row of this launch (its count is in row_start)
};

// One front's panel of a fused solve step (launch_panel_step): the panel's
// rows w0 .. w0+sbw-1 and the band rows [b0, b1) of the front.
constexpr int PANEL_MAX = 128;   // widest panel of a fused step
template <class T>
struct PanelStepDesc {
    const T* LW;     // the panel's inverted unit-lower diagonal block, sbw x sbw (ld sbw)
    const T* Lp;     // forward: L(0, w0) (element (r, k) at r + k*ldl); backward: L(w0, 0)
                     // (element (k, r) at k + r*ldl)
    T* Xin;          // forward: X, backward: X2 (the front's rows, ld ldx)
    T* Xout;         // forward: X2, backward: X
    int64_t ldx;
    int ldl;
    int w0, sbw;
    int b0, b1;      // band rows
    int k;           // right-hand sides
};

// A contiguous range to clear.
template <class T>
struct ZeroSeg {
    T* p;
    int64_t n;
};
// Lower trapezoid of a column-major matrix: rows j..rows-1 of column j
template <class T>
struct ZeroLowerDesc {
    T* p;
    int64_t ld;
    int rows, cols;
};

// Entries one workgroup clears (a launch has one workgroup per chunk of each
// segment): large enough that the segment lookup and workgroup dispatch are
// small next to the stores.
constexpr int ZERO_CHUNK = 16384;

// ---- intra-front Bunch-Kaufman pivoting (pivot.hip) --------------------------
// Widest panel of the pivoted LDL^T of ATL.
constexpr int BK_MAX_PANEL = 64;

// One front of a panel step of the pivoted LDL^T of ATL
// (El::LDL(ATL, dSub, p, conjugate, BUNCH_KAUFMAN_A), ldl/local_front.hpp:86).
template <class T>
struct BkDesc {
    T* A;          // frontL (ld); ATL is rows and columns 0..n-1
    T* W;          // W = L D of ATL's columns (ld), a level work buffer
    void* D;       // n FP64 field-type entries: the diagonal of D
    void* sub;     // n FP64 field-type entries: D's subdiagonal (0 after a 1x1 pivot)
    int* piv;      // n local indices: (P A P^T)(i, j) = A(piv[i], piv[j])
    int64_t ld;
    int n;
    int k0;        // first column of the panel (a multiple of the panel width)
    int front;     // for error reports
};
// Inverse of the unit lower triangle of one b x b diagonal block (ld), into
// an NB x NB block (ld NB)
template <class T>
struct TriInvDesc {
    const T* L;
    T* Linv;
    int64_t ld;
    int b;
};
// dst(i, c) = src(i, perm[c]) for i < rows, c < cols (El::PermuteCols)
template <class T>
struct GatherColsDesc {
    const T* src;
    T* dst;
    const int* perm;
    int64_t lds, ldd;
    int rows, cols;
};
// L(i, c) = (W D^{-1})(i, c), D block diagonal (diagonal D, subdiagonal sub)
template <class T>
struct QuasiDesc {
    const T* W;
    T* L;
    const void* D;
    const void* sub;
    int64_t ld;
    int rows, cols;
};
// Z = D^{-1} X for an n x n block X (rows of one or two, D block diagonal)
template <class T>
struct QuasiRowsDesc {
    const T* X;
    T* Z;
    const void* D;
    const void* sub;
    int64_t ldx, ldz;
    int n;
};
// A(piv[i], piv[j]) = Y(i, j) for i, j < n (PermuteRows and PermuteCols of
// a symmetric block)
template <class T>
struct PermSymDesc {
    const T* Y;
    T* A;
    const int* piv;
    int64_t ldy, lda;
    int n;
};
// The solves' inverted diagonal block of a front with n <= 2 NB columns in
// one workgroup (launch_wide_inverse_small): LW = [X0 0; -X1 L10 X0 X1] with
// X0, X1 the inverses of the unit-lower 64-blocks.
template <class T>
struct WideInvDesc {
    const T* X0;     // inverse of L11's first 64-block (NB x NB, ld NB)
    const T* X1;     // of its second block (rows and columns 64..n-1), if n > NB
    const T* L10;    // frontL rows 64..n-1 of columns 0..63 (ld ldl)
    T* LW;           // n x n, ld n
    int64_t ldl;
    int n;
};
// A front's rows [off, off + n) of the nodal right-hand sides
struct PermSeg {
    int off, n;
};

template <class T>
void launch_bk_panel(const BkDesc<T>* d_descs, int ndesc, int nbp, bool herm, int* d_status,
                     hipStream_t s);
template <class T>
void launch_unit_lower_inverse(const TriInvDesc<T>* d_descs, int ndesc, hipStream_t s);
template <class T>
void launch_wide_inverse_small(const WideInvDesc<T>* d_descs, int ndesc, hipStream_t s);
// start[i] = workgroups before descriptor i (one per column and 256-row chunk)
template <class T>
void launch_gather_cols(const GatherColsDesc<T>* d_descs, const int* d_start, int ndesc,
                        int nblocks, hipStream_t s);
template <class T>
void launch_quasi_scale(const QuasiDesc<T>* d_descs, const int* d_start, int ndesc, int nblocks,
                        bool herm, hipStream_t s);
// gather (scatter = false) dst(off + i) = src(off + piv[off + i]), or scatter
// dst(off + piv[off + i]) = src(off + i), for every segment and column
template <class T>
void launch_permute_front_rows(const PermSeg* d_segs, const int* d_row_start, int nseg, int nrows,
                               const T* src, T* dst, int N, int k, const int* d_piv, bool scatter,
                               hipStream_t s);
template <class T, class V>
void launch_quasi_diag_solve(T* xn, const V* D, const V* sub, int N, int k, bool herm,
                             int* d_status, hipStream_t s);
// start[i] = workgroups before descriptor i (one per column and 256-row chunk)
template <class T>
void launch_quasi_rows(const QuasiRowsDesc<T>* d_descs, const int* d_start, int ndesc,
                       int nblocks, bool herm, hipStream_t s);
template <class T>
void launch_permute_sym(const PermSymDesc<T>* d_descs, const int* d_start, int ndesc, int nblocks,
                        hipStream_t s);

// ---- launchers (stream-ordered, asynchronous) ----------------------------------
// Clears every segment; chunk_start[i] = number of 1024-element chunks before
// segment i.
template <class T>
void launch_zero_segments(const ZeroSeg<T>* d_segs, const int* d_chunk_start, int nseg,
                          int nchunks, hipStream_t s);
// role: 0 factor TRSM-like products, 1 solve, 2 factor with every descriptor
// scaled by scaleK, 3 factor updates (separate instantiations keep the
// unscaled kernels' register budget and name the launches in traces).
template <class T>
void launch_gemm(const GemmDesc<T>* d_descs, const int* d_tile_start, int ndesc, int ntiles,
                 int variant, int role, hipStream_t s);
// chunk_start[i] = number of 1024-element chunks before descriptor i
template <class T>
void launch_copy2d(const CopyDesc<T>* d_descs, const int* d_chunk_start, int ndesc, int nchunks,
                   hipStream_t s);
template <class T>
void launch_zero_lower(const ZeroLowerDesc<T>* d_descs, const int* d_col_start, int ndesc,
                       int ncols, hipStream_t s);
template <class T>
void launch_reduce_slices(const ReduceDesc<T>* d_descs, const int* d_chunk_start, int ndesc,
                          int nchunks, hipStream_t s);
// number of output tiles of a descriptor (host side, must match the kernel)
int gemm_tiles(int M, int N, int flags, int variant);

template <class T, class V>
void launch_assemble(const V* vals, const int64_t* dst, int64_t nnz, T* pool, bool conj,
                     hipStream_t s);

template <class T>
void launch_extend_add(const ExtendDesc<T>* d_descs, const int* d_col_start, int ndesc,
                       int ncols, hipStream_t s);

template <class T>
void launch_diag_ldl(const DiagDesc<T>* d_descs, int ndesc, bool herm, int* d_status,
                     hipStream_t s);

// solve helpers; T working type, V FP64 field type
template <class T, class V>
void launch_pull(const V* x, int64_t ldx, const int* inv_map, T* xn, int N, int k, hipStream_t s);
template <class T, class V>
void launch_push(const T* xn, const int* inv_map, V* x, int64_t ldx, int N, int k, hipStream_t s);
// W_s := X(lowerStruct_s, :) for every segment (backward sweep)
template <class T>
void launch_gather(const SolveSeg* segs, const int* row_start, int nseg, int nrows, T* wpool,
                   const T* xn, int N, int k, const int* lowpool, hipStream_t s);
template <class T>
void launch_solve_extend(const SolveExtend* d, const int* row_start, int nd, int nrows,
                         T* wpool, T* xn, int N, const int* relpool, int k, hipStream_t s);
// Fused panel step of the LDL solves' triangular sweeps (one task per 16
// right-hand sides of a panel, and per 64 band rows and 16 right-hand sides)
template <class T>
void launch_panel_step(const PanelStepDesc<T>* d, const int* task_start, int nd, int ntasks,
                       bool backward, bool conj, hipStream_t s);
template <class T, class V>
void launch_diag_solve(T* xn, const V* D, int N, int k, int* d_status, hipStream_t s);

// FP64 sparse kernels (original ordering, CSR)
template <class V>
void launch_spmv(int N, const int64_t* rowoff, const int* cols, const V* vals, double alpha,
                 const V* x, int64_t ldx, double beta, V* y, int64_t ldy, int k, hipStream_t s);
// y := alpha A x + beta b (b == nullptr: beta y); colmax (optional):
// colmax[j] = max(colmax[j], max_i |y(i, j)|) as FP64 bit patterns
template <class V>
void launch_spmv_ex(int N, const int64_t* rowoff, const int* cols, const V* vals, double alpha,
                    const V* x, int64_t ldx, double beta, const V* b, int64_t ldb, V* y,
                    int64_t ldy, int k, unsigned long long* colmax, hipStream_t s);
// xn(map[i], j) = b(i, j) - (A x)(i, j) in the working type, with
// colmax[j] = max(colmax[j], max_i |b - A x|(i, j)) (a refinement step's
// residual, handed to the solve in nodal order)
template <class T, class V>
void launch_residual_nodal(int N, const int64_t* rowoff, const int* cols, const V* vals,
                           const V* x, int64_t ldx, const V* b, int64_t ldb, int k,
                           unsigned long long* colmax, T* xn, const int* map, hipStream_t s);
// x(i, j) += xn(map[i], j) with colmax[j] = max(colmax[j], max_i |x(i, j)|)
template <class T, class V>
void launch_push_add_abs(const T* xn, const int* map, V* x, int64_t ldx, int N, int k,
                         unsigned long long* colmax, hipStream_t s);
template <class V>
void launch_row_abs_sum_max(int N, const int64_t* rowoff, const V* vals, unsigned long long* out,
                            hipStream_t s);
template <class V>
void launch_col_abs_max(const V* x, int64_t ldx, int N, int k, unsigned long long* out,
                        hipStream_t s);
template <class V>
void launch_copy_cols(V* y, int64_t ldy, const V* x, int64_t ldx, int N, int k, hipStream_t s);
template <class V>
void launch_adjoint_dA(int N, const int64_t* rowoff, const int* cols, const V* lambda,
                       int64_t ldl, const V* x, int64_t ldx, int k, bool conj, V* dA,
                       hipStream_t s);
template <class T, class V>
void launch_widen(const T* src, V* dst, int64_t n, hipStream_t s);
// Row permutation of an N x k block: gather dst(p, j) = src(perm[p], j), or
// (scatter) dst(perm[p], j) = src(p, j)
template <class V>
void launch_permute_rows(const V* src, int64_t lds, V* dst, int64_t ldd, const int* perm, int N,
                         int k, bool scatter, hipStream_t s);
// out[j] = sum_i |x(i, j)|^2, one workgroup per column (summation order fixed)
template <class V>
void launch_col_sumsq(const V* x, int64_t ldx, int N, int k, double* out, hipStream_t s);
// p[0..n) := value (a kernel, so it can be recorded in a graph)
void launch_fill_int(int* p, int n, int value, hipStream_t s);

}  // namespace cliqg
