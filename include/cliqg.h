/*
 * cliqg.h -- C ABI of the MI355X multifrontal LDL^T engine (libcliqg.so).
 *
 * Clique exposes its hot path only as header-only C++ templates in
 * namespace cliq (include/clique.hpp:19-90); there is no FFI. This ABI is
 * the boundary those templates (and any other host language) bind to: plain
 * pointers, sizes and integer status codes, no C++ or HIP types. The
 * header-only C++ surface in include/clique.hpp is written on top of it.
 *
 * Conventions
 *  - Every function returns a cliqg_status; CLIQG_OK is 0. The message of the
 *    last failure on the calling thread is returned by cliqg_last_error().
 *  - "d_" arguments are device pointers (HBM of the plan's device);
 *    "h_" arguments are host pointers.
 *  - Matrices are column-major with an explicit leading dimension, as
 *    El::Matrix is. Complex values are interleaved (re, im) doubles or
 *    floats, the layout of std::complex.
 *  - Indices and orderings: "original" is the user's ordering of the
 *    unknowns; "nodal" is the nested-dissection ordering in which front s
 *    owns the contiguous rows [off_s, off_s + size_s) (DistNodalMultiVec,
 *    include/clique/numeric/dist_nodal_multi_vec/decl.hpp:19-51).
 *  - One plan serves one device and one stream. Calls on one plan must not
 *    overlap (the reference's Solve is likewise not reentrant on one tree:
 *    it writes the front tree's mutable work matrices,
 *    include/clique/numeric/dist_symm_front_tree/decl.hpp:183, 202-203).
 */
#ifndef CLIQG_H
#define CLIQG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CLIQG_OK = 0,
    CLIQG_ERR_INVALID = 1,       /* bad argument / API misuse (El::LogicError) */
    CLIQG_ERR_SINGULAR = 2,      /* zero or non-finite pivot                   */
    CLIQG_ERR_NONCONVERGED = 3,  /* refinement did not reach the tolerance     */
    CLIQG_ERR_OOM = 4,           /* device or host allocation failed           */
    CLIQG_ERR_HIP = 5,           /* HIP runtime error                          */
    CLIQG_ERR_STATE = 6          /* wrong plan state (e.g. already factored)   */
} cliqg_status;

typedef enum { CLIQG_REAL = 0, CLIQG_COMPLEX = 1 } cliqg_field;
typedef enum { CLIQG_F64 = 0, CLIQG_F32 = 1 } cliqg_precision;

/* SymmFrontType of the factored tree
 * (include/clique/numeric/dist_symm_front_tree/decl.hpp:17-47). 1D/2D
 * distinctions are layout choices for P>1 and collapse at world size 1. */
typedef enum {
    CLIQG_FRONT_LDL = 0,          /* LDL_1D / LDL_2D: unit L, diagonal D            */
    CLIQG_FRONT_BLOCK_LDL = 1,    /* BLOCK_LDL_*: ATL := inverse, ABL original A21   */
    CLIQG_FRONT_LDL_INTRAPIV = 2, /* LDL_INTRAPIV_*: Bunch-Kaufman pivoting inside
                                     each front's ATL (FrontLDLIntraPiv,
                                     ldl/local_front.hpp:73-94): unit L, D with
                                     1x1 and 2x2 blocks, a permutation per front  */
    CLIQG_FRONT_BLOCK_LDL_INTRAPIV = 3 /* BLOCK_LDL_INTRAPIV_*: FrontBlockLDL with
                                     intra-front pivoting (ldl/local_front_block.hpp:
                                     35-57): ATL := inverse of the front's ATL,
                                     ABL original A21; no pivots are kept          */
} cliqg_front_type;

/* which-array selectors for cliqg_symbolic_node_array */
typedef enum {
    CLIQG_NODE_INDS = 0,            /* original indices of the node (SepOrLeaf::inds) */
    CLIQG_NODE_LOWER_STRUCT = 1,    /* SymmNodeInfo::lowerStruct                      */
    CLIQG_NODE_ORIG_LOWER_STRUCT = 2,
    CLIQG_NODE_ORIG_LOWER_REL = 3,  /* origLowerRelInds                               */
    CLIQG_NODE_LEFT_REL = 4,        /* leftRelInds                                    */
    CLIQG_NODE_RIGHT_REL = 5        /* rightRelInds                                   */
} cliqg_node_array;

typedef struct cliqg_symbolic_s* cliqg_symbolic;
typedef struct cliqg_plan_s* cliqg_plan;

/* Per-RHS outcome of iterative refinement. */
typedef struct {
    int iterations;          /* refinement steps taken (0 = the first solve sufficed) */
    int converged;           /* all columns reached tol                               */
    double max_backward_error;  /* max over columns of ||b-Ax||_inf/(||A||_inf||x||_inf+||b||_inf) */
} cliqg_refine_info;

/* ---- library ------------------------------------------------------------ */
const char* cliqg_last_error(void);
cliqg_status cliqg_version(int* major, int* minor);
/* Device-memory helpers so hosts need no HIP toolchain of their own.
 * kind: 0 host->device, 1 device->host, 2 device->device. */
cliqg_status cliqg_device_malloc(void** d_ptr, size_t bytes);
cliqg_status cliqg_device_free(void* d_ptr);
cliqg_status cliqg_memcpy(void* dst, const void* src, size_t bytes, int kind);
cliqg_status cliqg_device_synchronize(void);

/* ---- symbolic analysis (host) --------------------------------------------
 * Replaces NaturalNestedDissection (include/clique/symbolic/
 * natural_nested_dissection.hpp:17-26, 447-488) at world size 1: separator
 * tree, reordering map and DistSymmInfo (dist_symm_info.hpp:17-113). The
 * graph is the CSR adjacency of A's pattern in the original ordering (the
 * reference uses A.DistGraph(), tests/Solve.cpp:123-131). */
cliqg_status cliqg_symbolic_natural(int nx, int ny, int nz, int N,
                                    const int64_t* h_row_off, const int* h_cols,
                                    int cutoff, cliqg_symbolic* out);
/* Nested dissection of a general graph: replaces NestedDissection
 * (include/clique/symbolic/nested_dissection.hpp:35-44, 491-533) at world
 * size 1. The recursion, node numbering, map and analysis are the
 * reference's; each bisection is this library's own (a breadth-first level
 * structure from a pseudo-peripheral vertex, its best-balancing level as the
 * separator, separator vertices with neighbours on one side only moved to
 * that side) instead of METIS's, so separators differ from the reference's
 * METIS orderings. The graph is A's CSR pattern (symmetric), as for
 * cliqg_symbolic_natural. */
cliqg_status cliqg_symbolic_graph(int N, const int64_t* h_row_off, const int* h_cols, int cutoff,
                                  cliqg_symbolic* out);
/* A forest of independent problems: symbolic i occupies nodal and original
 * indices shifted by the sizes of symbolics 0..i-1 (batched solves of many
 * unrelated systems in one schedule; no reference counterpart). */
cliqg_status cliqg_symbolic_concat(const cliqg_symbolic* syms, int count, cliqg_symbolic* out);
cliqg_status cliqg_symbolic_destroy(cliqg_symbolic sym);
cliqg_status cliqg_symbolic_num_nodes(cliqg_symbolic sym, int* num_nodes, int* N);
/* info[0..7] = size, off, parent, child0, child1, |lowerStruct|,
 *              |origLowerStruct|, onLeft   (SymmNodeInfo fields) */
cliqg_status cliqg_symbolic_node(cliqg_symbolic sym, int s, int* h_info);
/* copies the selected array (length from cliqg_symbolic_node) into h_out */
cliqg_status cliqg_symbolic_node_array(cliqg_symbolic sym, int s, int which, int* h_out);
/* map[original] = nodal (DistMap returned by NaturalNestedDissection) */
cliqg_status cliqg_symbolic_map(cliqg_symbolic sym, int* h_map);
/* FactorizationWork / SolveWork (one sweep, one RHS) and sum of m*n
 * (include/clique/numeric/dist_symm_front_tree/impl.hpp:540-639). */
cliqg_status cliqg_symbolic_work(cliqg_symbolic sym, int is_complex, double* factor_flops,
                                 double* solve_flops_per_rhs, double* factor_entries);

/* ---- plan: device front tree --------------------------------------------
 * Replaces the DistSymmFrontTree constructor (dist_symm_front_tree/
 * decl.hpp:216-221, impl.hpp:22-384): records A's pattern (CSR in the
 * original ordering, rows sorted and de-duplicated), builds the
 * value-to-front scatter map and the static level schedule, and uploads
 * them. `conjugate` selects Hermitian (LDL^H) instead of symmetric.
 * stream may be NULL (the plan creates its own). */
cliqg_status cliqg_plan_create(cliqg_symbolic sym, const int64_t* h_row_off, const int* h_cols,
                               cliqg_field field, int conjugate, void* hip_stream,
                               cliqg_plan* out);
cliqg_status cliqg_plan_destroy(cliqg_plan plan);
/* bytes of device memory the plan holds now */
cliqg_status cliqg_plan_memory(cliqg_plan plan, size_t* bytes);

/* Numeric factorization. Replaces DistSymmFrontTree::Initialize's scatter of
 * A into the fronts (impl.hpp:241-297) followed by cliq::LDL (ldl.hpp:57-75;
 * LocalLDL ldl/local.hpp:25-122, FrontLDL ldl/local_front.hpp:29-71,
 * FrontBlockLDL ldl/local_front_block.hpp:25-73). d_vals holds A's nnz values
 * (double or interleaved complex double) in the CSR order given at plan
 * creation. precision F32 factors in (complex) FP32 with FP64 pivots D.
 * Refactoring with new values is allowed (no "already factored" error: the
 * pattern is unchanged, the values are re-scattered). */
cliqg_status cliqg_plan_factor(cliqg_plan plan, const void* d_vals, cliqg_precision precision,
                               cliqg_front_type front_type);

/* Solve with the factor, in place. Replaces DistNodalMultiVec::Pull
 * (dist_nodal_multi_vec/impl.hpp:67-217), cliq::Solve (solve.hpp:40-92) and
 * Push (impl.hpp:219-356). d_x is N x k, column-major (ld = ldx), FP64 or
 * complex FP64, original ordering. */
cliqg_status cliqg_plan_solve(cliqg_plan plan, void* d_x, int ldx, int k);
/* The same on a nodal-ordered right-hand side (no Pull/Push). */
cliqg_status cliqg_plan_solve_nodal(cliqg_plan plan, void* d_x, int ldx, int k);

/* cliqg_plan_solve without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and results as cliqg_plan_solve, but every workspace
 * buffer (nodal right-hand sides, update rows, split-K slices, launch
 * descriptors, status word) lives in the library's static device array of
 * CLIQG_NOALLOC_SCRATCH_BYTES per device, held until the process ends; the
 * call allocates and frees no device memory (the first call for a plan and
 * right-hand-side count builds that schedule's descriptors on the host).
 * A solve whose workspace exceeds the array is rejected with
 * CLIQG_ERR_INVALID. All calls on one device share the array: calls that
 * overlap in time on one device race (run them on one stream, or in
 * sequence). The plan's device must be current. */
#define CLIQG_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
cliqg_status cliqg_plan_solve_noalloc(cliqg_plan plan, void* d_x, int ldx, int k);

/* y := alpha A x + beta y with the factored values, FP64 (replaces
 * cliq::Multiply, numeric/multiply.hpp:27-166, at world size 1). */
cliqg_status cliqg_plan_multiply(cliqg_plan plan, double alpha, const void* d_x, int ldx,
                                 double beta, void* d_y, int ldy, int k);

/* Standalone FP64 sparse product y := alpha A x + beta y on device CSR
 * arrays (original ordering), for Multiply without a plan. stream may be
 * NULL (the null stream); the call is asynchronous on it. */
cliqg_status cliqg_spmv(int N, const int64_t* d_row_off, const int* d_cols, const void* d_vals,
                        cliqg_field field, double alpha, const void* d_x, int ldx, double beta,
                        void* d_y, int ldy, int k, void* hip_stream);

/* Row permutation of an N x k device block (FP64 or complex FP64), the
 * device side of DistNodalMultiVec::Pull / Push at world size 1
 * (dist_nodal_multi_vec/impl.hpp:67-217, 219-356): with scatter = 0,
 * dst(p, j) = src(perm[p], j) (Pull with the inverse map); with scatter = 1,
 * dst(perm[p], j) = src(p, j) (Push). d_perm holds N device ints; src and dst
 * must not overlap. Asynchronous on hip_stream (NULL: the null stream). */
cliqg_status cliqg_permute_rows(int N, int k, const int* d_perm, int scatter, const void* d_src,
                                int lds, void* d_dst, int ldd, cliqg_field field,
                                void* hip_stream);

/* Two-norms of the k columns of an N x k device block (FP64 or complex
 * FP64) into h_norms (host, k doubles), summed on the device in a fixed
 * order (the role of DistMultiVec Norms, core/dist_multi_vec/impl.hpp).
 * Synchronizes hip_stream. */
cliqg_status cliqg_col_norms(int N, int k, const void* d_x, int ldx, cliqg_field field,
                             double* h_norms, void* hip_stream);

/* FP64 iterative refinement (no reference counterpart): x := solve(b); while
 * some column has backward error > tol and iterations < max_iters:
 * r := b - A x (FP64), x := x + solve(r) (FP64 update). Columns already at
 * or below tol keep their x: the correction solves take only the others
 * (once at least 16 columns have converged). b is not modified; x and b
 * must not overlap (CLIQG_ERR_INVALID otherwise). */
cliqg_status cliqg_plan_refine(cliqg_plan plan, const void* d_b, int ldb, void* d_x, int ldx,
                               int k, double tol, int max_iters, cliqg_refine_info* info);

/* Adjoint of x = A^{-1} b (no reference counterpart). Given g = dL/dx and the
 * solution x (N x k, original ordering): lambda := A^{-T} g (A^{-H} g when
 * the plan is Hermitian) with the same factor, dL/db = lambda written to
 * d_lambda, and dL/dA on A's pattern, one value per stored nonzero
 * (i, j): dA[e] = -sum_r lambda(i, r) * x(j, r)   (conj(x) when Hermitian).
 * Stored entries are treated as independent parameters. */
cliqg_status cliqg_plan_adjoint(cliqg_plan plan, const void* d_g, int ldg, const void* d_x,
                                int ldx, int k, void* d_lambda, int ldl, void* d_dA);

/* Lazy host mirror of front s after factorization: h_L gets the m x n
 * factor (column-major, FP64/complex FP64 regardless of the factor
 * precision), h_diag the n pivots (LDL fronts only; may be NULL). Replaces
 * reading frontTree.localFronts[s].frontL/diag (tests/Refactor.cpp:157-160). */
cliqg_status cliqg_plan_get_front(cliqg_plan plan, int s, void* h_L, void* h_diag);

/* Pivots of front s of a CLIQG_FRONT_LDL_INTRAPIV factorization: h_subdiag
 * gets D's n subdiagonal entries (FP64 or complex FP64; entry i is D(i+1, i),
 * nonzero exactly where a 2x2 pivot starts at i), h_piv the n local indices
 * of the front's permutation, (P ATL P^T)(i, j) = ATL(piv[i], piv[j]).
 * h_diag of cliqg_plan_get_front gets D's diagonal. Replaces reading
 * localFronts[s].subdiag and .piv (dist_symm_front_tree/decl.hpp:174-185,
 * filled by ldl/local.hpp:108-115). Either pointer may be NULL. */
cliqg_status cliqg_plan_get_front_pivots(cliqg_plan plan, int s, void* h_subdiag, int* h_piv);

/* Stage front s's unfactored values for the next cliqg_plan_factor, which
 * factors them in place of A's assembled values for that front (the other
 * fronts are assembled from A as usual); the staging is cleared by the
 * factorization. h_L is the m x n front (column-major, ld m, FP64 or complex
 * FP64); its lower trapezoid is read (rows >= column inside the n x n top
 * block, all rows below it). Replaces writing
 * frontTree.localFronts[s].frontL before LDL (tests/Refactor.cpp:150-160). */
cliqg_status cliqg_plan_set_front(cliqg_plan plan, int s, const void* h_L);

/* GEMM tile autotuner (no reference counterpart; Elemental's dense calls have
 * fixed blockings). Factors d_vals with the given precision and front type
 * while timing every grouped GEMM launch in each tile shape it admits
 * (32x32, 64x64, 128x128 of four waves for real types, 128x128 of eight
 * waves for complex FP32) and, with solve_rhs > 0, does the same for a solve
 * with solve_rhs right-hand sides; the fastest shape per launch bucket
 * (working type, role, lower masking, log2 of the launch's 64x64 tile count
 * and of its mean K) goes into a table shared by every plan of the process,
 * and every schedule built afterwards uses it. The plan is then factored
 * again with the tuned shapes (as cliqg_plan_factor). h_buckets: buckets
 * measured, h_changed: buckets whose shape changed (either may be NULL).
 * With the environment variable CLIQG_TUNE_CACHE=<file> the table is read
 * from that file at first use and written back after each tuning. */
cliqg_status cliqg_plan_autotune(cliqg_plan plan, const void* d_vals, cliqg_precision precision,
                                 cliqg_front_type front_type, int solve_rhs, int* h_buckets,
                                 int* h_changed);
/* The tuning table as text, one line per tuned bucket: bucket, chosen shape
 * (0: 64x64, 4: 128x128, 5: 32x32, 6: 128x128 eight-wave, 7: 32x32 with
 * 64-deep K chunks) and the milliseconds of the 32x32, 64x64, 128x128,
 * eight-wave 128x128 and 64-deep 32x32 shapes over the bucket's launches
 * (-1 where not built). h_length gets the full
 * length; at most capacity - 1 characters and a terminating zero are
 * copied to h_buf (which may be NULL to query the length). */
cliqg_status cliqg_autotune_table(char* h_buf, size_t capacity, size_t* h_length);

/* Timing of the last factorization on the plan's stream (HIP events). */
cliqg_status cliqg_plan_last_factor_ms(cliqg_plan plan, double* ms);

/* Opt-in profiling (the role of the reference's per-phase timers and
 * FactorizationWork accounting, tests/Solve.cpp:223-251): HIP events around
 * every scheduled launch, accumulated per kernel family with the launch's
 * algorithmic flops (FactorizationWork convention) and bytes. Enabling
 * resets the accumulators. */
typedef struct {
    char name[32];
    int launches;
    double total_ms;
    double flops;
    double bytes;
} cliqg_profile_entry;
cliqg_status cliqg_plan_set_profiling(cliqg_plan plan, int enable);
cliqg_status cliqg_plan_get_profile(cliqg_plan plan, cliqg_profile_entry* h_out, int capacity,
                                    int* count);

/* Launch replay (no reference counterpart; Clique has no GPU path): with
 * enable = 1 the factorization's and each solve width's launch sequences are
 * recorded once as HIP graphs on the plan's stream and replayed; the default
 * (0) launches every kernel individually. Results are identical. Capture is
 * skipped while profiling and on the legacy default stream. On MI355X the
 * replay measured no faster than direct launches for the 120^3 workload. */
cliqg_status cliqg_plan_set_graphs(cliqg_plan plan, int enable);

#ifdef __cplusplus
}
#endif
#endif
