/*
 * CPU oracle for the multifrontal LDL^T path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Clique's world-size-1 algorithm, written
 * independently of the product library. Only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * never links or calls it.
 *
 * Reference functions restated (file:line in poulson/Clique):
 *   or_natural_nd      symbolic/natural_nested_dissection.hpp:60-189,447-619,
 *                      symbolic/nested_dissection.hpp:817-963,1206-1365,
 *                      src/symbolic/LocalSymmetricAnalysis.cpp:15-114
 *   or_initialize      numeric/dist_symm_front_tree/impl.hpp:241-297
 *   or_ldl             numeric/ldl/local.hpp:25-122 (LocalLDL),
 *                      numeric/ldl/local_front.hpp:29-71 (FrontLDL),
 *                      numeric/ldl/local_front.hpp:73-94 (FrontLDLIntraPiv;
 *                      Bunch-Kaufman "A" of Elemental's El::LDL restated),
 *                      numeric/ldl/local_front_block.hpp:25-73 (FrontBlockLDL)
 *   or_solve           numeric/solve.hpp:40-65, lower_solve/local.hpp:41-119,
 *                      202-277, lower_solve/local_front.hpp:35-82,
 *                      lower_solve/local_front_block.hpp:28-72,
 *                      lower_solve/dist.hpp:374-387, diagonal_solve.hpp:48-68
 *   or_multiply        numeric/multiply.hpp:149-165
 *   or_ldl_blas,       the same LocalLDL and solves with FrontLDL's Trsm, Gemm
 *   or_solve_blas      and Trrk and the front solves' Trsm and Gemm made
 *                      through a CPU BLAS (the CPU baseline of bench.py)
 *   or_factor_work     numeric/dist_symm_front_tree/impl.hpp:555-585, 607-635
 *
 * Pinning (no reference golden vectors exist; Elemental, which holds the
 * dense kernels, is an absent submodule): the tests check this oracle
 * against closed forms (box-separator structure, 1D-chain pivots, Laplacian
 * eigenpairs), dense brute-force LDL^T of the permuted matrix, scipy's
 * sparse solver and structural invariants. See tests/test_oracle.py.
 */
#ifndef CLIQ_ORACLE_H
#define CLIQ_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_tree or_tree;
typedef struct or_factor or_factor;

/* which-array selectors for or_tree_array */
enum {
    OR_INDS = 0,            /* original indices of the node's vertices     */
    OR_LOWER = 1,           /* lowerStruct                                 */
    OR_ORIG_LOWER = 2,      /* origLowerStruct                             */
    OR_ORIG_LOWER_REL = 3,  /* origLowerRelInds                            */
    OR_LEFT_REL = 4,        /* leftRelInds                                 */
    OR_RIGHT_REL = 5        /* rightRelInds                                */
};

or_tree* or_natural_nd(int nx, int ny, int nz, int N, const long* rowoff,
                       const int* cols, int cutoff);
void or_tree_free(or_tree* t);
int or_tree_num_nodes(const or_tree* t);
/* info[0..7] = size, off, parent, child0, child1, |lowerStruct|,
 *              |origLowerStruct|, onLeft */
void or_tree_node(const or_tree* t, int s, int* info);
int or_tree_array(const or_tree* t, int s, int which, int* out);
void or_tree_map(const or_tree* t, int* map);

/* numeric: values are double, or interleaved complex double when is_complex */
or_factor* or_factor_create(const or_tree* t, int is_complex, int conjugate);
void or_factor_free(or_factor* f);
/* DistSymmFrontTree::Initialize from a CSR matrix in the original ordering */
int or_initialize(or_factor* f, int N, const long* rowoff, const int* cols,
                  const double* vals);
/* LocalLDL; block = 1 selects FrontBlockLDL, block = 2 FrontLDLIntraPiv
 * (Bunch-Kaufman inside each front's ATL, ldl/local_front.hpp:73-94);
 * nthreads <= 0: default */
int or_ldl(or_factor* f, int nb, int block, int nthreads);
/* front s of a pivoted factor: D's subdiagonal (n entries, 0 after a 1x1
 * pivot) and the permutation p, (P ATL P^T)(i,j) = ATL(p[i], p[j]) */
int or_get_pivots(const or_factor* f, int s, double* sub, int* piv);
/* Test hook: the next or_ldl leaves out the trailing update of panel
 * `panel` of front `front` (front < 0 clears it). Used only to show that the
 * tests' tolerances detect such a mistake. */
void or_plant(or_factor* f, int front, int panel);
/* Solve in nodal ordering, X column-major N x k with leading dimension ldx */
int or_solve(or_factor* f, double* x, int ldx, int k);
/* copy front s: L is m x n column-major (ld = m), diag n entries */
int or_get_front(const or_factor* f, int s, double* L, double* diag);
/* CPU baseline: the same LocalLDL / solves with the dense calls of
 * FrontLDL and the front solves made through a multithreaded BLAS loaded from
 * `path` (symbol names prefixed by `prefix`, e.g. "" for MKL, "scipy_" for
 * scipy's OpenBLAS) with `threads` threads. 0 on success. */
int or_blas_load(const char* path, const char* prefix, int threads);
int or_ldl_blas(or_factor* f, int nb);
int or_solve_blas(or_factor* f, double* x, int ldx, int k);
/* y := alpha A x + beta y (real alpha, beta) in the original ordering */
void or_multiply(int N, const long* rowoff, const int* cols, const double* vals,
                 int is_complex, double alpha, const double* x, int ldx,
                 double beta, double* y, int ldy, int k);
/* FactorizationWork and SolveWork (one sweep, one right-hand side) */
void or_factor_work(const or_tree* t, int is_complex, double* factor_flops,
                    double* solve_flops, double* factor_entries);

#ifdef __cplusplus
}
#endif
#endif
