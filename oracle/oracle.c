/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for the reference
 * functions restated and how the oracle is pinned). Plain C99 + OpenMP.
 */
#include "oracle.h"

#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small int vectors                                                         */
/* ------------------------------------------------------------------------ */
typedef struct { int* v; int n, cap; } ivec;

static void iv_push(ivec* a, int x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 8;
        a->v = (int*)realloc(a->v, sizeof(int) * (size_t)a->cap);
    }
    a->v[a->n++] = x;
}
static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}
static void iv_sort_unique(ivec* a) {
    if (a->n < 2) return;
    qsort(a->v, (size_t)a->n, sizeof(int), cmp_int);
    int w = 1;
    for (int i = 1; i < a->n; ++i)
        if (a->v[i] != a->v[w - 1]) a->v[w++] = a->v[i];
    a->n = w;
}

/* ------------------------------------------------------------------------ */
/* graphs: CSR with targets >= n meaning "outside the subgraph"              */
/* ------------------------------------------------------------------------ */
typedef struct { int n; long* off; int* tgt; } graph;

static void graph_free(graph* g) { free(g->off); free(g->tgt); g->off = 0; g->tgt = 0; }

/* ------------------------------------------------------------------------ */
/* tree                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int size, off, parent, nchild, child[2], onLeft;
    ivec inds, lower, origLower, origLowerRel, leftRel, rightRel;
} onode;

struct or_tree {
    int N, num;
    onode* nodes;
    int* map;
};

typedef struct { onode* v; int n, cap; } nodelist;

static int nl_push(nodelist* l) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->v = (onode*)realloc(l->v, sizeof(onode) * (size_t)l->cap);
    }
    memset(&l->v[l->n], 0, sizeof(onode));
    l->v[l->n].parent = -1;
    return l->n++;
}

/* the set {off + t : t >= n} over the listed local sources */
static void ancestors(const graph* g, const int* src, int ns, int off, ivec* out) {
    for (int i = 0; i < ns; ++i) {
        int s = src ? src[i] : i;
        for (long k = g->off[s]; k < g->off[s + 1]; ++k)
            if (g->tgt[k] >= g->n) iv_push(out, off + g->tgt[k]);
    }
    iv_sort_unique(out);
}

/* Box bisection: split the longest dimension (ties x > y > z), left part
 * gets (d-1)/2 planes, separator numbered last. */
static int bisect(int nx, int ny, int nz, int* dl, int* dr, int* perm) {
    int L, R, S, o;
    if (nx >= ny && nx >= nz) {
        int a = (nx - 1) / 2;
        dl[0] = a; dl[1] = ny; dl[2] = nz;
        dr[0] = nx - 1 - a; dr[1] = ny; dr[2] = nz;
        L = dl[0] * ny * nz; R = dr[0] * ny * nz; S = ny * nz;
        o = 0;
        for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < a; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L;
        for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = a + 1; x < nx; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L + R;
        for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y)
            perm[a + nx * (y + ny * z)] = o++;
    } else if (ny >= nx && ny >= nz) {
        int a = (ny - 1) / 2;
        dl[0] = nx; dl[1] = a; dl[2] = nz;
        dr[0] = nx; dr[1] = ny - 1 - a; dr[2] = nz;
        L = nx * a * nz; R = nx * dr[1] * nz; S = nx * nz;
        o = 0;
        for (int z = 0; z < nz; ++z) for (int y = 0; y < a; ++y) for (int x = 0; x < nx; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L;
        for (int z = 0; z < nz; ++z) for (int y = a + 1; y < ny; ++y) for (int x = 0; x < nx; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L + R;
        for (int z = 0; z < nz; ++z) for (int x = 0; x < nx; ++x)
            perm[x + nx * (a + ny * z)] = o++;
    } else {
        int a = (nz - 1) / 2;
        dl[0] = nx; dl[1] = ny; dl[2] = a;
        dr[0] = nx; dr[1] = ny; dr[2] = nz - 1 - a;
        L = nx * ny * a; R = nx * ny * dr[2]; S = nx * ny;
        o = 0;
        for (int z = 0; z < a; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L;
        for (int z = a + 1; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x)
            perm[x + nx * (y + ny * z)] = o++;
        o = L + R;
        for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x)
            perm[x + nx * (y + ny * a)] = o++;
    }
    return S;
}

/* child graph: sources inv[shift .. shift+size), targets renumbered */
static void child_graph(const graph* g, const int* perm, const int* inv, int shift,
                        int size, graph* c) {
    c->n = size;
    c->off = (long*)malloc(sizeof(long) * (size_t)(size + 1));
    long cnt = 0;
    for (int s = 0; s < size; ++s) {
        int src = inv[s + shift];
        cnt += g->off[src + 1] - g->off[src];
    }
    c->tgt = (int*)malloc(sizeof(int) * (size_t)(cnt > 0 ? cnt : 1));
    long w = 0;
    for (int s = 0; s < size; ++s) {
        c->off[s] = w;
        int src = inv[s + shift];
        for (long k = g->off[src]; k < g->off[src + 1]; ++k) {
            int t = g->tgt[k];
            c->tgt[w++] = (t < g->n ? perm[t] : t) - shift;
        }
    }
    c->off[size] = w;
}

static void nd(int nx, int ny, int nz, const graph* g, const int* perm, nodelist* list,
               int parent, int off, int cutoff) {
    int n = g->n;
    if (n <= cutoff) {
        int me = nl_push(list);
        onode* nd_ = &list->v[me];
        nd_->size = n; nd_->off = off; nd_->parent = parent;
        for (int i = 0; i < n; ++i) iv_push(&nd_->inds, perm[i]);
        ancestors(g, NULL, n, off, &nd_->origLower);
        return;
    }
    int* map = (int*)malloc(sizeof(int) * (size_t)n);
    int* inv = (int*)malloc(sizeof(int) * (size_t)n);
    int dl[3], dr[3];
    int S = bisect(nx, ny, nz, dl, dr, map);
    for (int s = 0; s < n; ++s) inv[map[s]] = s;
    int L = dl[0] * dl[1] * dl[2], R = dr[0] * dr[1] * dr[2];
    int me = nl_push(list);
    {
        onode* sp = &list->v[me];
        sp->size = S; sp->off = off + (n - S); sp->parent = parent; sp->nchild = 2;
        ancestors(g, inv + (n - S), S, off, &sp->origLower);
        for (int s = 0; s < S; ++s) iv_push(&sp->inds, perm[inv[s + (n - S)]]);
    }
    graph lg, rg;
    child_graph(g, map, inv, 0, L, &lg);
    child_graph(g, map, inv, L, R, &rg);
    int* lp = (int*)malloc(sizeof(int) * (size_t)(L > 0 ? L : 1));
    int* rp = (int*)malloc(sizeof(int) * (size_t)(R > 0 ? R : 1));
    for (int s = 0; s < L; ++s) lp[s] = perm[inv[s]];
    for (int s = 0; s < R; ++s) rp[s] = perm[inv[s + L]];
    free(map); free(inv);
    list->v[me].child[1] = list->n;
    nd(dr[0], dr[1], dr[2], &rg, rp, list, me, off + L, cutoff);
    graph_free(&rg); free(rp);
    list->v[me].child[0] = list->n;
    nd(dl[0], dl[1], dl[2], &lg, lp, list, me, off, cutoff);
    graph_free(&lg); free(lp);
}

static void iv_union(const ivec* a, const ivec* b, ivec* out) {
    int i = 0, j = 0;
    out->n = 0;
    while (i < a->n || j < b->n) {
        int x;
        if (j >= b->n || (i < a->n && a->v[i] < b->v[j])) x = a->v[i++];
        else if (i >= a->n || b->v[j] < a->v[i]) x = b->v[j++];
        else { x = a->v[i]; ++i; ++j; }
        iv_push(out, x);
    }
}

static void rel_inds(const ivec* sub, const ivec* full, ivec* out) {
    out->n = 0;
    int p = 0;
    for (int i = 0; i < sub->n; ++i) {
        while (p < full->n && full->v[p] < sub->v[i]) ++p;
        iv_push(out, p);
    }
}

or_tree* or_natural_nd(int nx, int ny, int nz, int N, const long* rowoff,
                       const int* cols, int cutoff) {
    if ((long)nx * ny * nz != N || cutoff < 1) return NULL;
    graph g;
    g.n = N;
    g.off = (long*)malloc(sizeof(long) * (size_t)(N + 1));
    long nnz = rowoff[N] - rowoff[0];
    g.tgt = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    for (int i = 0; i <= N; ++i) g.off[i] = rowoff[i] - rowoff[0];
    memcpy(g.tgt, cols + rowoff[0], sizeof(int) * (size_t)nnz);
    int* perm = (int*)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; ++i) perm[i] = i;
    nodelist list = {0, 0, 0};
    nd(nx, ny, nz, &g, perm, &list, -1, 0, cutoff);
    free(perm);
    graph_free(&g);

    or_tree* t = (or_tree*)calloc(1, sizeof(or_tree));
    t->N = N;
    t->num = list.n;
    t->nodes = (onode*)calloc((size_t)list.n, sizeof(onode));
    int last = list.n - 1;
    for (int i = 0; i < list.n; ++i) {
        onode* s = &list.v[i];
        onode* d = &t->nodes[last - i];
        *d = *s;
        d->parent = s->parent < 0 ? -1 : last - s->parent;
        for (int c = 0; c < s->nchild; ++c) d->child[c] = last - s->child[c];
    }
    free(list.v);
    t->map = (int*)malloc(sizeof(int) * (size_t)N);
    for (int s = 0; s < t->num; ++s) {
        onode* d = &t->nodes[s];
        for (int k = 0; k < d->inds.n; ++k) t->map[d->inds.v[k]] = d->off + k;
    }
    /* local symmetric analysis */
    for (int s = 0; s < t->num; ++s) {
        onode* d = &t->nodes[s];
        if (d->nchild == 2) {
            onode* l = &t->nodes[d->child[0]];
            onode* r = &t->nodes[d->child[1]];
            l->onLeft = 1; r->onLeft = 0;
            ivec cs = {0, 0, 0}, ps = {0, 0, 0}, full = {0, 0, 0}, ni = {0, 0, 0};
            iv_union(&l->lower, &r->lower, &cs);
            iv_union(&d->origLower, &cs, &ps);
            for (int i = 0; i < d->size; ++i) iv_push(&ni, d->off + i);
            iv_union(&ps, &ni, &full);
            rel_inds(&d->origLower, &full, &d->origLowerRel);
            rel_inds(&l->lower, &full, &d->leftRel);
            rel_inds(&r->lower, &full, &d->rightRel);
            d->lower.n = 0;
            for (int i = d->size; i < full.n; ++i) iv_push(&d->lower, full.v[i]);
            free(cs.v); free(ps.v); free(full.v); free(ni.v);
        } else {
            d->lower.n = 0;
            for (int i = 0; i < d->origLower.n; ++i) {
                iv_push(&d->lower, d->origLower.v[i]);
                iv_push(&d->origLowerRel, i + d->size);
            }
        }
    }
    return t;
}

void or_tree_free(or_tree* t) {
    if (!t) return;
    for (int s = 0; s < t->num; ++s) {
        onode* d = &t->nodes[s];
        free(d->inds.v); free(d->lower.v); free(d->origLower.v);
        free(d->origLowerRel.v); free(d->leftRel.v); free(d->rightRel.v);
    }
    free(t->nodes); free(t->map); free(t);
}

int or_tree_num_nodes(const or_tree* t) { return t->num; }

void or_tree_node(const or_tree* t, int s, int* info) {
    const onode* d = &t->nodes[s];
    info[0] = d->size; info[1] = d->off; info[2] = d->parent;
    info[3] = d->nchild > 0 ? d->child[0] : -1;
    info[4] = d->nchild > 1 ? d->child[1] : -1;
    info[5] = d->lower.n; info[6] = d->origLower.n; info[7] = d->onLeft;
}

int or_tree_array(const or_tree* t, int s, int which, int* out) {
    const onode* d = &t->nodes[s];
    const ivec* v;
    switch (which) {
        case OR_INDS: v = &d->inds; break;
        case OR_LOWER: v = &d->lower; break;
        case OR_ORIG_LOWER: v = &d->origLower; break;
        case OR_ORIG_LOWER_REL: v = &d->origLowerRel; break;
        case OR_LEFT_REL: v = &d->leftRel; break;
        case OR_RIGHT_REL: v = &d->rightRel; break;
        default: return -1;
    }
    if (out && v->n) memcpy(out, v->v, sizeof(int) * (size_t)v->n);
    return v->n;
}

void or_tree_map(const or_tree* t, int* map) { memcpy(map, t->map, sizeof(int) * (size_t)t->N); }

void or_factor_work(const or_tree* t, int is_complex, double* ff, double* sf, double* fe) {
    double f = 0, s = 0, e = 0;
    for (int i = 0; i < t->num; ++i) {
        double n = t->nodes[i].size, m = n + t->nodes[i].lower.n;
        f += n * n * n / 3.0 + (m - n) * n * n + (m - n) * (m - n) * n;
        s += n * n + 2 * (m - n) * n;
        e += m * n;
    }
    if (is_complex) { f *= 4; s *= 4; }
    *ff = f; *sf = s; *fe = e;
}

/* ------------------------------------------------------------------------ */
/* numeric part, instantiated for double and double complex                  */
/* ------------------------------------------------------------------------ */
struct or_factor {
    const or_tree* t;
    int is_complex, conjugate, block, factored;
    int plant_front, plant_panel;   /* planted mistake of the next LDL (-1: none) */
    int pivoted;   /* LDL_INTRAPIV: Bunch-Kaufman inside each front's ATL */
    void** L;      /* per front m x n */
    void** diag;   /* per front n     */
    void** sub;    /* per front n: subdiagonal of the 2x2 pivots (0 after 1x1) */
    int** piv;     /* per front n: the ATL permutation                        */
};

/* A multithreaded BLAS loaded at run time for the CPU-baseline path
 * (Fortran interfaces; the trailing lengths are the hidden character-argument
 * lengths of gfortran-built libraries, ignored by the others). */
typedef void (*gemm_fn)(const char*, const char*, const int*, const int*, const int*, const void*,
                        const void*, const int*, const void*, const int*, const void*, void*,
                        const int*, size_t, size_t);
typedef void (*trsm_fn)(const char*, const char*, const char*, const char*, const int*,
                        const int*, const void*, const void*, const int*, void*, const int*,
                        size_t, size_t, size_t, size_t);
typedef void (*gemmt_fn)(const char*, const char*, const char*, const int*, const int*,
                         const void*, const void*, const int*, const void*, const int*,
                         const void*, void*, const int*, size_t, size_t, size_t);
typedef struct {
    gemm_fn dgemm, zgemm;
    trsm_fn dtrsm, ztrsm;
    gemmt_fn dgemmt, zgemmt;
    int loaded;
} or_blas;
static or_blas g_blas;

#define T double
#define SUF d
#define BLP d
#define ABSV(x) fabs(x)
#define CONJ(x) (x)
#include "oracle_num.inc"
#include "oracle_blas.inc"
#undef T
#undef SUF
#undef BLP
#undef ABSV
#undef CONJ

#define T double complex
#define SUF z
#define BLP z
#define ABSV(x) cabs(x)
#define CONJ(x) (f->conjugate ? conj(x) : (x))
#include "oracle_num.inc"
#include "oracle_blas.inc"
#undef T
#undef SUF
#undef BLP
#undef ABSV
#undef CONJ

int or_blas_load(const char* path, const char* prefix, int threads) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    char name[128];
#define OR_SYM(field, type, base)                                   \
    snprintf(name, sizeof(name), "%s" base, prefix);                \
    g_blas.field = (type)dlsym(h, name);                            \
    if (!g_blas.field) return -2;
    OR_SYM(dgemm, gemm_fn, "dgemm_")
    OR_SYM(zgemm, gemm_fn, "zgemm_")
    OR_SYM(dtrsm, trsm_fn, "dtrsm_")
    OR_SYM(ztrsm, trsm_fn, "ztrsm_")
    OR_SYM(dgemmt, gemmt_fn, "dgemmt_")
    OR_SYM(zgemmt, gemmt_fn, "zgemmt_")
#undef OR_SYM
    /* MKL keeps its default (Intel OpenMP) threading layer: its GNU layer
     * ignored the thread count here */
    void (*mkl_threads)(int) = (void (*)(int))dlsym(h, "MKL_Set_Num_Threads");
    snprintf(name, sizeof(name), "%sopenblas_set_num_threads", prefix);
    void (*ob_threads)(int) = (void (*)(int))dlsym(h, name);
    if (threads > 0) {
        if (mkl_threads) mkl_threads(threads);
        if (ob_threads) ob_threads(threads);
    }
    g_blas.loaded = 1;
    return 0;
}

int or_ldl_blas(or_factor* f, int nb) {
    if (!g_blas.loaded || nb < 1) return -1;
    f->block = 0;
    f->pivoted = 0;
    int r = f->is_complex ? ldl_blas_z(f, nb) : ldl_blas_d(f, nb);
    if (r == 0) f->factored = 1;
    return r;
}

int or_solve_blas(or_factor* f, double* x, int ldx, int k) {
    if (!g_blas.loaded || !f->factored || f->block) return -1;
    return f->is_complex ? solve_blas_z(f, x, ldx, k) : solve_blas_d(f, x, ldx, k);
}

or_factor* or_factor_create(const or_tree* t, int is_complex, int conjugate) {
    or_factor* f = (or_factor*)calloc(1, sizeof(or_factor));
    f->t = t;
    f->is_complex = is_complex;
    f->conjugate = conjugate;
    f->plant_front = f->plant_panel = -1;
    f->L = (void**)calloc((size_t)t->num, sizeof(void*));
    f->diag = (void**)calloc((size_t)t->num, sizeof(void*));
    f->sub = (void**)calloc((size_t)t->num, sizeof(void*));
    f->piv = (int**)calloc((size_t)t->num, sizeof(int*));
    return f;
}

void or_factor_free(or_factor* f) {
    if (!f) return;
    for (int s = 0; s < f->t->num; ++s) {
        free(f->L[s]); free(f->diag[s]); free(f->sub[s]); free(f->piv[s]);
    }
    free(f->L); free(f->diag); free(f->sub); free(f->piv); free(f);
}

int or_initialize(or_factor* f, int N, const long* rowoff, const int* cols, const double* vals) {
    return f->is_complex ? init_z(f, N, rowoff, cols, vals) : init_d(f, N, rowoff, cols, vals);
}

int or_ldl(or_factor* f, int nb, int block, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (nb < 1) return -1;
    /* 0: FrontLDL, 1: FrontBlockLDL, 2: FrontLDLIntraPiv, 3: FrontBlockLDL
     * with intra-front pivoting */
    f->block = block == 1 || block == 3;
    f->pivoted = block == 2 || block == 3;
    int r = f->is_complex ? ldl_z(f, nb, f->block) : ldl_d(f, nb, f->block);
    if (r == 0) f->factored = 1;
    return r;
}

void or_plant(or_factor* f, int front, int panel) {
    f->plant_front = front;
    f->plant_panel = panel;
}

int or_solve(or_factor* f, double* x, int ldx, int k) {
    if (!f->factored) return -1;
    return f->is_complex ? solve_z(f, x, ldx, k) : solve_d(f, x, ldx, k);
}

int or_get_pivots(const or_factor* f, int s, double* sub, int* piv) {
    const size_t n = (size_t)f->t->nodes[s].size;
    const size_t es = f->is_complex ? 2 * sizeof(double) : sizeof(double);
    if (sub && n && f->sub[s]) memcpy(sub, f->sub[s], es * n);
    if (piv && n && f->piv[s]) memcpy(piv, f->piv[s], sizeof(int) * n);
    return 0;
}

int or_get_front(const or_factor* f, int s, double* L, double* diag) {
    const onode* d = &f->t->nodes[s];
    size_t n = (size_t)d->size, m = n + (size_t)d->lower.n;
    size_t es = f->is_complex ? 2 * sizeof(double) : sizeof(double);
    if (L && m * n > 0) memcpy(L, f->L[s], es * m * n);
    if (diag && n && f->diag[s]) memcpy(diag, f->diag[s], es * n);
    return 0;
}

void or_multiply(int N, const long* rowoff, const int* cols, const double* vals,
                 int is_complex, double alpha, const double* x, int ldx,
                 double beta, double* y, int ldy, int k) {
    if (is_complex) {
        const double complex* v = (const double complex*)vals;
        const double complex* X = (const double complex*)x;
        double complex* Y = (double complex*)y;
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < N; ++i) {
                double complex acc = 0;
                for (long e = rowoff[i]; e < rowoff[i + 1]; ++e)
                    acc += v[e] * X[cols[e] + (long)j * ldx];
                Y[i + (long)j * ldy] = beta * Y[i + (long)j * ldy] + alpha * acc;
            }
    } else {
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < N; ++i) {
                double acc = 0;
                for (long e = rowoff[i]; e < rowoff[i + 1]; ++e)
                    acc += vals[e] * x[cols[e] + (long)j * ldx];
                y[i + (long)j * ldy] = beta * y[i + (long)j * ldy] + alpha * acc;
            }
    }
}
