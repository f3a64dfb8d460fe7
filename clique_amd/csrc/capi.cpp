// extern "C" boundary of libcliqg (declared in include/cliqg.h).
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/cliqg.h"
#include "plan.hpp"

struct cliqg_symbolic_s {
    std::shared_ptr<const cliqg::Symbolic> sym;
};
struct cliqg_plan_s {
    std::unique_ptr<cliqg::Plan> plan;
};

namespace {

thread_local std::string g_last_error;

template <class F>
cliqg_status guarded(F&& f) {
    try {
        f();
        return CLIQG_OK;
    } catch (const cliqg::Error& e) {
        g_last_error = e.what();
        return static_cast<cliqg_status>(e.code);
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return CLIQG_ERR_OOM;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return CLIQG_ERR_INVALID;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CLIQG_ERR_INVALID;
    }
}

void require(bool cond, const char* msg) {
    if (!cond) throw cliqg::Error(CLIQG_ERR_INVALID, msg);
}

const cliqg::Symbolic& S(cliqg_symbolic s) {
    require(s && s->sym, "null symbolic handle");
    return *s->sym;
}
cliqg::Plan& P(cliqg_plan p) {
    require(p && p->plan, "null plan handle");
    return *p->plan;
}

}  // namespace

extern "C" {

const char* cliqg_last_error(void) { return g_last_error.c_str(); }

cliqg_status cliqg_version(int* major, int* minor) {
    return guarded([&] {
        require(major && minor, "null argument");
        *major = 0;
        *minor = 1;
    });
}

cliqg_status cliqg_device_malloc(void** d_ptr, size_t bytes) {
    return guarded([&] {
        require(d_ptr, "null argument");
        *d_ptr = nullptr;
        if (bytes) CQ_HIP(hipMalloc(d_ptr, bytes));
    });
}
cliqg_status cliqg_device_free(void* d_ptr) {
    return guarded([&] {
        if (d_ptr) CQ_HIP(hipFree(d_ptr));
    });
}
cliqg_status cliqg_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    return guarded([&] {
        require(kind >= 0 && kind <= 2, "bad copy kind");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                : kind == 1 ? hipMemcpyDeviceToHost
                                            : hipMemcpyDeviceToDevice;
        if (bytes) CQ_HIP(hipMemcpy(dst, src, bytes, k));
    });
}
cliqg_status cliqg_device_synchronize(void) {
    return guarded([&] { CQ_HIP(hipDeviceSynchronize()); });
}

cliqg_status cliqg_symbolic_natural(int nx, int ny, int nz, int N, const int64_t* h_row_off,
                                    const int* h_cols, int cutoff, cliqg_symbolic* out) {
    return guarded([&] {
        require(out && h_row_off && (h_cols || N == 0), "null argument");
        require(N >= 0 && (int64_t)nx * ny * nz == N, "grid dimensions do not match N");
        auto sym = std::make_shared<cliqg::Symbolic>();
        const cliqg::Graph g = cliqg::graph_from_csr(N, h_row_off, h_cols);
        cliqg::natural_nested_dissection(nx, ny, nz, g, cutoff, *sym);
        *out = new cliqg_symbolic_s{std::move(sym)};
    });
}

cliqg_status cliqg_symbolic_graph(int N, const int64_t* h_row_off, const int* h_cols, int cutoff,
                                  cliqg_symbolic* out) {
    return guarded([&] {
        require(out && h_row_off && h_cols && N > 0, "null argument");
        auto sym = std::make_shared<cliqg::Symbolic>();
        const cliqg::Graph g = cliqg::graph_from_csr(N, h_row_off, h_cols);
        cliqg::graph_nested_dissection(g, cutoff, *sym);
        *out = new cliqg_symbolic_s{std::move(sym)};
    });
}

cliqg_status cliqg_symbolic_concat(const cliqg_symbolic* syms, int count, cliqg_symbolic* out) {
    return guarded([&] {
        require(out && syms && count > 0, "null argument");
        auto all = std::make_shared<cliqg::Symbolic>();
        int Noff = 0, nodeOff = 0;
        for (int i = 0; i < count; ++i) {
            const cliqg::Symbolic& s = S(syms[i]);
            for (const auto& nd : s.nodes) {
                cliqg::SymNode c = nd;
                c.off += Noff;
                if (c.parent >= 0) c.parent += nodeOff;
                for (auto& x : c.children) x += nodeOff;
                for (auto& x : c.inds) x += Noff;
                for (auto& x : c.origLowerStruct) x += Noff;
                for (auto& x : c.lowerStruct) x += Noff;
                all->nodes.push_back(std::move(c));
            }
            for (int x : s.map) all->map.push_back(x + Noff);
            for (int x : s.inverseMap) all->inverseMap.push_back(x + Noff);
            Noff += s.N;
            nodeOff += (int)s.nodes.size();
        }
        all->N = Noff;
        int myOff = 0;
        for (auto& nd : all->nodes) {
            nd.myOff = myOff;
            myOff += nd.size;
        }
        *out = new cliqg_symbolic_s{std::move(all)};
    });
}

cliqg_status cliqg_symbolic_destroy(cliqg_symbolic sym) {
    return guarded([&] { delete sym; });
}

cliqg_status cliqg_symbolic_num_nodes(cliqg_symbolic sym, int* num_nodes, int* N) {
    return guarded([&] {
        const auto& s = S(sym);
        if (num_nodes) *num_nodes = (int)s.nodes.size();
        if (N) *N = s.N;
    });
}

cliqg_status cliqg_symbolic_node(cliqg_symbolic sym, int s, int* h_info) {
    return guarded([&] {
        const auto& sy = S(sym);
        require(h_info && s >= 0 && s < (int)sy.nodes.size(), "bad node index");
        const auto& nd = sy.nodes[s];
        h_info[0] = nd.size;
        h_info[1] = nd.off;
        h_info[2] = nd.parent;
        h_info[3] = nd.children.size() > 0 ? nd.children[0] : -1;
        h_info[4] = nd.children.size() > 1 ? nd.children[1] : -1;
        h_info[5] = (int)nd.lowerStruct.size();
        h_info[6] = (int)nd.origLowerStruct.size();
        h_info[7] = nd.onLeft ? 1 : 0;
    });
}

cliqg_status cliqg_symbolic_node_array(cliqg_symbolic sym, int s, int which, int* h_out) {
    return guarded([&] {
        const auto& sy = S(sym);
        require(s >= 0 && s < (int)sy.nodes.size(), "bad node index");
        const auto& nd = sy.nodes[s];
        const std::vector<int>* v = nullptr;
        switch (which) {
            case CLIQG_NODE_INDS: v = &nd.inds; break;
            case CLIQG_NODE_LOWER_STRUCT: v = &nd.lowerStruct; break;
            case CLIQG_NODE_ORIG_LOWER_STRUCT: v = &nd.origLowerStruct; break;
            case CLIQG_NODE_ORIG_LOWER_REL: v = &nd.origLowerRelInds; break;
            case CLIQG_NODE_LEFT_REL: v = &nd.leftRelInds; break;
            case CLIQG_NODE_RIGHT_REL: v = &nd.rightRelInds; break;
            default: require(false, "unknown node array");
        }
        if (!v->empty()) {
            require(h_out, "null output");
            std::memcpy(h_out, v->data(), v->size() * sizeof(int));
        }
    });
}

cliqg_status cliqg_symbolic_map(cliqg_symbolic sym, int* h_map) {
    return guarded([&] {
        const auto& sy = S(sym);
        require(h_map || sy.N == 0, "null output");
        if (sy.N) std::memcpy(h_map, sy.map.data(), sizeof(int) * sy.N);
    });
}

cliqg_status cliqg_symbolic_work(cliqg_symbolic sym, int is_complex, double* factor_flops,
                                 double* solve_flops_per_rhs, double* factor_entries) {
    return guarded([&] {
        const auto& sy = S(sym);
        double f = 0, s = 0, e = 0;
        for (const auto& nd : sy.nodes) {
            const double n = nd.size, m = n + (double)nd.lowerStruct.size();
            f += n * n * n / 3.0 + (m - n) * n * n + (m - n) * (m - n) * n;
            s += n * n + 2 * (m - n) * n;
            e += m * n;
        }
        if (is_complex) { f *= 4; s *= 4; }
        if (factor_flops) *factor_flops = f;
        if (solve_flops_per_rhs) *solve_flops_per_rhs = s;
        if (factor_entries) *factor_entries = e;
    });
}

cliqg_status cliqg_plan_create(cliqg_symbolic sym, const int64_t* h_row_off, const int* h_cols,
                               cliqg_field field, int conjugate, void* hip_stream,
                               cliqg_plan* out) {
    return guarded([&] {
        require(out && h_row_off, "null argument");
        require(field == CLIQG_REAL || field == CLIQG_COMPLEX, "unknown field");
        require(sym && sym->sym, "null symbolic handle");
        auto p = std::make_unique<cliqg::Plan>(sym->sym, h_row_off, h_cols, field == CLIQG_COMPLEX,
                                               conjugate != 0 && field == CLIQG_COMPLEX,
                                               static_cast<hipStream_t>(hip_stream));
        *out = new cliqg_plan_s{std::move(p)};
    });
}

cliqg_status cliqg_plan_destroy(cliqg_plan plan) {
    return guarded([&] { delete plan; });
}

cliqg_status cliqg_plan_memory(cliqg_plan plan, size_t* bytes) {
    return guarded([&] {
        require(bytes, "null argument");
        *bytes = P(plan).device_bytes();
    });
}

cliqg_status cliqg_plan_factor(cliqg_plan plan, const void* d_vals, cliqg_precision precision,
                               cliqg_front_type front_type) {
    return guarded([&] {
        auto& p = P(plan);
        require(d_vals || p.nnz == 0, "null values");
        p.factor(d_vals, (int)precision, (int)front_type);
    });
}

cliqg_status cliqg_plan_solve(cliqg_plan plan, void* d_x, int ldx, int k) {
    return guarded([&] {
        require(k >= 0 && (d_x || k == 0), "bad right-hand side");
        P(plan).solve(d_x, ldx, k, false);
    });
}

cliqg_status cliqg_plan_solve_nodal(cliqg_plan plan, void* d_x, int ldx, int k) {
    return guarded([&] {
        require(k >= 0 && (d_x || k == 0), "bad right-hand side");
        P(plan).solve(d_x, ldx, k, true);
    });
}

cliqg_status cliqg_plan_solve_noalloc(cliqg_plan plan, void* d_x, int ldx, int k) {
    return guarded([&] {
        require(k >= 0 && (d_x || k == 0), "bad right-hand side");
        P(plan).solve_noalloc(d_x, ldx, k);
    });
}

cliqg_status cliqg_plan_multiply(cliqg_plan plan, double alpha, const void* d_x, int ldx,
                                 double beta, void* d_y, int ldy, int k) {
    return guarded([&] {
        auto& p = P(plan);
        require(ldx >= p.N && ldy >= p.N && k >= 0, "bad leading dimension");
        p.multiply(alpha, d_x, ldx, beta, d_y, ldy, k);
        CQ_HIP(hipStreamSynchronize(p.stream));
    });
}

cliqg_status cliqg_spmv(int N, const int64_t* d_row_off, const int* d_cols, const void* d_vals,
                        cliqg_field field, double alpha, const void* d_x, int ldx, double beta,
                        void* d_y, int ldy, int k, void* hip_stream) {
    return guarded([&] {
        require(N >= 0 && k >= 0 && ldx >= N && ldy >= N, "bad dimensions");
        if (N == 0 || k == 0) return;
        require(d_row_off && d_cols && d_vals && d_x && d_y, "null argument");
        hipStream_t st = static_cast<hipStream_t>(hip_stream);
        if (field == CLIQG_COMPLEX)
            cliqg::launch_spmv<cliqg::cdouble>(N, d_row_off, d_cols, (const cliqg::cdouble*)d_vals,
                                               alpha, (const cliqg::cdouble*)d_x, ldx, beta,
                                               (cliqg::cdouble*)d_y, ldy, k, st);
        else
            cliqg::launch_spmv<double>(N, d_row_off, d_cols, (const double*)d_vals, alpha,
                                       (const double*)d_x, ldx, beta, (double*)d_y, ldy, k, st);
        CQ_HIP(hipGetLastError());
    });
}

cliqg_status cliqg_permute_rows(int N, int k, const int* d_perm, int scatter, const void* d_src,
                                int lds, void* d_dst, int ldd, cliqg_field field,
                                void* hip_stream) {
    return guarded([&] {
        require(N >= 0 && k >= 0 && lds >= N && ldd >= N, "bad dimensions");
        if (N == 0 || k == 0) return;
        require(d_perm && d_src && d_dst, "null argument");
        require(field == CLIQG_REAL || field == CLIQG_COMPLEX, "unknown field");
        const size_t vb = field == CLIQG_COMPLEX ? 16 : 8;
        const char* s0 = static_cast<const char*>(d_src);
        const char* d0 = static_cast<const char*>(d_dst);
        const char* s1 = s0 + ((size_t)lds * (k - 1) + N) * vb;
        const char* d1 = d0 + ((size_t)ldd * (k - 1) + N) * vb;
        require(s1 <= d0 || d1 <= s0, "source and destination overlap");
        hipStream_t st = static_cast<hipStream_t>(hip_stream);
        if (field == CLIQG_COMPLEX)
            cliqg::launch_permute_rows<cliqg::cdouble>((const cliqg::cdouble*)d_src, lds,
                                                       (cliqg::cdouble*)d_dst, ldd, d_perm, N, k,
                                                       scatter != 0, st);
        else
            cliqg::launch_permute_rows<double>((const double*)d_src, lds, (double*)d_dst, ldd,
                                               d_perm, N, k, scatter != 0, st);
        CQ_HIP(hipGetLastError());
    });
}

cliqg_status cliqg_col_norms(int N, int k, const void* d_x, int ldx, cliqg_field field,
                             double* h_norms, void* hip_stream) {
    return guarded([&] {
        require(N >= 0 && k >= 0 && ldx >= N, "bad dimensions");
        if (k == 0) return;
        require(d_x || N == 0, "null argument");
        require(h_norms, "null output");
        require(field == CLIQG_REAL || field == CLIQG_COMPLEX, "unknown field");
        hipStream_t st = static_cast<hipStream_t>(hip_stream);
        cliqg::DevBuf out((size_t)k * sizeof(double));
        if (field == CLIQG_COMPLEX)
            cliqg::launch_col_sumsq<cliqg::cdouble>((const cliqg::cdouble*)d_x, ldx, N, k,
                                                    out.as<double>(), st);
        else
            cliqg::launch_col_sumsq<double>((const double*)d_x, ldx, N, k, out.as<double>(), st);
        CQ_HIP(hipGetLastError());
        CQ_HIP(hipMemcpyAsync(h_norms, out.p, (size_t)k * sizeof(double), hipMemcpyDeviceToHost,
                              st));
        CQ_HIP(hipStreamSynchronize(st));
        for (int j = 0; j < k; ++j) h_norms[j] = std::sqrt(h_norms[j]);
    });
}

cliqg_status cliqg_plan_refine(cliqg_plan plan, const void* d_b, int ldb, void* d_x, int ldx,
                               int k, double tol, int max_iters, cliqg_refine_info* info) {
    return guarded([&] {
        auto& p = P(plan);
        require(ldb >= p.N && ldx >= p.N && k >= 0 && max_iters >= 0, "bad arguments");
        if (k > 0) {
            // x is written before every residual reads b: the two ranges must not overlap
            const size_t vb = p.complex ? 16 : 8;
            const char* b0 = static_cast<const char*>(d_b);
            const char* x0 = static_cast<const char*>(d_x);
            const char* b1 = b0 + ((size_t)ldb * (k - 1) + p.N) * vb;
            const char* x1 = x0 + ((size_t)ldx * (k - 1) + p.N) * vb;
            require(b1 <= x0 || x1 <= b0, "refinement needs x and b in distinct device memory");
        }
        int it = 0, conv = 0;
        double eta = 0.0;
        p.refine(d_b, ldb, d_x, ldx, k, tol, max_iters, &it, &conv, &eta);
        if (info) {
            info->iterations = it;
            info->converged = conv;
            info->max_backward_error = eta;
        }
        if (!conv && k > 0)
            throw cliqg::Error(CLIQG_ERR_NONCONVERGED,
                               "refinement stopped at backward error " + cliqg::sci(eta));
    });
}

cliqg_status cliqg_plan_adjoint(cliqg_plan plan, const void* d_g, int ldg, const void* d_x,
                                int ldx, int k, void* d_lambda, int ldl, void* d_dA) {
    return guarded([&] {
        auto& p = P(plan);
        require(ldg >= p.N && ldx >= p.N && ldl >= p.N && k >= 0, "bad leading dimension");
        p.adjoint(d_g, ldg, d_x, ldx, k, d_lambda, ldl, d_dA);
        CQ_HIP(hipStreamSynchronize(p.stream));
    });
}

cliqg_status cliqg_plan_set_front(cliqg_plan plan, int s, const void* h_L) {
    return guarded([&] { P(plan).set_front(s, h_L); });
}

cliqg_status cliqg_plan_get_front(cliqg_plan plan, int s, void* h_L, void* h_diag) {
    return guarded([&] { P(plan).get_front(s, h_L, h_diag); });
}

cliqg_status cliqg_plan_autotune(cliqg_plan plan, const void* d_vals, cliqg_precision precision,
                                 cliqg_front_type front_type, int solve_rhs, int* h_buckets,
                                 int* h_changed) {
    return guarded([&] {
        auto& p = P(plan);
        require(d_vals || p.nnz == 0, "null values");
        int nb = 0, nc = 0;
        p.autotune(d_vals, (int)precision, (int)front_type, solve_rhs, &nb, &nc);
        if (h_buckets) *h_buckets = nb;
        if (h_changed) *h_changed = nc;
    });
}

cliqg_status cliqg_autotune_table(char* h_buf, size_t capacity, size_t* h_length) {
    return guarded([&] {
        const std::string t = cliqg::tune_table_text();
        if (h_length) *h_length = t.size();
        if (h_buf && capacity > 0) {
            const size_t n = std::min(capacity - 1, t.size());
            std::memcpy(h_buf, t.data(), n);
            h_buf[n] = 0;
        }
    });
}

cliqg_status cliqg_plan_get_front_pivots(cliqg_plan plan, int s, void* h_subdiag, int* h_piv) {
    return guarded([&] { P(plan).get_pivots(s, h_subdiag, h_piv); });
}

cliqg_status cliqg_plan_last_factor_ms(cliqg_plan plan, double* ms) {
    return guarded([&] {
        require(ms, "null argument");
        *ms = P(plan).last_factor_ms();
    });
}

cliqg_status cliqg_plan_set_profiling(cliqg_plan plan, int enable) {
    return guarded([&] {
        auto& p = P(plan);
        p.profiling = enable != 0;
        p.profile.clear();
        p.pending.clear();
    });
}

cliqg_status cliqg_plan_set_graphs(cliqg_plan plan, int enable) {
    return guarded([&] { P(plan).useGraphs = enable != 0; });
}

cliqg_status cliqg_plan_get_profile(cliqg_plan plan, cliqg_profile_entry* h_out, int capacity,
                                    int* count) {
    return guarded([&] {
        auto& p = P(plan);
        require(count, "null argument");
        int n = 0;
        for (const auto& kv : p.profile) {
            if (h_out && n < capacity) {
                cliqg_profile_entry& e = h_out[n];
                std::memset(e.name, 0, sizeof(e.name));
                std::strncpy(e.name, kv.first.c_str(), sizeof(e.name) - 1);
                e.launches = kv.second.launches;
                e.total_ms = kv.second.ms;
                e.flops = kv.second.flops;
                e.bytes = kv.second.bytes;
            }
            ++n;
        }
        *count = n;
    });
}

}  // extern "C"
