"""Benchmark: multifrontal LDL^T factor TFLOP/s and time-to-1e-10 backward error,
FP32 factor + FP64 iterative refinement versus the plain FP64 path, on one GPU.

One step = factor A in (complex) FP32 on the GPU and refine the k right-hand
sides in FP64 until every column has backward error
||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= 1e-10 (inputs resident
in HBM). The same number of plain-FP64 steps (factor + solve, refined if a
column misses 1e-10) is timed after the FP32 steps. Factor flops follow
Clique's FactorizationWork convention (complex counted x4).

Default workload: 120^3 complex PML Helmholtz (7-point, 10-cell PML,
s = 1 + 20i (d/L)^2, omega = 2 pi 10), natural nested dissection with cutoff
128, 64 unit point sources at seeded interior nodes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LDL^T factor TFLOP/s & time-to-1e-10 backward error, FP32+IR vs FP64, 1 GPU"
PEAK = {"fp32": 157.3, "fp64": 78.6}   # MI355X dense MFMA TFLOP/s
HBM_PEAK = 8000.0                      # GB/s


def workload(cfg, n_override=None):
    from clique_amd import problems as P
    rng = np.random.default_rng(2024)
    if cfg == 1:
        n = n_override or 30
        ip, ind, v, dims = P.laplacian7(n)
        N = ip.size - 1
        b = rng.standard_normal((N, 1))
        desc = f"{n}^3 real 7-point Laplacian (diag 6.05), natural ND cutoff 128, 1 RHS"
        return ip, ind, v, dims, b, False, desc
    if cfg == 2:
        n = n_override or 64
        ip, ind, v, dims = P.helmholtz7(n)
        N = ip.size - 1
        b = rng.standard_normal((N, 4)) + 1j * rng.standard_normal((N, 4))
        desc = f"{n}^3 complex-symmetric Helmholtz (omega 2pi*6), natural ND cutoff 128, 4 RHS"
        return ip, ind, v, dims, b, True, desc
    if cfg == 3:
        import scipy.sparse as sp
        count = n_override or 512
        sizes = rng.integers(32, 193, size=(count, 2))
        shifts = rng.uniform(0.01, 0.5, size=count)
        mats, dims_list = [], []
        for (nx, ny), s_ in zip(sizes, shifts):
            ip_, ind_, v_, d_ = P.laplacian5(int(nx), int(ny), float(s_))
            mats.append(sp.csr_matrix((v_, ind_, ip_)))
            dims_list.append(d_)
        A = sp.block_diag(mats, format="csr")
        A.sort_indices()
        N = A.shape[0]
        b = rng.standard_normal((N, 16))
        desc = (f"{count} independent 2D 5-point problems, nx,ny ~ U{{32..192}}, diag 4+s, "
                f"one forest (natural ND cutoff 128 each), 16 RHS")
        return (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, dims_list, b,
                False, desc)
    if cfg == 4:
        n = n_override or 80
        ip, ind, v, dims = P.diffusion_grf(n, seed=0)
        N = ip.size - 1
        b = rng.standard_normal((N, 8))
        desc = f"{n}^3 GRF diffusion (7-point, kappa=exp(g)), natural ND cutoff 128, 8 RHS"
        return ip, ind, v, dims, b, False, desc
    n = n_override or 120
    ip, ind, v, dims = P.helmholtz_pml(n, omega=2 * np.pi * 10, pml=10, strength=20.0)
    N = ip.size - 1
    k = 64
    # interior nodes: at least the PML width from every face
    pts = []
    while len(pts) < k:
        x, y, z = rng.integers(10, n - 10, size=3)
        idx = int(x + n * (y + n * z))
        if idx not in pts:
            pts.append(idx)
    b = np.zeros((N, k), dtype=np.complex128)
    b[pts, np.arange(k)] = 1.0
    desc = (f"{n}^3 complex PML Helmholtz (7-point, 10-cell PML, s=1+20i(d/L)^2, omega 2pi*10), "
            f"natural ND cutoff 128, {k} point-source RHS")
    return ip, ind, v, dims, b, True, desc


def symbolic_for(ip, ind, dims, cutoff=128):
    """Natural nested dissection of a workload; a list of grid dimensions is
    a forest of independent problems (one symbolic per block, concatenated)."""
    import clique_amd as cq
    if not isinstance(dims, list):
        return cq.NaturalNestedDissection(*dims, ip, ind, cutoff=cutoff)
    syms, o = [], 0
    for d in dims:
        n_i = int(np.prod(d))
        sub_ip = ip[o:o + n_i + 1] - ip[o]
        sub_ind = ind[ip[o]:ip[o + n_i]] - o
        syms.append(cq.NaturalNestedDissection(*d, sub_ip, sub_ind, cutoff=cutoff))
        o += n_i
    return cq.Symbolic.concat(syms)


def pml_sample(n):
    """The config-5 operator scaled down to n^3: omega keeps config 5's points
    per wavelength (omega h = 2 pi 10 / 121), the PML its share of the grid."""
    from clique_amd import problems as P
    return P.helmholtz_pml(n, omega=2 * np.pi * 10 * (n + 1) / 121.0, pml=max(2, round(10 * n / 120)),
                           strength=20.0)


def cpu_baseline(budget_s=20.0, k=64, target=None):
    """The reference's CPU path restated (oracle/oracle_blas.inc: sequential
    postorder front loop, FrontLDL with nb = 96 making its Trsm / Gemm / Trrk
    calls through a multithreaded CPU BLAS, BLAS-backed front solves), FP64
    complex, on the largest sample of the config-5 operator family the budget
    allows. LDL and Solve (k point sources) are timed separately, as
    tests/Solve.cpp:223-302 does; target = (FactorizationWork, SolveWork per
    RHS) of the benchmarked workload, whose CPU times are extrapolated by the
    work ratio."""
    import oracle
    threads = min(16, len(os.sched_getaffinity(0)))
    blas = oracle.blas_load(threads)
    rng = np.random.default_rng(7)
    n, best = 28, None
    while True:
        ip, ind, v, dims = pml_sample(n)
        t = oracle.Tree(*dims, ip, ind, cutoff=128)
        f = oracle.Factor(t, is_complex=True)
        f.initialize(ip, ind, v)
        t0 = time.perf_counter()
        f.ldl_blas(nb=96)
        dl = time.perf_counter() - t0
        N = ip.size - 1
        b = np.zeros((N, k), dtype=np.complex128)
        b[rng.choice(N, size=k, replace=False), np.arange(k)] = 1.0
        t0 = time.perf_counter()
        f.solve_nodal_blas(b)
        ds = time.perf_counter() - t0
        fw, sw, _ = t.work(True)
        best = (n, fw, sw, dl, ds)
        del f, t
        nxt = int(n * 1.2)
        if (dl + ds) * (nxt / n) ** 6 > budget_s or n >= 120:
            break
        n = nxt
    n, fw, sw, dl, ds = best
    out = {"value": fw / dl / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
           "blas": blas, "ldl_s": dl, "solve_s": ds, "solve_rhs": k,
           "sample": (f"reference CPU path restated (FrontLDL nb=96 through {blas}), FP64 "
                      f"complex, {n}^3 PML Helmholtz sample: LDL {fw / 1e9:.1f} GFLOP in "
                      f"{dl:.2f} s, Solve of {k} point sources in {ds:.2f} s")}
    if target:
        ft, st = target
        out["extrapolated_to_workload"] = {"ldl_s": ft / (fw / dl), "solve_s": ds * st / sw,
                                           "time_to_solution_s": ft / (fw / dl) + ds * st / sw}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_gemm_update_pml120_fp32_traffic.json")


def measured_traffic(cfg, n, family):
    """HBM bytes per launch of the factor's update GEMM, from the committed
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE over `bench.py
    --factor-only`, each scaled by the same counter on a copy of known size);
    None for other workloads, for another kernel than the one measured, or
    when no measurement is committed."""
    if cfg != 5 or n not in (None, 120) or not os.path.exists(TRAFFIC_FILE):
        return None, None
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)
    if t.get("family", "gemm_update") != family:
        return None, None
    return t["bytes_per_launch"], os.path.relpath(TRAFFIC_FILE, ROOT)


TUNING_FILE = os.path.join(ROOT, "clique_amd", "tuning_mi355x.txt")


def use_tuning(args):
    """Point the library at a copy of the tile table (CLIQG_TUNE_CACHE, read
    at the first plan; a copy, so that a --retune pass does not rewrite the
    committed file). Returns the table's path relative to the repository, or
    None (cost-model shapes, or a table set in the environment)."""
    if args.no_tune or os.environ.get("CLIQG_TUNE_CACHE") or not os.path.exists(args.tuning):
        return None
    import shutil
    import tempfile
    import atexit
    fd, tmp = tempfile.mkstemp(prefix="cliqg_tune_", suffix=".txt")
    os.close(fd)
    shutil.copyfile(args.tuning, tmp)
    atexit.register(lambda: os.path.exists(tmp) and os.remove(tmp))
    os.environ["CLIQG_TUNE_CACHE"] = tmp
    return os.path.relpath(os.path.abspath(args.tuning), ROOT)


def write_tuning(path):
    """cliqg_plan_autotune over the FP32 and FP64 plans of configurations 1-5
    (configuration 5 last, so that the headline workload's buckets keep its
    own choices; its FP32 plan also at 32 right-hand sides, the width of the
    refinement's narrowed correction solves); the process-wide table is
    written to `path` after every pass."""
    os.environ["CLIQG_TUNE_CACHE"] = os.path.abspath(path)
    if os.path.exists(path):
        os.remove(path)
    import torch
    import clique_amd as cq
    for cfg in (1, 2, 3, 4, 5):
        ip, ind, v, dims, b, cplx, desc = workload(cfg)
        sym = symbolic_for(ip, ind, dims)
        vals = torch.tensor(v, device="cuda:0")
        for prec in ("f32", "f64"):
            p = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
            widths = (b.shape[1], 32) if (cfg == 5 and prec == "f32") else (b.shape[1],)
            for k in widths:
                nb, nc = p.autotune(vals, precision=prec, rhs=k)
                print(f"# config {cfg} {prec} rhs {k}: {nb} buckets, {nc} changed",
                      file=sys.stderr, flush=True)
            del p
            torch.cuda.empty_cache()
    print(json.dumps({"table": path, "buckets": len(cq.autotune_table())}))


def factor_only(args):
    """Two profiled FP32 factorizations (launches serialized on one stream, as
    in the profiled factorization of the benchmark line) and a 1 GiB device
    copy, nothing else: the program the kernel statistics and byte counters
    are collected on (the copy calibrates the counters)."""
    import torch
    import clique_amd as cq
    ip, ind, v, dims, b, cplx, desc = workload(args.config, args.n)
    sym = cq.NaturalNestedDissection(*dims, ip, ind, cutoff=128)
    vals = torch.tensor(v, device="cuda:0")
    p = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
    if args.retune and not args.no_tune:
        p.autotune(vals, precision="f32")
    p.set_profiling(True)
    for _ in range(2):
        p.factor(vals, precision="f32")
    prof = p.profile()
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    name = max((n_ for n_ in prof if n_.startswith("gemm_update")), key=lambda n_: prof[n_]["ms"])
    g = prof[name]
    print(json.dumps({"factor_ms": p.last_factor_ms(), "copy_bytes": src.numel() * 4,
                      "family": name, "launches": g["launches"],
                      "avg_ms": g["ms"] / g["launches"],
                      "tflops": g["flops"] / (g["ms"] * 1e-3) / 1e12,
                      "shapes": {kk: e["shape"] for kk, e in cq.autotune_table().items()}}))


def timed_solve(p, src, reps=3):
    """Median of `reps` solves of the right-hand sides src (original
    ordering, Pull and Push included) on the plan's stream, in ms."""
    import torch
    x = src.clone()
    ts = []
    for r in range(reps + 1):
        x.copy_(src)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        p.solve(x)
        e1.record()
        torch.cuda.synchronize()
        if r:
            ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def solve_work(sym, flops_per_rhs, N, k, cplx, fac_bytes, rhs_bytes, ms):
    """Algorithmic work of one solve with k right-hand sides: flops =
    2 SolveWork k + N k (x4 complex for the diagonal), bytes = 2 sum(m n)
    factor entries + 4 N k + 2 sum(u) k right-hand-side entries."""
    _, _, entries = sym.work(cplx)
    sum_u = sum(sym.node(s)["lower_size"] for s in range(sym.num_nodes))
    fl = 2.0 * flops_per_rhs * k + N * k * (4 if cplx else 1)
    by = 2.0 * entries * fac_bytes + 4.0 * N * k * rhs_bytes + 2.0 * sum_u * k * rhs_bytes
    return {"ms": ms, "rhs": k, "flops": fl, "bytes": by, "tflops": fl / (ms * 1e-3) / 1e12,
            "gbps": by / (ms * 1e-3) / 1e9, "hbm_frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=5, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--n", type=int, default=None, help="override the grid size")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="no tuning table and no tuning pass (cost-model tile shapes)")
    ap.add_argument("--tuning", default=TUNING_FILE, metavar="FILE",
                    help="GEMM tile table written by cliqg_plan_autotune (bench.py "
                         "--write-tuning) that the plans use; default: the committed MI355X table")
    ap.add_argument("--retune", action="store_true",
                    help="run the autotuner on both plans before the timed steps (starting "
                         "from the table)")
    ap.add_argument("--write-tuning", default=None, metavar="FILE",
                    help="autotune the FP32 and FP64 plans of every configuration and write "
                         "the table to FILE")
    ap.add_argument("--block", action="store_true",
                    help="add a BLOCK_LDL (selective inversion) FP32+IR line")
    ap.add_argument("--no-block", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--factor-only", action="store_true",
                    help="two FP32 factorizations and a calibration copy (for counter runs)")
    ap.add_argument("--cpu-baseline-only", type=str, default=None, metavar="K,FLOPS,SOLVE",
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_only:
        # the CPU baseline runs in a process of its own, without torch (whose
        # bundled MKL would stand in for the CPU BLAS's symbols)
        kk, ft, st = args.cpu_baseline_only.split(",")
        tgt = (float(ft), float(st)) if float(ft) > 0 else None
        print(json.dumps(cpu_baseline(args.cpu_budget, k=int(kk), target=tgt)))
        return
    if args.write_tuning:
        write_tuning(args.write_tuning)
        return
    tuning = use_tuning(args)
    if args.gpus != 1:
        print("bench.py measures one GPU only: the multifrontal factor and solves run on "
              "one MI355X here, and runs on several GPUs are not implemented", file=sys.stderr)
        sys.exit(2)
    if args.factor_only:
        factor_only(args)
        return

    import torch
    import clique_amd as cq

    t_setup = time.perf_counter()
    ip, ind, v, dims, b, cplx, desc = workload(args.config, args.n)
    N = ip.size - 1
    k = b.shape[1]
    sym = symbolic_for(ip, ind, dims)
    flops, solve_flops, entries = sym.work(cplx)
    dev = torch.device("cuda:0")
    vals = torch.tensor(v, device=dev)
    bd = cq.colmajor(torch.tensor(b, device=dev))
    x32 = cq.colmajor(torch.empty_like(bd))
    print(f"# {desc}: N={N} fronts={sym.num_nodes} factor {flops / 1e9:.1f} GFLOP, "
          f"setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    p32 = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
    tune = {"table": tuning} if tuning else {}
    if args.retune and not args.no_tune:
        # the project's tile autotuner, once per plan before the timed steps
        t_tune = time.perf_counter()
        nb, nc = p32.autotune(vals, precision="f32", rhs=k)
        tune["fp32"] = {"buckets": nb, "changed": nc, "s": time.perf_counter() - t_tune}

    def step32():
        p32.factor(vals, precision="f32")
        fms = p32.last_factor_ms()
        _, info = p32.refine(bd, x32, tol=args.tol, max_iters=30)
        return fms, info

    for _ in range(args.warmup):
        step32()
    torch.cuda.synchronize()
    f32_ms, infos = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fms, info = step32()
        f32_ms.append(fms)
        infos.append(info)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    # One more step of the same workload with HIP events around every launch
    # (the library's profiler), right after the timed steps: the events around
    # ~3,300 launches add about 7 % to a step, so they stay out of the timed
    # region; kernel durations are the same.
    p32.set_profiling(True)
    step32()
    prof32 = p32.profile()
    p32.set_profiling(False)
    print(f"# fp32+IR: {t32 * 1e3:.1f} ms/step, factor {np.mean(f32_ms):.1f} ms, "
          f"IR iterations {[i['iterations'] for i in infos]}, eta {infos[-1]['backward_error']:.2e}",
          file=sys.stderr, flush=True)
    solve32_ms = timed_solve(p32, bd)
    fac_bytes = 8 if cplx else 4
    sw32 = solve_work(sym, flops_per_rhs=solve_flops, N=N, k=k, cplx=cplx, fac_bytes=fac_bytes,
                      rhs_bytes=fac_bytes, ms=solve32_ms)
    print(f"# fp32 solve ({k} RHS): {solve32_ms:.2f} ms, {sw32['gbps']:.0f} GB/s, "
          f"{sw32['tflops']:.1f} TFLOP/s", file=sys.stderr, flush=True)

    mem32 = p32.memory()
    del p32
    torch.cuda.empty_cache()

    # BLOCK_LDL fronts (inverted diagonal blocks, GEMM-only solves), FP32 + IR
    blk = None
    if args.block:
        pb = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
        xb = cq.colmajor(torch.empty_like(bd))

        def stepb():
            pb.factor(vals, precision="f32", front_type="block_ldl")
            fms = pb.last_factor_ms()
            _, info = pb.refine(bd, xb, tol=args.tol, max_iters=30)
            return fms, info

        for _ in range(args.warmup):
            stepb()
        torch.cuda.synchronize()
        fb_ms, infob = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fms, info = stepb()
            fb_ms.append(fms)
            infob.append(info)
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / args.steps
        pb.set_profiling(True)
        xb.copy_(bd)
        pb.solve(xb, nodal=False)
        profbs = pb.profile()
        pb.set_profiling(False)
        blk = {"time_to_tol_ms": tb * 1e3, "factor_ms": float(np.mean(fb_ms)),
               "refine_iterations": infob[-1]["iterations"],
               "backward_error": infob[-1]["backward_error"],
               "solve_ms": sum(e["ms"] for e in profbs.values())}
        print(f"# block fp32+IR: {tb * 1e3:.1f} ms/step, factor {np.mean(fb_ms):.1f} ms, "
              f"solve {blk['solve_ms']:.1f} ms, IR iterations {[i['iterations'] for i in infob]}",
              file=sys.stderr, flush=True)
        del pb
        torch.cuda.empty_cache()

    fp64 = None
    if not args.no_fp64:
        p64 = cq.DistSymmFrontTree(sym, ip, ind, is_complex=cplx)
        x64 = cq.colmajor(torch.empty_like(bd))
        if args.retune and not args.no_tune:
            t_tune = time.perf_counter()
            nb, nc = p64.autotune(vals, precision="f64", rhs=k)
            tune["fp64"] = {"buckets": nb, "changed": nc, "s": time.perf_counter() - t_tune}

        def step64():
            p64.factor(vals, precision="f64")
            fms = p64.last_factor_ms()
            _, info = p64.refine(bd, x64, tol=args.tol, max_iters=5)
            return fms, info

        for _ in range(args.warmup):
            step64()
        torch.cuda.synchronize()
        f64_ms, infos64 = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fms, info = step64()
            f64_ms.append(fms)
            infos64.append(info)
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / args.steps
        solve64_ms = timed_solve(p64, bd)
        sw64 = solve_work(sym, flops_per_rhs=solve_flops, N=N, k=k, cplx=cplx,
                          fac_bytes=2 * fac_bytes, rhs_bytes=2 * fac_bytes, ms=solve64_ms)
        # forward error of the FP32+IR solution against the FP64 one, per column
        fwd = (torch.linalg.vector_norm(x32 - x64, dim=0) /
               torch.linalg.vector_norm(x64, dim=0)).cpu().numpy()
        p64.set_profiling(True)
        p64.factor(vals, precision="f64")
        prof64 = p64.profile()
        fp64 = {"time_to_tol_ms": t64 * 1e3, "factor_ms": float(np.mean(f64_ms)),
                "solve": sw64,
                "factor_tflops": flops / (np.mean(f64_ms) * 1e-3) / 1e12,
                "refine_iterations": infos64[-1]["iterations"],
                "backward_error": infos64[-1]["backward_error"],
                "gemm_update_tflops": sum(e["flops"] for n_, e in prof64.items()
                                          if n_.startswith("gemm_update"))
                / (sum(e["ms"] for n_, e in prof64.items() if n_.startswith("gemm_update")) * 1e-3)
                / 1e12}
        print(f"# fp64: {t64 * 1e3:.1f} ms/step, factor {np.mean(f64_ms):.1f} ms",
              file=sys.stderr, flush=True)
        del p64

    # the dominant kernel: the update family (one tile shape) with the most time
    upd_name = max((n_ for n_ in prof32 if n_.startswith("gemm_update")),
                   key=lambda n_: prof32[n_]["ms"])
    g = prof32[upd_name]
    achieved = g["flops"] / (g["ms"] * 1e-3) / 1e12
    kernel_of = {"gemm_update": "k_gemm<cfloat,64,64,3>", "gemm_update_w8": "k_gemm<cfloat,128,128,3,16,2,4>",
                 "gemm_update_small": "k_gemm<cfloat,32,32,3>", "gemm_update_large": "k_gemm<float,128,128,3>",
                 "gemm_update_k64": "k_gemm<cfloat,32,32,3,64>"}
    # complex FP32 products issue three real MFMAs per complex block (Gauss)
    # where the flop convention counts four real multiplications
    mfma_share = 0.75 if cplx else 1.0
    traffic, traffic_src = measured_traffic(args.config, args.n, upd_name)
    solve_names = ("solve_gemm", "solve_gemm_stream", "solve_extend_add", "solve_gather",
                   "solve_zero", "solve_reduce", "solve_panel")

    def families(prof, solve):
        # one step (factor + refinement)
        return {name: {"launches": e["launches"], "ms": round(e["ms"], 3),
                       "tflops": (e["flops"] / (e["ms"] * 1e-3) / 1e12) if e["flops"] else None,
                       "gbps": (e["bytes"] / (e["ms"] * 1e-3) / 1e9) if e["bytes"] else None}
                for name, e in prof.items() if (name in solve_names) == solve}
    breakdown = families(prof32, False)
    fac32 = float(np.mean(f32_ms))
    out = {
        "metric": METRIC,
        "value": flops / (fac32 * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "c32" if cplx else "f32",
        "data": "synthetic",
        "config": {"workload": desc, "N": N, "fronts": sym.num_nodes, "rhs": k,
                   "factor_gflop": flops / 1e9, "factor_entries": entries,
                   "parallelism": "single GPU"},
        "fp32_ir": {"time_to_tol_ms": t32 * 1e3, "factor_ms": fac32,
                    "refine_iterations": infos[-1]["iterations"],
                    "backward_error": infos[-1]["backward_error"], "tol": args.tol,
                    "device_bytes": mem32, "solve": sw32,
                    "forward_error_vs_fp64": ({"max": float(fwd.max()),
                                               "median": float(np.median(fwd))}
                                              if fp64 else None)},
        "fp64": fp64,
        "speedup_fp32ir_vs_fp64": (fp64["time_to_tol_ms"] / (t32 * 1e3)) if fp64 else None,
        "roofline": {"bound": "mfma",
                     "kernel": kernel_of.get(upd_name, upd_name) +
                     " (Schur, panel and outer-panel updates, FP32 factor)",
                     "family": upd_name,
                     "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK["fp32"],
                     "mfma_issue_frac": achieved * mfma_share / PEAK["fp32"],
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "launches": g["launches"], "avg_launch_ms": g["ms"] / g["launches"],
                     "flops_per_launch": g["flops"] / g["launches"]},
        "breakdown_fp32_factor": breakdown,
        "breakdown_fp32_solves": families(prof32, True),
        "block_ldl_fp32_ir": blk,
        "autotune": tune or None,
        "autotune_shapes": {kk: e["shape"] for kk, e in cq.autotune_table().items()},
    }
    if not args.no_cpu:
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-budget",
                            str(args.cpu_budget), "--cpu-baseline-only",
                            f"{k},{flops if cplx else 0},{solve_flops}"],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline failed: " + r.stderr[-2000:])
        cb = json.loads(r.stdout.strip().splitlines()[-1])
        out["cpu_baseline"] = cb
        if "extrapolated_to_workload" in cb:
            out["speedup_fp32ir_vs_cpu_extrapolated"] = (
                cb["extrapolated_to_workload"]["time_to_solution_s"] / t32)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
