"""Benchmark of the MI355X P1 assembly + PCG path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu-baseline]

Workload (one GPU): 3D Poisson P1 on the unit cube, 128^3 hexes split into 6
Kuhn tets each (12,582,912 tets, 2,146,689 nodes), interior nodes jittered by
+-0.15h (seed 0), f = 1, u = 0 on z = 0 (row elimination), CSR (~14.8 nnz/row).

A step is one assembly pass over that mesh with the inputs resident in HBM:
the FP64 matrix values of the precomputed pattern plus the constant-source
RHS, in one kernel.  value = assembled DoFs per second = rows / step time.
The pattern build, FP64 SpMV and the two solvers (FP64 Jacobi-PCG and the
mixed-precision PCG, both to 1e-8 FP64 true relative residual) are measured
in the same run and reported beside it.

One JSON line is printed on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "assembled DoFs/s; solve time to 1e-8 relres, learned-PCG vs plain CG"
HBM_PEAK_GBS = 8000.0

# assembly variant id (afem_bsr_set_assembly_variant) -> kernel instantiation
ASSEMBLY_KERNELS = {
    0: "k_assemble_laplace<4, 4>", 1: "k_assemble_laplace_plan<4, 4>", 2: "k_assemble_laplace_plan<4, 8>",
    3: "k_assemble_laplace_cplan<4, 4>", 4: "k_assemble_laplace_cplan<4, 8>",
    5: "k_assemble_laplace_lplan<4, 4>", 6: "k_assemble_laplace_lplan<4, 8>", 7: "k_assemble_laplace_lplan<4, 16>",
}
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
# workload (scaled by a copy of known size), written by arcanefem_amd.profiling
PMC_SUMMARY = os.path.join(ROOT, "profiles", "assembly_spmv_tet128.pmc.json")


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed PMC summary, or None."""
    try:
        with open(PMC_SUMMARY) as f:
            summ = json.load(f)
    except (OSError, ValueError):
        return None
    if summ.get("workload") != workload:
        return None
    k = summ.get("kernels", {}).get(kernel)
    return k["traffic_bytes"] if k else None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_avg_ms(A, name):
    n, t = A.profile_query(name)
    return (t / n if n else float("nan")), n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="hexes per side")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=9, help="timed repetitions of the CPU baseline")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one 256 MiB device-to-device copy first (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design only, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import arcanefem_amd as A

    A.init(0)
    if args.calibrate:
        nbytes = 256 << 20
        src, dst = A.DeviceArray(nbytes, np.uint8), A.DeviceArray(nbytes, np.uint8)
        src.zero()
        A.check(A.lib().afem_memcpy_dtod(dst.c, src.c, nbytes))
        A.synchronize()
        log(f"[bench] calibration copy of {nbytes} bytes done")
        del src, dst
    n = args.n
    t0 = time.time()
    X, C = A.box_mesh(3, n, 0.15, seed=args.seed)
    N, T = X.shape[0], C.shape[0]
    log(f"[bench] mesh {n}^3: N={N} T={T} ({time.time() - t0:.1f}s)")
    mesh = A.Mesh(3, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    A.synchronize()
    t0 = time.perf_counter()
    bsr.compute_sparsity()
    A.synchronize()
    pattern_ms = (time.perf_counter() - t0) * 1e3
    info = bsr.info()
    nnz = int(info.nb_col)
    log(f"[bench] pattern nnz={nnz} ({nnz / N:.2f}/row) in {pattern_ms:.1f} ms")
    sysm = A.DoFLinearSystem(N)
    rhs = sysm.rhs()
    variant, tune_ms = bsr.autotune_assembly()
    log(f"[bench] assembly autotuner: variant {variant} ({tune_ms:.3f} ms)")

    # ---- timed assembly steps -------------------------------------------------------------
    for _ in range(args.warmup):
        bsr.assemble_laplace(None, 1.0, rhs)
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    A.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bsr.assemble_laplace(None, 1.0, rhs)
    A.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    A.profile_enable(False)
    ms_per_step = wall_ms / args.steps
    kname = "assemble_laplace_tet"
    k_ms, k_n = kernel_avg_ms(A, kname)
    alg_bytes = 16 * T + 24 * N + 12 * nnz + 4 * (N + 1)
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    value = N / (ms_per_step * 1e-3)
    log(f"[bench] assembly: {ms_per_step:.3f} ms/step (kernel {k_ms:.3f} ms x{k_n}), {value / 1e9:.2f} GDoF/s, "
        f"{achieved:.0f} GB/s algorithmic")

    workload_key = f"tet{n}_seed{args.seed}"
    akernel = ASSEMBLY_KERNELS.get(variant)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Poisson P1 assembly (CSR values + constant-source RHS), {n}^3 hexes x 6 Kuhn tets, "
                        f"jitter 0.15h seed {args.seed}",
            "nodes": N, "tets": T, "nnz": nnz, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload_key, akernel),
            "kernel": akernel, "kernel_ms": k_ms, "algorithmic_bytes": alg_bytes,
            "bytes_formula": "16*T + 24*N + 12*nnz + 4*(N+1)", "assembly_variant": variant,
            "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT),
        },
        "pattern_ms": pattern_ms,
    }

    # ---- SpMV and solves -------------------------------------------------------------------------
    bsr.assemble_laplace(None, 1.0, rhs)
    sysm.set_bsr(bsr)
    bottom = np.nonzero(X[:, 2] == 0.0)[0]
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bottom, 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    sysm.apply_rhs_transformation()
    x = A.DeviceArray.from_numpy(np.random.default_rng(1).normal(size=N))
    y = A.DeviceArray(N, np.float64)
    x32 = A.DeviceArray.from_numpy(np.random.default_rng(1).normal(size=N).astype(np.float32))
    y32 = A.DeviceArray(N, np.float32)
    for _ in range(3):
        sysm.spmv(x, y)
        sysm.spmv_fp32(x32, y32)
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    for _ in range(20):
        sysm.spmv(x, y)
        sysm.spmv_fp32(x32, y32)
    A.synchronize()
    A.profile_enable(False)
    s64, _ = kernel_avg_ms(A, "spmv_fp64")
    s32, _ = kernel_avg_ms(A, "spmv_fp32")
    b64 = 12 * nnz + 20 * N
    b32 = 8 * nnz + 12 * N
    fmt = sysm.spmv_format()
    k64 = "k_spmv_sell<double, double, 0>" if fmt[0] == "sell" else f"k_spmv_csr<{fmt[1]}, double, double, 0>"
    k32 = "k_spmv_sell<float, float, 0>" if fmt[0] == "sell" else f"k_spmv_csr<{fmt[1]}, float, float, 0>"
    out["spmv"] = {
        "format": fmt[0] if fmt[0] == "sell" else f"csr{fmt[1]}",
        "fp64": {"kernel_ms": s64, "algorithmic_bytes": b64, "GBs": b64 / (s64 * 1e-3) / 1e9,
                 "frac": b64 / (s64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": k64,
                 "traffic": measured_traffic(workload_key, k64)},
        "fp32": {"kernel_ms": s32, "algorithmic_bytes": b32, "GBs": b32 / (s32 * 1e-3) / 1e9,
                 "frac": b32 / (s32 * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": k32,
                 "traffic": measured_traffic(workload_key, k32)},
    }
    log(f"[bench] spmv fp64 {s64:.3f} ms ({out['spmv']['fp64']['GBs']:.0f} GB/s), "
        f"fp32 {s32:.3f} ms ({out['spmv']['fp32']['GBs']:.0f} GB/s)")

    if not args.no_solve:
        solves = {}
        for name, method in (("jacobi_pcg_fp64", A.SOLVER_PCG_FP64), ("mixed_pcg", A.SOLVER_PCG_MIXED)):
            best = None
            for rep in range(2):
                r = sysm.solve(raise_on_failure=False, method=method, rtol=1e-8, max_iter=20000, check_every=32)
                if best is None or r.solve_ms < best.solve_ms:
                    best = r
            solves[name] = {"iterations": best.iterations, "solve_ms": best.solve_ms, "relres_fp64": best.relres,
                            "converged": bool(best.converged), "reliable_updates": best.reliable_updates}
            log(f"[bench] {name}: {best.iterations} it, {best.solve_ms:.1f} ms, relres {best.relres:.2e}")
        solves["speedup_mixed_vs_fp64"] = solves["jacobi_pcg_fp64"]["solve_ms"] / solves["mixed_pcg"]["solve_ms"]
        solves["tolerance"] = 1e-8
        out["solve"] = solves

    # ---- CPU baseline: the oracle's assembly of the same mesh, one core --------------------------
    if not args.no_cpu_baseline:
        import oracle as O
        row, col, _, _ = bsr.download()
        times = []
        O.assemble(3, X, C, 4, row, col)  # warm-up (page-in)
        for _ in range(args.cpu_reps):
            t0 = O.wtime()
            O.assemble(3, X, C, 4, row, col)
            O.rhs_constant_source(3, X, C, 4, 1.0)
            times.append(O.wtime() - t0)
        t = float(np.median(times))
        out["cpu_baseline"] = {
            "value": N / t, "unit": "DoF/s", "cores": 1, "kind": "port",
            "sample": f"full {n}^3 assembly (cell-wise, matrix values with per-entry row search + RHS), "
                      f"1 warm-up + median of {args.cpu_reps} ({t:.2f} s each, {sum(times):.1f} s total), "
                      f"one thread on {cpu_model()}",
        }
        log(f"[bench] cpu baseline: {N / t / 1e6:.2f} MDoF/s ({t:.2f} s)")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
