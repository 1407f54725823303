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
    0: "k_assemble_laplace<4, 4>", 1: "k_assemble_laplace_plan<4, 8>",
    2: "k_assemble_laplace_lplan<4, 8>", 3: "k_assemble_laplace_lplan<4, 16>",
    4: "k_assemble_laplace_brick<4, 8>", 5: "k_assemble_laplace_brick<4, 16>",
    6: "k_assemble_laplace_cell<4>",
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


def cpu_solve_baseline(A, args, bsr, sysm, N):
    """FP64 Jacobi-PCG of the same transformed system on the host (the
    PETSc cg + jacobi defaults, femutils/PetscDoFLinearSystemFactory.axl:5-10,
    restated by the oracle): on all host threads, 1 warm-up + median of
    --cpu-solve-reps full solves to 1e-8; on one thread a bounded sample of
    40 iterations (median of 5), extrapolated to the all-cores iteration
    count."""
    import oracle as O
    row, col, val, _ = bsr.download()
    b = sysm.rhs().numpy()
    nth = O.max_threads()
    O.pcg_jacobi_threads(row, col, val, b, nth, rtol=1e-8, max_iter=20000)
    times, its, rr = [], None, None
    for _ in range(args.cpu_solve_reps):
        t0 = O.wtime()
        _, its, rr = O.pcg_jacobi_threads(row, col, val, b, nth, rtol=1e-8, max_iter=20000)
        times.append(O.wtime() - t0)
    t_all = float(np.median(times))
    sample_it = 40
    t1 = []
    for _ in range(5):
        t0 = O.wtime()
        O.pcg_jacobi_threads(row, col, val, b, 1, rtol=1e-30, max_iter=sample_it)
        t1.append(O.wtime() - t0)
    per_it_1 = float(np.median(t1)) / sample_it
    out = {"kind": "port", "cpu": cpu_model(),
           "all_cores": {"cores": nth, "solve_s": t_all, "iterations": int(its), "relres_fp64": rr,
                         "reps": args.cpu_solve_reps},
           "one_core": {"cores": 1, "ms_per_iteration": per_it_1 * 1e3,
                        "solve_s_extrapolated": per_it_1 * max(int(its), 1),
                        "sample": f"{sample_it} iterations, median of 5"}}
    log(f"[bench] cpu jacobi-pcg fp64: {t_all:.2f} s on {nth} threads ({its} it), "
        f"{per_it_1 * 1e3:.1f} ms/iteration on one thread")
    return out


def measure_copy(A, nbytes=1 << 30, reps=10):
    """Device-to-device copy by the library's streaming kernel (event-timed):
    the measured HBM bandwidth the roofline figures are set beside."""
    src, dst = A.DeviceArray(nbytes, np.uint8), A.DeviceArray(nbytes, np.uint8)
    src.zero()
    for _ in range(3):
        A.check(A.lib().afem_stream_copy(dst.c, src.c, nbytes))
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    for _ in range(reps):
        A.check(A.lib().afem_stream_copy(dst.c, src.c, nbytes))
    A.synchronize()
    A.profile_enable(False)
    ms, n = kernel_avg_ms(A, "stream_copy")
    del src, dst
    return {"bytes_moved": 2 * nbytes, "kernel_ms": ms, "GBs": 2 * nbytes / (ms * 1e-3) / 1e9, "reps": n,
            "kernel": "k_stream_copy"}


def rel_diff(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def solve_best(A, sysm, reps=2, x_fp64=None, **kw):
    """Best of `reps` solves; with x_fp64 also ||x - x_fp64|| / ||x_fp64||
    (the error against the FP64 Jacobi-PCG solution of the same system)."""
    best = None
    for _ in range(reps):
        r = sysm.solve(raise_on_failure=False, **kw)
        if best is None or r.solve_ms < best.solve_ms:
            best = r
    out = {"iterations": best.iterations, "solve_ms": best.solve_ms, "relres_fp64": best.relres,
           "converged": bool(best.converged), "reliable_updates": best.reliable_updates,
           "fallback": bool(best.fallback)}
    if x_fp64 is not None:
        out["error_vs_fp64_solution"] = rel_diff(sysm.solution().numpy(), x_fp64)
    return out


def learned_section(A, args, X, C, mesh, bsr, N, nnz):
    """Heterogeneous 3D Poisson: lambda = exp(g), g a Gaussian random field
    (squared-exponential covariance, sigma 1, correlation length 0.1) at the
    cell centroids.  The learned preconditioner is trained on the GPU on
    systems of a smaller mesh generated by the FP64 assembly (one Adam step per
    system, 8 Rademacher probes), then every solver runs on a held-out field on
    the full mesh to 1e-8 FP64 true relative residual."""
    sigma, ell, nfeat = 1.0, 0.1, 512
    # training systems
    nt = args.train_n
    Xt, Ct = A.box_mesh(3, nt, 0.15, seed=args.seed + 1)
    mt = A.Mesh(3, Xt, Ct)
    bt = A.BSRFormat(mt, 1, use_csr=True)
    bt.compute_sparsity()
    st = A.DoFLinearSystem(Xt.shape[0])
    lam_t = A.DeviceArray(Ct.shape[0], np.float64)
    bottom_t = np.nonzero(Xt[:, 2] == 0.0)[0].astype(np.int32)
    st.set_bsr(bt)
    st.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bottom_t, 1, 0, 0.0, 0.0)
    net = A.LearnedSpai(seed=args.seed)

    def make_training_system(k):
        A.lognormal_field(mt, sigma, ell, nfeat, 10_000 + k, out=lam_t)
        bt.assemble_laplace(lam_t, 1.0, st.rhs())
        st.set_bsr(bt)
        st.apply_matrix_transformation()

    make_training_system(0)
    loss_before = net.loss(st, probe_seed=0)
    A.synchronize()
    t0 = time.perf_counter()
    losses = []
    steps = 0
    for epoch in range(args.train_epochs):
        for k in range(args.train_systems):
            make_training_system(k)
            losses.append(net.train_step(st, probe_seed=epoch * args.train_systems + k, lr=args.lr))
            steps += 1
    A.synchronize()
    train_s = time.perf_counter() - t0
    log(f"[bench] learned SPAI: {steps} steps on {nt}^3 systems in {train_s:.1f} s, "
        f"loss {loss_before:.4f} -> {np.mean(losses[-32:]):.4f}")
    del st, bt, mt

    # held-out field on the full mesh
    lam = A.lognormal_field(mesh, sigma, ell, nfeat, 424242)
    sysm = A.DoFLinearSystem(N)
    bsr.assemble_laplace(lam, 1.0, sysm.rhs())
    sysm.set_bsr(bsr)
    bottom = np.nonzero(X[:, 2] == 0.0)[0].astype(np.int32)
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bottom, 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    sysm.apply_rhs_transformation()
    lamv = lam.numpy()
    res = {
        "workload": f"3D Poisson, {args.n}^3 hexes x 6 tets, lambda = exp(g), g GRF sigma {sigma} "
                    f"corr length {ell} ({nfeat} random Fourier features), f = 1, u = 0 on z = 0",
        "lambda_range": [float(lamv.min()), float(lamv.max())],
        "training": {"mesh": f"{nt}^3", "systems": args.train_systems, "epochs": args.train_epochs,
                     "steps": steps, "probes": 8, "lr": args.lr, "time_s": train_s,
                     "loss_untrained": loss_before, "loss_last32": float(np.mean(losses[-32:]))},
        "tolerance": 1e-8,
    }
    kw = dict(rtol=1e-8, max_iter=40000, check_every=32)
    res["jacobi_pcg_fp64"] = solve_best(A, sysm, method=A.SOLVER_PCG_FP64, **kw)
    x64 = sysm.solution().numpy()
    res["cg_fp64"] = solve_best(A, sysm, x_fp64=x64, method=A.SOLVER_CG_FP64, **kw)
    res["jacobi_pcg_mixed"] = solve_best(A, sysm, x_fp64=x64, method=A.SOLVER_PCG_MIXED, **kw)
    # inference: the network's forward pass on the system and the installation
    # of M (its SELL copy and BF16 words, made by a zero-iteration solve).  The
    # first application at this mesh size also allocates the activations;
    # the second, on the same buffers, is the per-system cost.
    for key in ("learned_inference_first_ms", "learned_inference_ms"):
        A.synchronize()
        t0 = time.perf_counter()
        net.apply(sysm)
        sysm.solve(raise_on_failure=False, method=A.SOLVER_FPCG_SPAI, precond_bf16=1, max_iter=0)
        A.synchronize()
        res[key] = (time.perf_counter() - t0) * 1e3
    res["learned_fpcg_mixed_fp32m"] = solve_best(A, sysm, x_fp64=x64, method=A.SOLVER_FPCG_SPAI, **kw)
    res["learned_fpcg_mixed"] = solve_best(A, sysm, x_fp64=x64, method=A.SOLVER_FPCG_SPAI, precond_bf16=1, **kw)
    res["learned_fpcg_mixed"]["preconditioner_values"] = "bf16"
    lt = res["learned_fpcg_mixed"]["solve_ms"]
    res["speedup_vs_jacobi_pcg_fp64"] = res["jacobi_pcg_fp64"]["solve_ms"] / lt
    res["speedup_vs_cg_fp64"] = res["cg_fp64"]["solve_ms"] / lt
    res["speedup_vs_jacobi_pcg_fp64_incl_inference"] = (res["jacobi_pcg_fp64"]["solve_ms"]
                                                        / (lt + res["learned_inference_ms"]))
    for name in ("cg_fp64", "jacobi_pcg_fp64", "jacobi_pcg_mixed", "learned_fpcg_mixed_fp32m", "learned_fpcg_mixed"):
        r = res[name]
        err = r.get("error_vs_fp64_solution")
        log(f"[bench] heterogeneous {name}: {r['iterations']} it, {r['solve_ms']:.1f} ms, "
            f"relres {r['relres_fp64']:.2e}{' (fallback)' if r['fallback'] else ''}"
            + (f", |x - x_fp64|/|x_fp64| {err:.1e}" if err is not None else ""))
    return res


def batch_section(A, args):
    """64 independent 2D Poisson/heat systems: unit squares with n x n quads
    split into 2 triangles, n log-uniform in [64, 512] (seeded), interior
    nodes jittered by 0.2h, conductivity piecewise constant on an 8x8
    checkerboard with values log-uniform in [0.1, 10], f = 1, u = 0 on the
    boundary.  All systems are assembled in one launch on the disjoint union
    of the meshes and solved together to 1e-8 FP64 true relative residual
    (FP32-SpMV PCG with FP64 reliable updates) by the default chunked solve
    (hipGraph chunks of ordinary kernels); the one-persistent-launch solve is
    timed beside it."""
    rng = np.random.default_rng(args.seed)
    ns = np.rint(np.exp(rng.uniform(np.log(64), np.log(512), args.batch_systems))).astype(int)
    parts, lams, bnds = [], [], []
    for k, n in enumerate(ns):
        X, C = A.box_mesh(2, int(n), 0.2, seed=args.seed + 100 + k)
        vals = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 64))
        cen = X[C[:, :3]].mean(axis=1)
        ix = np.minimum((cen[:, 0] * 8).astype(int), 7)
        iy = np.minimum((cen[:, 1] * 8).astype(int), 7)
        parts.append((X, C))
        lams.append(vals[iy * 8 + ix])
        on = (X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0)
        bnds.append(np.nonzero(on)[0])
    X, C, offs = A.union_mesh(parts)
    nodes = np.concatenate([b + offs[k] for k, b in enumerate(bnds)]).astype(np.int32)
    N, T = X.shape[0], C.shape[0]
    mesh = A.Mesh(2, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    bsr.compute_sparsity()
    lam = A.DeviceArray.from_numpy(np.concatenate(lams))
    sysm = A.DoFLinearSystem(N)
    sysm.set_bsr(bsr)
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    bs = A.BatchSolver(sysm, offs)

    def assemble():
        bsr.assemble_laplace(lam, 1.0, sysm.rhs())
        sysm.set_bsr(bsr)
        sysm.apply_matrix_transformation()
        sysm.apply_rhs_transformation()

    variant, _ = bsr.autotune_assembly()  # times the variants with lambda = 1: assemble afterwards
    assemble()
    lanes, tiles, it_ms = bs.autotune()
    warm, _ = bs.solve(raise_on_failure=False, rtol=1e-8, max_iter=args.batch_max_iter)  # graph instantiation
    for k, r in enumerate(warm):
        if not r.converged:
            log(f"[bench] batch: system {k} (n={ns[k]}) not converged: {r.iterations} it, relres {r.relres:.2e}")
    best = None
    for rep in range(2):
        A.synchronize()
        t0 = time.perf_counter()
        assemble()
        A.synchronize()
        t1 = time.perf_counter()
        res, solve_ms = bs.solve(raise_on_failure=False, rtol=1e-8, max_iter=args.batch_max_iter)
        t2 = time.perf_counter()
        cur = ((t1 - t0) * 1e3, (t2 - t1) * 1e3, res)
        if best is None or cur[0] + cur[1] < best[0] + best[1]:
            best = cur
    asm_ms, sol_ms, res = best
    # the one-persistent-launch solve on the same systems: XCD-local teams
    # (persistent=1, where the device allows them) and the coherent exchange
    pers = {}
    for mode in (1, 2):
        bs.solve(raise_on_failure=False, rtol=1e-8, max_iter=args.batch_max_iter, persistent=mode)
        best_p = None
        for rep in range(2):
            assemble()
            A.synchronize()
            t1 = time.perf_counter()
            pres, _ = bs.solve(raise_on_failure=False, rtol=1e-8, max_iter=args.batch_max_iter, persistent=mode)
            ms = (time.perf_counter() - t1) * 1e3
            if best_p is None or ms < best_p[0]:
                best_p = (ms, pres)
        pers[mode] = (best_p[0], best_p[1], bs.persistent_mode())
    pers_ms, pers_res, pers_mode = pers[1]
    its = [r.iterations for r in res]
    log("[bench] batch iterations by n: " + ", ".join(f"{n}:{i}" for n, i in sorted(zip(ns.tolist(), its))))
    out = {
        "workload": f"{args.batch_systems} 2D Poisson systems, n log-uniform in [64, 512] (seed {args.seed}), "
                    "8x8 checkerboard conductivity log-uniform in [0.1, 10], f = 1, u = 0 on the boundary",
        "systems": int(args.batch_systems), "nodes_total": int(N), "triangles_total": int(T),
        "n_min": int(ns.min()), "n_max": int(ns.max()),
        "assembly_ms": asm_ms, "assembly_variant": variant,
        "solve_ms": sol_ms, "total_ms": asm_ms + sol_ms,
        "iterations_max": int(max(its)), "iterations_mean": float(np.mean(its)),
        "converged": int(sum(r.converged for r in res)),
        "relres_max": float(max(r.relres for r in res)),
        "tuned": {"lanes": lanes, "tiles": tiles, "ms_per_iteration": it_ms},
        "dofs_per_s": N / ((asm_ms + sol_ms) * 1e-3),
        "solver": "chunks of 32 iterations of ordinary kernels, each chunk a replayed hipGraph",
        "persistent": {"solve_ms": pers_ms, "iterations_max": int(max(r.iterations for r in pers_res)),
                       "converged": int(sum(r.converged for r in pers_res)),
                       "relres_max": float(max(r.relres for r in pers_res)),
                       "exchange": {1: "XCD-local teams", 2: "coherence point"}.get(pers_mode, "none"),
                       "coherent_solve_ms": pers[2][0],
                       "what": "one persistent launch in a hipGraph: per-system workgroup teams, bounded "
                               "barrier waits in device memory"},
    }
    log(f"[bench] batch: {args.batch_systems} systems, {N} DoFs, assembly {asm_ms:.2f} ms, solve {sol_ms:.1f} ms "
        f"(chunked graphs; one persistent launch {pers_ms:.1f} ms in exchange mode {pers_mode}, "
        f"coherent {pers[2][0]:.1f} ms), "
        f"iterations max {max(its)} mean {np.mean(its):.0f}, converged {out['converged']}")
    # the same systems one at a time with the single-system mixed PCG
    if not args.no_sequential:
        A.synchronize()
        t0 = time.perf_counter()
        seq_its, seq_solve = [], 0.0
        for k, (Xk, Ck) in enumerate(parts):
            mk = A.Mesh(2, Xk, Ck)
            bk = A.BSRFormat(mk, 1, use_csr=True)
            bk.compute_sparsity()
            sk = A.DoFLinearSystem(Xk.shape[0])
            bk.assemble_laplace(A.DeviceArray.from_numpy(lams[k]), 1.0, sk.rhs())
            sk.set_bsr(bk)
            sk.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bnds[k].astype(np.int32), 1, 0, 0.0, 0.0)
            sk.apply_matrix_transformation()
            sk.apply_rhs_transformation()
            r = sk.solve(raise_on_failure=False, method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=args.batch_max_iter)
            seq_its.append(r.iterations)
            seq_solve += r.solve_ms
        A.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3
        out["sequential_solve_ms"] = seq_solve
        out["sequential_ms_incl_setup"] = seq_ms
        out["speedup_solve_vs_sequential"] = seq_solve / sol_ms
        log(f"[bench] batch: one system at a time: solves {seq_solve:.0f} ms "
            f"({seq_ms:.0f} ms incl. per-mesh setup)")
    if args.heat_steps > 0:
        out["heat"] = batch_heat(A, args, mesh, bsr, sysm, bs, lam, nodes)
    return out


def batch_heat(A, args, mesh, bsr, sysm, bs, lam, nodes):
    """The same 64 meshes and conductivities as heat problems: backward Euler
    (lambda K + M/dt assembled once on the union mesh), f = 1, T = 0 at t = 0
    and on the boundary, every time step one warm-started batched solve of
    all systems to 1e-8."""
    dt, steps = args.heat_dt, args.heat_steps
    N = sysm.rhs().shape[0]
    A.synchronize()
    t0 = time.perf_counter()
    bsr.assemble_heat(lam, dt)
    sysm.set_bsr(bsr)
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, nodes, 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    source = A.DeviceArray(N, np.float64)
    A.check(A.lib().afem_rhs_constant_source(mesh.h, 1.0, None, 1, 0, source.c))
    sysm.solution().zero()
    A.synchronize()
    t1 = time.perf_counter()
    res = A.heat_batch_steps(mesh, sysm, bs, source, dt, steps, raise_on_failure=False, rtol=1e-8,
                             max_iter=args.batch_max_iter)
    A.synchronize()
    t2 = time.perf_counter()
    its = [[r.iterations for r in rs] for rs, _ in res]
    conv = all(r.converged for rs, _ in res for r in rs)
    loop_ms = (t2 - t1) * 1e3
    out = {"dt": dt, "steps": steps, "setup_ms": (t1 - t0) * 1e3, "loop_ms": loop_ms,
           "ms_per_step": loop_ms / steps, "solve_ms_per_step": float(np.mean([ms for _, ms in res])),
           "iterations_max_per_step": [int(max(i)) for i in its],
           "iterations_mean": float(np.mean(its)), "converged": bool(conv),
           "dof_steps_per_s": N * steps / (loop_ms * 1e-3)}
    log(f"[bench] batch heat: {steps} steps of dt {dt}: {loop_ms:.0f} ms ({loop_ms / steps:.1f} ms/step), "
        f"max iterations per step {out['iterations_max_per_step']}, converged {conv}")
    return out


def poisson2d_section(A, args):
    """2D Poisson P1: unit square, 256 x 256 quads split into 2 triangles,
    interior nodes jittered by 0.2h, f = 1, u = 0 on the whole boundary, CSR,
    FP64 Jacobi-PCG to 1e-10; the oracle's CPU assembly and Jacobi-PCG of the
    same system timed beside it (one thread)."""
    import oracle as O
    n = 256
    X, C = A.box_mesh(2, n, 0.2, seed=args.seed)
    N, T = X.shape[0], C.shape[0]
    mesh = A.Mesh(2, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    bsr.compute_sparsity()
    nnz = int(bsr.info().nb_col)
    sysm = A.DoFLinearSystem(N)
    bsr.autotune_assembly()
    for _ in range(3):
        bsr.assemble_laplace(None, 1.0, sysm.rhs())
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(20):
        bsr.assemble_laplace(None, 1.0, sysm.rhs())
    A.synchronize()
    asm_ms = (time.perf_counter() - t0) * 1e3 / 20
    A.profile_enable(False)
    k_ms, _ = kernel_avg_ms(A, "assemble_laplace_tri")
    sysm.set_bsr(bsr)
    on = np.nonzero((X[:, 0] == 0.0) | (X[:, 0] == 1.0) | (X[:, 1] == 0.0) | (X[:, 1] == 1.0))[0]
    sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, on.astype(np.int32), 1, 0, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    sysm.apply_rhs_transformation()
    sol = solve_best(A, sysm, method=A.SOLVER_PCG_FP64, rtol=1e-10, max_iter=20000, check_every=32)
    out = {
        "workload": "2D Poisson P1, 256^2 quads x 2 triangles, jitter 0.2h, f = 1, u = 0 on the boundary",
        "nodes": N, "triangles": T, "nnz": nnz,
        "assembly_ms": asm_ms, "assembly_kernel_ms": k_ms,
        "assembled_dofs_per_s": N / (asm_ms * 1e-3),
        "jacobi_pcg_fp64": sol,
    }
    if not args.no_cpu_baseline:
        row, col, _, _ = bsr.download()
        t0 = O.wtime()
        vals = O.assemble(2, X, C, 3, row, col)
        rhs = O.rhs_constant_source(2, X, C, 3, 1.0)
        t_asm = O.wtime() - t0
        elim = np.zeros(N, np.uint8)
        elim[on] = O.ELIMINATE_ROW
        O.apply_transformations(row, col, vals, rhs, elim_info=elim, elim_value=np.zeros(N))
        x0 = np.zeros(N)
        t0 = O.wtime()
        _, its, _ = O.pcg_jacobi(row, col, vals, rhs, rtol=1e-10, x0=x0)
        t_sol = O.wtime() - t0
        out["cpu"] = {"assembly_s": t_asm, "solve_s": t_sol, "iterations": int(its), "cores": 1, "kind": "port"}
        out["speedup_vs_cpu"] = (t_asm + t_sol) * 1e3 / (asm_ms + sol["solve_ms"])
    log(f"[bench] 2D 256^2: assembly {asm_ms:.3f} ms, PCG {sol['iterations']} it {sol['solve_ms']:.1f} ms"
        + (f", CPU {out['cpu']['assembly_s'] * 1e3:.0f} + {out['cpu']['solve_s'] * 1e3:.0f} ms"
           if "cpu" in out else ""))
    return out


def elasticity_section(A, args, assembly_only=False):
    """3D linear elasticity P1: 96^3 hexes x 6 tets (jitter 0.15h), E = 21e5,
    nu = 0.28, body force (0, 0, -1), face x = 0 clamped; BSR with 3x3 FP64
    blocks; mixed-precision block-Jacobi PCG to 1e-8.  assembly_only: the
    timed assembly steps alone (the entry point of counter runs)."""
    n = 96
    X, C = A.box_mesh(3, n, 0.15, seed=args.seed)
    N, T = X.shape[0], C.shape[0]
    mesh = A.Mesh(3, X, C)
    bsr = A.BSRFormat(mesh, 3, use_csr=False)
    bsr.compute_sparsity()
    nb = int(bsr.info().nb_col)
    E, nu = 21e5, 0.28
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    for _ in range(2):
        bsr.assemble_elasticity(lam, mu)
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(10):
        bsr.assemble_elasticity(lam, mu)
    A.synchronize()
    asm_ms = (time.perf_counter() - t0) * 1e3 / 10
    A.profile_enable(False)
    k_ms, _ = kernel_avg_ms(A, "assemble_elasticity_tet")
    alg = 16 * T + 24 * N + 76 * nb + 4 * (N + 1)
    if assembly_only:
        return {"workload": "3D linear elasticity P1 assembly, 96^3 hexes x 6 tets", "assembly_ms": asm_ms,
                "assembly_kernel_ms": k_ms, "algorithmic_bytes": alg,
                "assembly_GBs": alg / (k_ms * 1e-3) / 1e9 if k_ms == k_ms else None}
    sysm = A.DoFLinearSystem(3 * N)
    rhs = sysm.rhs()
    rhs.zero()
    A.check(A.lib().afem_rhs_body_force(mesh.h, A._ptr(np.array([0.0, 0.0, -1.0])), rhs.c))
    sysm.set_bsr(bsr)
    left = np.nonzero(X[:, 0] == 0.0)[0].astype(np.int32)
    for c in range(3):
        sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, left, 3, c, 0.0, 0.0)
    sysm.apply_matrix_transformation()
    sysm.apply_rhs_transformation()
    spmv = block_spmv_roofline(A, sysm, N, nb)
    sol64 = solve_best(A, sysm, reps=1, method=A.SOLVER_PCG_FP64, rtol=1e-8, max_iter=40000, check_every=32)
    x64 = sysm.solution().numpy()
    sol = solve_best(A, sysm, reps=1, x_fp64=x64, method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=40000,
                     check_every=32)
    out = {
        "workload": "3D linear elasticity P1, 96^3 hexes x 6 tets, E 21e5, nu 0.28, body force (0,0,-1), "
                    "x = 0 clamped, BSR 3x3",
        "nodes": N, "dofs": 3 * N, "tets": T, "blocks": nb,
        "assembly_ms": asm_ms, "assembly_kernel_ms": k_ms, "algorithmic_bytes": alg,
        "assembly_GBs": alg / (k_ms * 1e-3) / 1e9 if k_ms == k_ms else None,
        "assembled_dofs_per_s": 3 * N / (asm_ms * 1e-3),
        "block_jacobi_pcg_fp64": sol64,
        "block_jacobi_pcg_mixed": sol,
        "spmv_bsr3": spmv,
    }
    log(f"[bench] elasticity 96^3: assembly {asm_ms:.2f} ms ({out['assembly_GBs']:.0f} GB/s), "
        f"PCG {sol['iterations']} it {sol['solve_ms']:.1f} ms relres {sol['relres_fp64']:.1e}")
    return out


def block_spmv_roofline(A, sysm, N, nb):
    """FP64 and FP32 SpMV of a 3x3-block system (the elasticity matrix):
    algorithmic bytes 8 nnz_s + 4 nnz_b + 4 (N + 1) + 16 DoFs (FP64 values
    and vectors) and 4 nnz_s + 4 nnz_b + 4 (N + 1) + 8 DoFs (FP32)."""
    d = 3 * N
    x = A.DeviceArray.from_numpy(np.random.default_rng(1).normal(size=d))
    y = A.DeviceArray(d, np.float64)
    x32 = A.DeviceArray.from_numpy(np.random.default_rng(1).normal(size=d).astype(np.float32))
    y32 = A.DeviceArray(d, np.float32)
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
    nnz_s = 9 * nb
    b64 = 8 * nnz_s + 4 * nb + 4 * (N + 1) + 16 * d
    b32 = 4 * nnz_s + 4 * nb + 4 * (N + 1) + 8 * d
    fmt = sysm.spmv_format()
    out = {"format": ("bsr3_csr%d" % fmt[1]) if fmt[0] == "csr" else "sell3",
           "fp64": {"kernel_ms": s64, "algorithmic_bytes": b64, "GBs": b64 / (s64 * 1e-3) / 1e9,
                    "frac": b64 / (s64 * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "fp32": {"kernel_ms": s32, "algorithmic_bytes": b32, "GBs": b32 / (s32 * 1e-3) / 1e9,
                    "frac": b32 / (s32 * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    log(f"[bench] elasticity spmv ({out['format']}): fp64 {s64:.3f} ms ({out['fp64']['GBs']:.0f} GB/s), "
        f"fp32 {s32:.3f} ms ({out['fp32']['GBs']:.0f} GB/s)")
    return out


def q1_section(A, args):
    """Q1 elements: Laplace assembly (values + constant-source RHS) on a
    jittered 128^3 Hexa8 mesh and a 1024^2 Quad4 mesh (jitter 0.15h), and a
    mixed-precision PCG solve of the Hexa8 system (f = 1, u = 0 on z = 0) to
    1e-8."""
    out = {}
    for dim, n, kname in ((3, 128, "assemble_laplace_hexa8"), (2, 1024, "assemble_laplace_quad4")):
        X, C = A.box_mesh_q1(dim, n, 0.15, seed=args.seed)
        N, T, npc = X.shape[0], C.shape[0], C.shape[1]
        mesh = A.Mesh(dim, X, C, npc=npc)
        bsr = A.BSRFormat(mesh, 1, use_csr=True)
        bsr.compute_sparsity()
        nnz = int(bsr.info().nb_col)
        sysm = A.DoFLinearSystem(N)
        for _ in range(2):
            bsr.assemble_laplace(None, 1.0, sysm.rhs())
        A.synchronize()
        A.profile_reset()
        A.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(10):
            bsr.assemble_laplace(None, 1.0, sysm.rhs())
        A.synchronize()
        asm_ms = (time.perf_counter() - t0) * 1e3 / 10
        A.profile_enable(False)
        k_ms, _ = kernel_avg_ms(A, kname)
        alg = 4 * max(4, npc) * T + 24 * N + 12 * nnz + 4 * (N + 1) + 8 * N
        rec = {"nodes": N, "cells": T, "nnz": nnz, "assembly_ms": asm_ms, "assembly_kernel_ms": k_ms,
               "algorithmic_bytes": alg, "assembly_GBs": alg / (k_ms * 1e-3) / 1e9 if k_ms == k_ms else None,
               "assembled_dofs_per_s": N / (asm_ms * 1e-3)}
        if dim == 3:
            sysm.set_bsr(bsr)
            bottom = np.nonzero(X[:, 2] == 0.0)[0].astype(np.int32)
            sysm.apply_dirichlet(A.DIRICHLET_ROW_ELIMINATION, bottom, 1, 0, 0.0, 0.0)
            sysm.apply_matrix_transformation()
            sysm.apply_rhs_transformation()
            rec["mixed_pcg"] = solve_best(A, sysm, reps=1, method=A.SOLVER_PCG_MIXED, rtol=1e-8, max_iter=40000)
        key = "hexa8_128" if dim == 3 else "quad4_1024"
        out[key] = rec
        log(f"[bench] {key}: assembly {asm_ms:.2f} ms (kernel {k_ms:.3f} ms, {rec['assembly_GBs']:.0f} GB/s), "
            f"{N} nodes, nnz {nnz}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="hexes per side")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=9, help="timed repetitions of the CPU baseline")
    ap.add_argument("--no-learned", action="store_true", help="skip the heterogeneous / learned-PCG section")
    ap.add_argument("--train-n", type=int, default=64, help="hexes per side of the training meshes")
    ap.add_argument("--train-systems", type=int, default=512, help="training systems (one Adam step each)")
    ap.add_argument("--train-epochs", type=int, default=1)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D systems section")
    ap.add_argument("--batch-systems", type=int, default=64)
    ap.add_argument("--batch-max-iter", type=int, default=100000)
    ap.add_argument("--no-sequential", action="store_true", help="skip the one-at-a-time comparison")
    ap.add_argument("--heat-steps", type=int, default=10, help="time steps of the batched heat loop (0 = skip)")
    ap.add_argument("--heat-dt", type=float, default=0.01)
    ap.add_argument("--no-extra", action="store_true", help="skip the 2D Poisson and elasticity sections")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one 256 MiB device-to-device copy first (reference for HBM byte counters)")
    ap.add_argument("--assembly-only", action="store_true",
                    help="pattern and the timed assembly steps only (the entry point of counter runs)")
    ap.add_argument("--elasticity-assembly-only", action="store_true",
                    help="the timed elasticity assembly steps only (the entry point of counter runs)")
    ap.add_argument("--variant", type=int, default=None,
                    help="pin the Laplace assembly variant instead of autotuning")
    ap.add_argument("--cpu-solve-reps", type=int, default=5,
                    help="timed repetitions of the all-cores CPU Jacobi-PCG (0 = skip the CPU solve)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this implementation runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the path across GPUs is designed but not built)", file=sys.stderr)
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
    if args.elasticity_assembly_only:
        print(json.dumps(elasticity_section(A, args, assembly_only=True)), flush=True)
        return
    copy = measure_copy(A) if not args.assembly_only else None
    if copy:
        log(f"[bench] stream copy: {copy['GBs']:.0f} GB/s ({copy['bytes_moved'] / 1e9:.2f} GB moved in "
            f"{copy['kernel_ms']:.3f} ms)")
    n = args.n
    t0 = time.time()
    X, C = A.box_mesh(3, n, 0.15, seed=args.seed)
    N, T = X.shape[0], C.shape[0]
    # smallest signed 6|T| over the unjittered cell's h^3: the jittered mesh stays positively oriented
    e1, e2, e3 = X[C[:, 1]] - X[C[:, 0]], X[C[:, 2]] - X[C[:, 0]], X[C[:, 3]] - X[C[:, 0]]
    min_det_ratio = float(np.einsum("ij,ij->i", e1, np.cross(e2, e3)).min()) * float(n) ** 3
    del e1, e2, e3
    assert min_det_ratio > 0.0, "inverted cell in the benchmark mesh"
    log(f"[bench] mesh {n}^3: N={N} T={T}, smallest signed 6|T| / h^3 = {min_det_ratio:.3f} "
        f"({time.time() - t0:.1f}s)")
    mesh = A.Mesh(3, X, C)
    bsr = A.BSRFormat(mesh, 1, use_csr=True)
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    t0 = time.perf_counter()
    bsr.compute_sparsity()
    A.synchronize()
    pattern_ms = (time.perf_counter() - t0) * 1e3
    A.profile_enable(False)
    prow_n, prow_ms = A.profile_query("pattern_rows")
    info = bsr.info()
    nnz = int(info.nb_col)
    pat_bytes = 16 * T + 8 * nnz + 4 * N
    log(f"[bench] pattern nnz={nnz} ({nnz / N:.2f}/row) in {pattern_ms:.1f} ms "
        f"(row kernel {prow_ms:.2f} ms over {prow_n} passes, {pat_bytes / (prow_ms * 1e-3) / 1e9:.0f} GB/s)")
    sysm = A.DoFLinearSystem(N)
    rhs = sysm.rhs()
    if args.variant is not None:
        bsr.set_assembly_variant(args.variant)
        variant, tune_ms = args.variant, float("nan")
    else:
        variant, tune_ms = bsr.autotune_assembly()
    log(f"[bench] assembly variant {variant} ({'pinned' if args.variant is not None else 'autotuned'}, "
        f"{tune_ms:.3f} ms)")
    # set-up a caller pays with the variant known: the pattern, then the first
    # assembly (which builds that variant's plan), on a fresh handle
    A.synchronize()
    A.profile_reset()
    A.profile_enable(True)
    t0 = time.perf_counter()
    bsr2 = A.BSRFormat(mesh, 1, use_csr=True)
    bsr2.compute_sparsity()
    bsr2.set_assembly_variant(variant)
    bsr2.assemble_laplace(None, 1.0, rhs)
    A.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    A.profile_enable(False)
    plan_ms = {k: A.profile_query(k)[1] for k in ("plan_brick", "plan_sliced", "plan_cell")}
    del bsr2
    log(f"[bench] set-up with the variant known: {setup_ms:.1f} ms (pattern + plan "
        + ", ".join(f"{k} {v:.1f} ms" for k, v in plan_ms.items() if v > 0) + " + first assembly)")

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
            "min_signed_det_over_h3": min_det_ratio,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload_key, akernel),
            "kernel": akernel, "kernel_ms": k_ms, "algorithmic_bytes": alg_bytes,
            "bytes_formula": "16*T + 24*N + 12*nnz + 4*(N+1)", "assembly_variant": variant,
            "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT),
        },
        "pattern_ms": pattern_ms,
        "setup_ms": setup_ms,
        "setup": {"total_ms": setup_ms, "plan_ms": plan_ms,
                  "what": "fresh handle: computeSparsity (count, scan, fill, pair plan) + the chosen variant's "
                          "plan + first assembly"},
        "pattern": {"total_ms": pattern_ms, "row_kernel_ms": prow_ms, "row_kernel_passes": prow_n,
                    "algorithmic_bytes": pat_bytes, "bytes_formula": "16*T + 8*nnz + 4*N (count + fill passes)",
                    "GBs": pat_bytes / (prow_ms * 1e-3) / 1e9,
                    "frac": pat_bytes / (prow_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "kernel": "k_pattern_rows"},
    }
    if copy:
        out["hbm_copy"] = copy
        out["roofline"]["copy_measured_GBs"] = copy["GBs"]
        out["roofline"]["frac_of_copy"] = achieved / copy["GBs"]
    if args.assembly_only:
        print(json.dumps(out), flush=True)
        return

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
    if fmt[0] == "sell":
        k64, k32 = "k_spmv_sell<double, double, 0>", "k_spmv_sell<float, float, 0>"
    elif fmt[0].startswith("sell16"):
        (u64, bs64), (u32, bs32) = sysm.spmv_tuning()
        k64 = f"k_spmv_sell16<double, double, 0, {u64}, {bs64}>"
        k32 = f"k_spmv_sell16<float, float, 0, {u32}, {bs32}>"
    else:
        k64, k32 = f"k_spmv_csr<{fmt[1]}, double, double, 0>", f"k_spmv_csr<{fmt[1]}, float, float, 0>"
    out["spmv"] = {
        "format": fmt[0] if fmt[0] != "csr" else f"csr{fmt[1]}",
        "fp64": {"kernel_ms": s64, "algorithmic_bytes": b64, "GBs": b64 / (s64 * 1e-3) / 1e9,
                 "frac": b64 / (s64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": k64,
                 "traffic": measured_traffic(workload_key, k64)},
        "fp32": {"kernel_ms": s32, "algorithmic_bytes": b32, "GBs": b32 / (s32 * 1e-3) / 1e9,
                 "frac": b32 / (s32 * 1e-3) / 1e9 / HBM_PEAK_GBS, "kernel": k32,
                 "traffic": measured_traffic(workload_key, k32)},
        "bytes_note": "algorithmic bytes are CSR-equivalent (12 nnz + 20 N FP64, 8 nnz + 12 N FP32); "
                      "traffic_GBs is the format's measured HBM traffic over the same time",
    }
    for key, ms in (("fp64", s64), ("fp32", s32)):
        rec = out["spmv"][key]
        if rec["traffic"]:
            rec["traffic_GBs"] = rec["traffic"] / (ms * 1e-3) / 1e9
            rec["traffic_frac"] = rec["traffic_GBs"] / HBM_PEAK_GBS
        if copy:
            rec["frac_of_copy"] = rec["GBs"] / copy["GBs"]
    log(f"[bench] spmv fp64 {s64:.3f} ms ({out['spmv']['fp64']['GBs']:.0f} GB/s), "
        f"fp32 {s32:.3f} ms ({out['spmv']['fp32']['GBs']:.0f} GB/s)")

    if not args.no_solve:
        solves = {}
        kw = dict(rtol=1e-8, max_iter=20000, check_every=32)
        solves["jacobi_pcg_fp64"] = solve_best(A, sysm, method=A.SOLVER_PCG_FP64, **kw)
        x64 = sysm.solution().numpy()
        solves["mixed_pcg"] = solve_best(A, sysm, x_fp64=x64, method=A.SOLVER_PCG_MIXED, **kw)
        for name in ("jacobi_pcg_fp64", "mixed_pcg"):
            r = solves[name]
            log(f"[bench] {name}: {r['iterations']} it, {r['solve_ms']:.1f} ms, relres {r['relres_fp64']:.2e}"
                + (f", |x - x_fp64|/|x_fp64| {r['error_vs_fp64_solution']:.1e}"
                   if "error_vs_fp64_solution" in r else ""))
        solves["speedup_mixed_vs_fp64"] = solves["jacobi_pcg_fp64"]["solve_ms"] / solves["mixed_pcg"]["solve_ms"]
        solves["tolerance"] = 1e-8
        if not args.no_cpu_baseline and args.cpu_solve_reps > 0:
            solves["cpu_jacobi_pcg_fp64"] = cpu_solve_baseline(A, args, bsr, sysm, N)
            solves["speedup_gpu_fp64_vs_cpu_all_cores"] = (solves["cpu_jacobi_pcg_fp64"]["all_cores"]["solve_s"] * 1e3
                                                         / solves["jacobi_pcg_fp64"]["solve_ms"])
        out["solve"] = solves

    # ---- heterogeneous Poisson with the learned preconditioner -------------------------------
    if not args.no_solve and not args.no_learned:
        out["learned"] = learned_section(A, args, X, C, mesh, bsr, N, nnz)

    if not args.no_solve and not args.no_batch:
        out["batch"] = batch_section(A, args)
    if not args.no_solve and not args.no_extra:
        out["poisson2d"] = poisson2d_section(A, args)
        out["elasticity"] = elasticity_section(A, args)
        out["q1"] = q1_section(A, args)

    # ---- CPU baseline: the oracle's assembly of the same mesh, all host threads and one core ------
    if not args.no_cpu_baseline:
        import oracle as O
        row, col, _, _ = bsr.download()
        nthreads = O.max_threads()
        O.assemble_laplace_threads(X, C, 4, row, col, 1.0, nthreads)  # warm-up (page-in)
        times = []
        for _ in range(args.cpu_reps):
            t0 = O.wtime()
            O.assemble_laplace_threads(X, C, 4, row, col, 1.0, nthreads)
            times.append(O.wtime() - t0)
        t = float(np.median(times))
        times1 = []
        for _ in range(max(5, args.cpu_reps // 3)):
            t0 = O.wtime()
            O.assemble(3, X, C, 4, row, col)
            O.rhs_constant_source(3, X, C, 4, 1.0)
            times1.append(O.wtime() - t0)
        t1 = float(np.median(times1))
        out["cpu_baseline"] = {
            "value": N / t, "unit": "DoF/s", "cores": nthreads, "kind": "port",
            "sample": f"full {n}^3 assembly (cell-wise, matrix values with per-entry row search + RHS) on "
                      f"{nthreads} threads (atomic scatter), 1 warm-up + median of {args.cpu_reps} "
                      f"({t:.3f} s each); one thread (serial oracle): median of {len(times1)} ({t1:.2f} s each); "
                      f"{cpu_model()}",
            "value_1core": N / t1,
        }
        log(f"[bench] cpu baseline: {N / t / 1e6:.2f} MDoF/s on {nthreads} threads ({t:.3f} s), "
            f"{N / t1 / 1e6:.2f} MDoF/s on one ({t1:.2f} s)")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
