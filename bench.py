#!/usr/bin/env python3
"""Time-steps/s of the 3D hex lid-driven cavity at Re=1000: mixed precision vs FP64 on one GPU.

Workload (default ``c3``): unit-cube cavity of 115^3 trilinear hexes (1,520,875
elements, 1,560,896 nodes), nodes graded toward every wall by
x_i = 0.5 (1 - cos(pi i / 115)), rho = 1, mu = 1e-3, lid u = 1 on z = 1 ramped by
min(t, 1), p pinned at (0, 0, 0), generalised-alpha rho_inf = 0, dt = 0.1.
``c2`` is the same at 32^3.

Three runs of K time steps from the zero state, each after W warm-up steps (the
state is reset after warm-up):
  * FP64 plain: the reference's solver configuration, FP64 tangent and CSR, BiCGStab +
    ILU(0) at rtol 1e-8, the factor rebuilt for every new tangent;
  * FP64: the same with the solver options of the mixed run: ILU(0) with the
    multiplicative coarse pressure correction on 8^3 boxes, the factor rebuilt on every
    6th tangent;
  * mixed: FP32 element tangents and FP32 CSR, FP32 BiCGStab + ILU(0) (factor and
    tile inverses stored in BF16) with the coarse correction without its global
    constant mode, inside an FP64 residual / Newton update loop (defect correction) to
    Newton 1e-9 whose every iteration first takes the FP64 coarse step
    x += Z E^-1 Z^T F(x); each FP32 solve is asked for
    min(0.5, max(4e-5, 0.8 * 1e-9 / ||F||)) (inexact Newton).
``value`` is the mixed run's time steps per second; the JSON line also carries
both FP64 rates, the speed-ups, the relative L2 error of u and p (mixed vs FP64),
per-kernel-class device times measured with HIP events on the library's stream,
the roofline of the dominant kernel and a CPU baseline.
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

# HBM traffic per launch of the dominant kernel (bytes read + bytes written), from the
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes in profiles/c3_bsr_spmv_hbm_traffic.md,
# calibrated against a 1 GiB copy; keyed by (kernel, value type, mesh size n).
MEASURED_TRAFFIC = {("k_bsr_spmv", "float", 115): 3.4262e9 + 0.0250e9,
                    ("k_bsr_spmv", "double", 115): 5.9633e9 + 0.0506e9}
TRAFFIC_SOURCE = "profiles/c3_bsr_spmv_hbm_traffic.md"

METRIC = "time-steps/s, 3D hex cavity Re1000, mixed vs FP64 GPU at <=1e-6 rel. L2 error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HEXA8_FLOPS_KF = 62723   # reference arithmetic per Hexa8 element, Ke + Fe (oracle op count)
HEXA8_FLOPS_F = 7939     # residual only


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(n):
    from cfdsfemmpi_amd import deck, meshgen
    mesh = meshgen.ldc3d_hex(n, graded=True)
    ctl = deck.Control(np.array([1.0, 1e-3, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0]), 2, 0.0, 0.1, 1e9, 10, 1, 1e-6)
    return mesh, ctl


def run_path(mesh, ctl, prec, opts, newton_tol, steps, warmup, label, predictor=0, pc_lag=1, inner=(0.0, 0.5),
             tuning=0, coarse=0, coarse_mult=False):
    from cfdsfemmpi_amd import sfem
    t0 = time.time()
    fem = sfem.StabFEM(device=0)
    fem.setMesh(mesh)
    fem.setControl(ctl)
    fem.precision = prec
    fem.tuning = tuning
    fem.coarse_boxes = coarse
    fem.coarse_mult = coarse_mult
    fem.prepareInputData()
    fem.setSolver(opts)
    ctx = fem.ctx
    ctx.set_predictor(predictor)
    ctx.set_pc_lag(pc_lag)
    ctx.set_inner_tolerance(*inner)
    info = ctx.info()
    log(f"{label}: setup {time.time() - t0:.1f} s, nfree={info['nfree']} nnz={info['nnz']} colours={info['ncolors']}")

    def advance(k, record):
        timeNow = ctl.dt
        out = []
        for s in range(k):
            tf = timeNow if timeNow <= 1.0 else 1.0
            st = ctx.step(tf, opts, newton_tol, 10)
            out.append((st.newton_iters, st.krylov_iters, st.rhs_norm0, st.rhs_norm))
            if record:
                log(f"{label}: step {s + 1}/{k}: newton {st.newton_iters} krylov {st.krylov_iters} "
                    f"|F0| {st.rhs_norm0:.3e} |F| {st.rhs_norm:.3e}")
            timeNow += ctl.dt
        return out

    if warmup:
        advance(warmup, False)
        ctx.reset_state()
    ctx.enable_kernel_timing(True)
    ctx.get_solution()  # drains the stream
    t1 = time.perf_counter()
    hist = advance(steps, True)
    s = ctx.get_solution()
    t2 = time.perf_counter()
    tune = ctx.tuning()
    log(f"{label}: autotuned kernel shapes {tune}")
    kt = {name: ctx.kernel_timing(k) for name, k in
          (("assembly", sfem.KT_ASSEMBLY), ("spmv", sfem.KT_SPMV), ("pc", sfem.KT_PC))}
    log(f"{label}: {steps} steps in {t2 - t1:.3f} s; device ms by class: " +
        ", ".join(f"{k} {v[1]:.1f} ({v[0]} launches)" for k, v in kt.items()))
    # one tangent + residual pass and one residual-only pass timed on their own (best of
    # three), after the timed steps (the state is not advanced by them)
    def pass_ms(want_k):
        best = None
        for _ in range(3):
            c0, m0 = ctx.kernel_timing(sfem.KT_ASSEMBLY)
            ctx.assemble(1, want_K=want_k)
            c1, m1 = ctx.kernel_timing(sfem.KT_ASSEMBLY)
            best = (m1 - m0) if best is None else min(best, m1 - m0)
        return best
    asm_pass = {"K_F_ms": pass_ms(True), "F_ms": pass_ms(False)}
    res = dict(label=label, seconds=t2 - t1, steps_per_s=steps / (t2 - t1), hist=hist, kt=kt, info=info,
               soln=s, tune=tune, block=ctx.block_info(), tiles=ctx.tile_info(), stats=ctx.solver_stats(),
               asm_pass=asm_pass, coarse=ctx.coarse_info())
    ctx.close()
    return res


def cpu_threads():
    """Host threads for the CPU baseline: the OMP_NUM_THREADS the host grants (a machine's
    share of cores per GPU), else every core."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "") or (os.cpu_count() or 1)))
    except ValueError:
        return max(1, os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(n_target, c2=True):
    """The oracle (C++ CPU restatement of the reference path: same element arithmetic, DOF
    numbering and CSR, MatSetValues-style binary-search ADD assembly, BiCGStab + ILU(0)
    at rtol 1e-8, Newton 1e-8) on all granted host threads: element colour classes in
    parallel, level-scheduled ILU(0), row-parallel products.  Measured, not scaled: the
    first time step of the target cavity (set-up -- pattern and prepareElemData -- not
    timed, as the reference's time-loop timer excludes it), and the 32^3 cavity's 20 steps
    in full.  The tangent is assembled once per step, as the HIP path does."""
    import oracle
    from cfdsfemmpi_amd import deck
    th = cpu_threads()
    oracle.set_threads(th)
    so = deck.SolverOptions("bcgs", "ilu", 1e-8, 5000)
    mesh, ctl = make_case(n_target)
    t0 = time.perf_counter()
    f = oracle.OracleFEM(mesh, ctl, so)
    f.prepare()
    t_setup = time.perf_counter() - t0
    tm = {}
    t0 = time.perf_counter()
    f.solve_fully_implicit(1e-8, 10, steps=1, reuse_K=True, timers=tm)
    t_step = time.perf_counter() - t0
    log(f"cpu baseline: {n_target}^3 step 1 in {t_step:.1f} s (assembly {tm['assembly']:.1f} s, solver "
        f"{tm['solver']:.1f} s, {tm.get('krylov_its', 0)} BiCGStab its) on {th} threads; set-up {t_setup:.1f} s")
    del f
    c2_res = None
    if c2:
        mesh2, ctl2 = make_case(32)
        f2 = oracle.OracleFEM(mesh2, ctl2, so)
        f2.prepare()
        tm2 = {}
        t0 = time.perf_counter()
        f2.solve_fully_implicit(1e-8, 10, steps=20, reuse_K=True, timers=tm2)
        t_c2 = time.perf_counter() - t0
        c2_res = {"workload": "32^3 graded Hexa8 cavity, 20 steps", "seconds": t_c2, "steps_per_s": 20 / t_c2,
                  "assembly_s": tm2["assembly"], "solver_s": tm2["solver"], "krylov_its": tm2.get("krylov_its")}
        log(f"cpu baseline: 32^3, 20 steps in {t_c2:.1f} s ({20 / t_c2:.3f} steps/s)")
        del f2
    oracle.set_threads(1)
    return {"value": 1.0 / t_step, "unit": "time-steps/s", "cores": th, "kind": "port",
            "sample": (f"oracle (C++ CPU restatement of the reference path) on {th} threads, measured: the first "
                       f"time step of the {n_target}^3 cavity (assembly {tm['assembly']:.2f} s, BiCGStab + "
                       f"natural-order ILU(0) {tm.get('krylov_its', 0)} its {tm['solver']:.2f} s); set-up "
                       f"{t_setup:.1f} s not timed"),
            "phases_s": {"assembly": tm["assembly"], "solver": tm["solver"]},
            "c2_full_run": c2_res, "cpu": cpu_model(), "host_threads": os.cpu_count()}


C4_METRIC = "time-steps/s of a 256-cavity Tria3 batch (every mesh advances one step), batched vs one mesh at a time"
C4_DT = 0.05
TRIA3_FLOPS_KF = 1010    # reference arithmetic per Tria3 element, Ke + Fe (hand count)


def c4_case(count, seed):
    from cfdsfemmpi_amd import meshgen
    return meshgen.c4_batch(count, seed=seed)


def c4_ctx(mesh, em, mu, prec, opts, coarse=0, mult=True):
    from cfdsfemmpi_amd import sfem
    ctx = sfem.Context(mesh, device=0)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(np.array([1.0, float(np.mean(mu)), 0, 0, 0, 0, 0, 0, 1, 1, 0]), 2, 0.0, C4_DT)
    ctx.set_precision(prec)
    if em is not None:
        ctx.set_batch(em, mu)
    if coarse:
        ctx.set_coarse_correction(coarse, mult)  # per mesh for a batch
    ctx.setup_solver(opts)
    return ctx


def run_batch(case, prec, opts, newton_tol, steps, warmup, label, graph=True, coarse=0, mult=True):
    """Every mesh of the batch advances ``steps`` time steps together (sfem_step_batch)."""
    from cfdsfemmpi_amd import sfem
    mesh, em, mu, ns, meshes = case
    t0 = time.time()
    ctx = c4_ctx(mesh, em, mu, prec, opts, coarse, mult)
    ctx.batch_use_graph(graph)
    info = ctx.info()
    log(f"{label}: setup {time.time() - t0:.1f} s, nfree={info['nfree']} nnz={info['nnz']} colours={info['ncolors']}")

    def advance(k, record):
        t = C4_DT
        out = []
        for s in range(k):
            ni, ki, rn = ctx.step_batch(min(t, 1.0), opts, newton_tol, 10)
            out.append((ni.copy(), ki.copy(), rn.copy()))
            if record:
                log(f"{label}: step {s + 1}/{k}: newton max {ni.max()} krylov max {ki.max()} "
                    f"(mean {ki.mean():.1f}) max |F| {rn.max():.3e}")
            t += C4_DT
        return out

    if warmup:
        advance(warmup, False)
        ctx.reset_state()
    ctx.get_solution()
    t1 = time.perf_counter()
    hist = advance(steps, True)
    s = ctx.get_solution()
    t2 = time.perf_counter()
    log(f"{label}: {steps} steps in {t2 - t1:.3f} s")
    # one more step, outside the timed region, issued launch by launch over every mesh
    # (no graph, no narrowing) for per-class device times of full-size launches
    ctx.batch_use_graph(False)
    ctx.batch_narrowing(False)
    ctx.enable_kernel_timing(True)
    ctx.step_batch(min(C4_DT * (steps + 1), 1.0), opts, newton_tol, 10)
    kt = {name: ctx.kernel_timing(k) for name, k in
          (("assembly", sfem.KT_ASSEMBLY), ("spmv", sfem.KT_SPMV), ("pc", sfem.KT_PC))}
    res = dict(label=label, seconds=t2 - t1, steps_per_s=steps / (t2 - t1), hist=hist, kt=kt,
               info=info, soln=s, block=ctx.block_info(), tune=ctx.tuning(), stats=ctx.solver_stats(),
               element_variant=ctx.element_tuning())
    ctx.close()
    return res


def run_sequential_sample(case, prec, opts, newton_tol, steps, stride, coarse=0):
    """Meshes 0, stride, 2*stride, ... stepped one after another through the single-mesh
    path (one context each, setup untimed): the rate a caller without batching gets."""
    from cfdsfemmpi_amd import sfem
    mesh, em, mu, ns, meshes = case
    idx = list(range(0, len(meshes), stride))
    ctxs = [c4_ctx(meshes[k], None, mu[k:k + 1], prec, opts, coarse) for k in idx]
    for c in ctxs:
        c.get_solution()
    t1 = time.perf_counter()
    for c in ctxs:
        t = C4_DT
        for s in range(steps):
            c.step(min(t, 1.0), opts, newton_tol, 10)
            t += C4_DT
        c.get_solution()
    t2 = time.perf_counter()
    for c in ctxs:
        c.close()
    sec_all = (t2 - t1) * len(meshes) / len(idx)
    return {"meshes_sampled": idx, "seconds_sample": t2 - t1, "steps_per_s_estimate": steps / sec_all}


def cpu_baseline_c4(case, stride):
    """Oracle (single-threaded CPU restatement) on one FP64 time step of meshes 0, stride,
    ... (BiCGStab + ILU(0) at rtol 1e-8, Newton 1e-8), scaled to the whole batch by count."""
    import oracle
    from cfdsfemmpi_amd import deck
    mesh, em, mu, ns, meshes = case
    idx = list(range(0, len(meshes), stride))
    so = deck.SolverOptions("bcgs", "ilu", 1e-8, 5000)
    t = 0.0
    for k in idx:
        ctl = deck.Control(np.array([1.0, mu[k], 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0]), 2, 0.0, C4_DT, 1e9, 1, 1, 1e-8)
        f = oracle.OracleFEM(meshes[k], ctl, so)
        f.prepare()
        t0 = time.perf_counter()
        f.solve_fully_implicit(1e-8, 10, steps=1)
        t += time.perf_counter() - t0
    per_step = t * len(meshes) / len(idx)
    return {"value": 1.0 / per_step, "unit": "time-steps/s", "cores": 1, "kind": "port",
            "sample": (f"oracle (C++ CPU restatement, 1 thread), first time step of {len(idx)} of the "
                       f"{len(meshes)} meshes (k % {stride} == 0, n = {[int(ns[k]) for k in idx]}) in "
                       f"{t:.2f} s, scaled to the batch by mesh count"),
            "cpu": platform.processor() or platform.machine(), "host_threads": os.cpu_count()}


def main_c4(a):
    from cfdsfemmpi_amd import sfem
    steps = a.steps if a.steps is not None else 20
    case = c4_case(a.c4_count, a.c4_seed)
    mesh, em, mu, ns, meshes = case
    log(f"workload: {len(ns)} Tria3 cavities, n in [{ns.min()}, {ns.max()}], {mesh.nElem} elements, "
        f"{mesh.nNode} nodes, Re in [{1 / mu.max():.0f}, {1 / mu.min():.0f}]")
    o64 = sfem.solver_opts("bcgs", "ilu", 1e-8, 5000, reorder=a.reorder)
    omx = sfem.solver_opts("bcgs", "ilu", a.inner_rtol, 5000, reorder=a.reorder, factor_bf16=a.bf16_factors)
    co = a.c4_coarse
    rpl = None
    if co and a.plain:
        # the reference's solver configuration: ILU(0) alone
        rpl = run_batch(case, sfem.PREC_FP64, o64, a.newton_tol, steps, a.warmup, "fp64-batch-plain",
                        graph=bool(a.c4_graph), mult=bool(a.c4_coarse_mult))
        log(f"fp64-batch-plain: {rpl['steps_per_s']:.3f} steps/s")
    r64 = run_batch(case, sfem.PREC_FP64, o64, a.newton_tol, steps, a.warmup, "fp64-batch", coarse=co,
                    graph=bool(a.c4_graph), mult=bool(a.c4_coarse_mult))
    log(f"fp64-batch: {r64['steps_per_s']:.3f} steps/s")
    node_mesh = np.concatenate([np.full(m.nNode, k) for k, m in enumerate(meshes)])

    def mesh_errors(s):
        a64, b = r64["soln"].reshape(-1, 3), s.reshape(-1, 3)
        eu = np.sqrt(np.bincount(node_mesh, ((b[:, :2] - a64[:, :2]) ** 2).sum(1)) /
                     np.bincount(node_mesh, (a64[:, :2] ** 2).sum(1)))
        ep = np.sqrt(np.bincount(node_mesh, (b[:, 2] - a64[:, 2]) ** 2) /
                     np.maximum(np.bincount(node_mesh, a64[:, 2] ** 2), 1e-300))
        return float(eu.max()), float(ep.max())

    r32 = None
    if a.c4_store32:
        o32 = sfem.solver_opts("bcgs", "ilu", a.c4_store32_rtol, 5000, reorder=a.reorder, factor_bf16=a.bf16_factors)
        r32 = run_batch(case, sfem.PREC_STORE32, o32, a.c4_store32_newton, steps, a.warmup, "store32-batch",
                        coarse=co, graph=bool(a.c4_graph), mult=bool(a.c4_coarse_mult))
        r32["err"] = mesh_errors(r32["soln"])
        log(f"store32-batch: {r32['steps_per_s']:.3f} steps/s, per-mesh rel L2 vs fp64: u max {r32['err'][0]:.3e} "
            f"p max {r32['err'][1]:.3e}")
    rmx = err_u = err_p = None
    if a.c4_mixed:
        rmx = run_batch(case, sfem.PREC_MIXED, omx, a.newton_tol_mixed, steps, a.warmup, "mixed-batch", coarse=co,
                        graph=bool(a.c4_graph), mult=bool(a.c4_coarse_mult))
        log(f"mixed-batch: {rmx['steps_per_s']:.3f} steps/s")
        err_u, err_p = mesh_errors(rmx["soln"])
        log(f"per-mesh rel L2 mixed vs fp64: u max {err_u:.3e}, p max {err_p:.3e}")
    seq = None
    if a.c4_seq_stride:
        seq = run_sequential_sample(case, sfem.PREC_FP64, o64, a.newton_tol, steps, a.c4_seq_stride, co)
        log(f"one mesh at a time (fp64), {len(seq['meshes_sampled'])} meshes in {seq['seconds_sample']:.2f} s: "
            f"~{seq['steps_per_s_estimate']:.3f} batch-steps/s")
    kt = r64["kt"]
    bi, info = r64["block"], r64["info"]
    nnz, nf = info["nnz"], info["nfree"]
    cnt, ms = kt["spmv"]
    sbytes = nnz * 8 + 4 * bi["npairs"] + 16 * bi["nblk"] + 2 * 8 * nf
    roof = None
    if cnt:
        avg = ms / cnt
        roof = {"bound": "hbm", "kernel": "k_bsr_spmv<double> (node-block SpMV over the whole batch)",
                "achieved": sbytes / (avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "avg_launch_ms": avg, "launches": cnt, "algorithmic_bytes_per_launch": sbytes,
                "traffic": None, "timed_on": "one step after the timed region, every launch over every mesh"}
        roof["frac"] = roof["achieved"] / roof["peak"]
    cpu = None if a.no_cpu else cpu_baseline_c4(case, a.c4_cpu_stride)
    # the reported rate: the FP32-stored run when it meets the velocity gate, else FP64
    best = r32 if (r32 is not None and r32["err"][0] <= 1e-6 and r32["steps_per_s"] > r64["steps_per_s"]) else r64
    out = {
        "metric": C4_METRIC,
        "value": best["steps_per_s"],
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * best["seconds"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best is r64 else "f64 Krylov / f32-stored tangent+ILU" + (" (bf16 factors)" if a.bf16_factors else ""),
        "data": f"synthetic (generated meshes, seed {a.c4_seed}, zero initial state)",
        "config": {"workload": f"{len(ns)} jittered unit-square Tria3 cavities, n in [16, 160], Re log-uniform "
                               f"in [50, 2000], dt={C4_DT}, {steps} steps, one batched solve per step",
                   "n_elements": mesh.nElem, "n_nodes": mesh.nNode, "nfree": nf, "nnz": nnz,
                   "seed": a.c4_seed, "reported_run": best["label"],
                   "solver": "BiCGStab + ILU(0), per-mesh scalars and convergence" +
                             (f", multiplicative coarse pressure correction on {co}^2 boxes per mesh" if co else ""),
                   "fp64_run": {"rtol": 1e-8, "newton_tol": a.newton_tol},
                   "graph": "one BiCGStab iteration replayed as a HIP graph",
                   "narrowing": "iteration restricted to unfinished meshes", "parallelism": "single GPU"},
        "fp64_batch_steps_per_s": r64["steps_per_s"],
        "fp64_batch_plain_steps_per_s": rpl["steps_per_s"] if rpl else None,
        "fp64_batch_plain_krylov_iters_max": [int(h[1].max()) for h in rpl["hist"]] if rpl else None,
        "one_mesh_at_a_time_fp64": seq,
        "speedup_fp64_batch_vs_one_at_a_time": (r64["steps_per_s"] / seq["steps_per_s_estimate"]) if seq else None,
        "store32_batch": None if r32 is None else {
            "steps_per_s": r32["steps_per_s"], "inner_rtol": a.c4_store32_rtol, "newton_tol": a.c4_store32_newton,
            "ilu_factors": "bf16" if a.bf16_factors else "fp32",
            "speedup_vs_fp64_batch": r32["steps_per_s"] / r64["steps_per_s"],
            "error_gate_1e-6_passed": r32["err"][0] <= 1e-6,
            "rel_l2_u_vs_fp64_max_over_meshes": r32["err"][0], "rel_l2_p_vs_fp64_max_over_meshes": r32["err"][1],
            "newton_iters_max": [int(h[0].max()) for h in r32["hist"]],
            "krylov_iters_max": [int(h[1].max()) for h in r32["hist"]]},
        "mixed_batch": None if rmx is None else {
            "steps_per_s": rmx["steps_per_s"], "inner_rtol": a.inner_rtol, "newton_tol": a.newton_tol_mixed,
            "ilu_factors": "bf16" if a.bf16_factors else "fp32",
            "speedup_vs_fp64_batch": rmx["steps_per_s"] / r64["steps_per_s"],
            "error_gate_1e-6_passed": err_u <= 1e-6,
            "rel_l2_u_vs_fp64_max_over_meshes": err_u, "rel_l2_p_vs_fp64_max_over_meshes": err_p,
            "fp32_mesh_solves_kept_best_iterate": rmx["stats"][1],
            "newton_iters_max": [int(h[0].max()) for h in rmx["hist"]],
            "krylov_iters_max": [int(h[1].max()) for h in rmx["hist"]]},
        "krylov_iters_max": [int(h[1].max()) for h in r64["hist"]],
        "krylov_iters_mean": [round(float(h[1].mean()), 1) for h in r64["hist"]],
        "newton_iters_max": [int(h[0].max()) for h in r64["hist"]],
        "element_loop_variant": {r["label"]: r["element_variant"] for r in (r64, r32, rmx) if r},
        "kernel_ms_full_step": {r["label"]: {k: {"launches": v[0], "total_ms": v[1]} for k, v in r["kt"].items()}
                                for r in (r64, r32, rmx) if r},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


C5_METRIC = "C5: Newton/Krylov iterations and time-steps/s, learned vs extrapolated initial iterate"


def c5_context(mesh, mu, opts, coarse=0):
    from cfdsfemmpi_amd import deck, sfem
    ctl = deck.Control(np.array([1.0, mu, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0]), 2, 0.0, 0.1, 1e9, 100000, 1, 1e-8)
    fem = sfem.StabFEM(device=0)
    fem.setMesh(mesh)
    fem.setControl(ctl)
    fem.coarse_boxes = coarse  # no prescribed pressure: the correction is E's pseudo-inverse
    fem.coarse_mult = True
    fem.prepareInputData()
    fem.setSolver(opts)
    fem.ctx.set_predictor(sfem.PRED_EXTRAPOLATE)
    return fem


def c5_run(ctx, opts, steps, newton_tol, record=None, record_from=0, label="", judged=False):
    """steps time steps from the zero state (t = dt, 2 dt, ..., inflow ramped by min(t, 1));
    per-step Newton iterations, Krylov iterations and iteration-0 residual norms.  With
    a learner, the converged states of steps record_from, record_from + 1, ... are
    recorded as one trajectory."""
    newton, krylov, r0, cand = [], [], [], []
    t, dt = 0.1, 0.1
    t0 = time.perf_counter()
    for k in range(steps):
        st = ctx.step(min(t, 1.0), opts, newton_tol, 10, allow_notconv=False)
        newton.append(st.newton_iters)
        krylov.append(st.krylov_iters)
        r0.append(st.rhs_norm0)
        if judged:
            cand.append([float(v) for v in ctx.predictor_stats()[1]])
        if record is not None and k + 1 >= record_from:
            record.record(ctx, new_trajectory=(k + 1 == record_from))
        t += dt
    sec = time.perf_counter() - t0
    log(f"{label}: {steps} steps in {sec:.2f} s ({steps / sec:.2f} steps/s), Newton {sum(newton)}, "
        f"Krylov {sum(krylov)}")
    return {"seconds": sec, "steps_per_s": steps / sec, "newton": newton, "krylov": krylov, "r0": r0,
            "candidates": cand}


def c5_train_batch(meshes, res, a):
    """Every training trajectory in one batched context (the meshes side by side, one
    viscosity each, sfem_step_batch per time step; FP64 BiCGStab + ILU(0) at the relative
    tolerance --c5-train-rtol, Newton to the absolute tolerance per mesh), the batch's
    converged states recorded from step --c5-record-from on.  The learner's samples are
    (window, node) pairs, so one store over the union is the trajectories' stores side by
    side.  Returns a learner trained on nothing yet (the union's store) and the rate."""
    from cfdsfemmpi_amd import meshgen, sfem
    mesh, em = meshgen.disjoint_union(meshes)
    mu = 1.0 / np.asarray(res, np.float64)
    opts = sfem.solver_opts("bcgs", "ilu", a.c5_train_rtol, 20000, reorder=a.reorder)
    ctx = sfem.Context(mesh, device=0)
    ctx.set_dirichlet(mesh.dbc_node, mesh.dbc_dof, mesh.dbc_value)
    ctx.prepare_elements()
    ctx.build_pattern()
    ctx.set_params(np.array([1.0, float(np.mean(mu)), 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0]), 2, 0.0, 0.1)
    ctx.set_batch(em, mu)
    if a.c5_coarse:
        ctx.set_coarse_correction(a.c5_coarse, True)
    ctx.setup_solver(opts)
    ctx.batch_use_graph(True)
    nrec = a.c5_train_steps - a.c5_record_from + 1
    learner = sfem.Learner(ctx, max(nrec, 4), seed=a.c5_seed)
    log(f"training data: {len(meshes)} meshes in one batch ({mesh.nNode} nodes), store of {nrec} states "
        f"({nrec * mesh.nNode * 12 / 2**30:.1f} GiB)")
    t, kry, nwt = 0.1, 0, 0
    t0 = t1 = time.perf_counter()
    for k in range(a.c5_train_steps):
        ni, ki, rn = ctx.step_batch(min(t, 1.0), opts, a.newton_tol, 10)
        kry += int(ki.max())
        nwt += int(ni.max())
        if k + 1 >= a.c5_record_from:
            learner.record(ctx, new_trajectory=(k + 1 == a.c5_record_from))
        if (k + 1) % 50 == 0 or k + 1 == a.c5_train_steps:
            t2 = time.perf_counter()
            n = (k + 1) % 50 or 50
            log(f"training data: steps {k + 2 - n}-{k + 1} of {a.c5_train_steps}: {n / (t2 - t1):.2f} batch-steps/s, "
                f"Newton {nwt / n:.2f} and Krylov {kry / n:.1f} per step (largest of the meshes), max |F| {rn.max():.2e}")
            t1, kry, nwt = t2, 0, 0
        t += 0.1
    rate = a.c5_train_steps / (time.perf_counter() - t0)
    ctx.close()
    return learner, rate


def main_c5(a):
    """Learned initial iterate on the cylinder-in-channel Quad4 mesh: trajectories at
    several Reynolds numbers with seeded 1% inlet perturbations train the per-node MLP,
    then a run at Re = 100 with an unseen perturbation is timed from the extrapolated
    start and from the learned start (judged per step by the iteration-0 residual)."""
    from cfdsfemmpi_amd import meshgen, sfem
    steps = a.steps if a.steps is not None else 200
    if ":" in a.c5_train_re:  # lo:hi:count, evenly spaced
        lo, hi, cnt = a.c5_train_re.split(":")
        res = [float(v) for v in np.linspace(float(lo), float(hi), int(cnt))]
    else:
        res = [float(r) for r in a.c5_train_re.split(",")]
    opts = sfem.solver_opts("bcgs", "ilu", 1e-12, 20000, reorder=a.reorder, atol=0.5 * a.newton_tol)
    meshes = [meshgen.cylinder_channel(perturb=0.01, seed=1000 + k) for k in range(len(res))]
    m0 = meshes[0]
    log(f"workload: cylinder in channel, {m0.nElem} Quad4, {m0.nNode} nodes; training Re {res}, "
        f"{a.c5_train_steps} steps each; evaluation Re 100, {steps} steps")
    learner = None
    train_rates = []
    t_gen = time.perf_counter()
    if a.c5_train_batch:
        learner, gen = c5_train_batch(meshes, res, a)
        train_rates.append(gen)
    else:
        for k, (mesh, re) in enumerate(zip(meshes, res)):
            fem = c5_context(mesh, 1.0 / re, opts, a.c5_coarse)
            if learner is None:
                learner = sfem.Learner(fem.ctx, len(res) * a.c5_train_steps, seed=a.c5_seed)
            r = c5_run(fem.ctx, opts, a.c5_train_steps, a.newton_tol, record=learner, record_from=a.c5_record_from,
                       label=f"train Re {re:g}")
            train_rates.append(r["steps_per_s"])
            fem.ctx.close()
    t_gen = time.perf_counter() - t_gen
    # one network per component setting (--c5-components), each from the same initial
    # weights and trained on the same store; four equal phases, the learning rate halved
    # after each (Adam's moments carry over)
    comps = list(dict.fromkeys(int(c) for c in a.c5_components.split(",")))
    w_init = learner.get_weights()
    chunk = max(1, a.c5_train_iters // 4)
    weights, losses, t_fit = {}, {}, 0.0
    for j, nc in enumerate(comps):
        if j:
            learner.set_weights(w_init)
        learner.set_components(nc)  # a change of components takes the output scales again
        t0 = time.perf_counter()
        loss = np.concatenate([learner.train(chunk, a.c5_batch, a.c5_lr * 0.5 ** k) for k in range(4)])
        t_fit += time.perf_counter() - t0
        log(f"training ({nc} components): {len(loss)} Adam steps x {a.c5_batch} samples in "
            f"{time.perf_counter() - t0:.2f} s, loss {loss[:50].mean():.4e} -> {loss[-50:].mean():.4e}")
        weights[nc], losses[nc] = learner.get_weights(), loss
    # the trained networks move to a learner of the evaluation mesh (weights and output
    # scales); the training store is released
    windows = learner.info()[1]
    learner.close()
    evalmesh = meshgen.cylinder_channel(perturb=0.01, seed=99)
    runs = {}
    for kind in ["extrapolated", "quadratic"] + [f"learned{nc}" for nc in comps]:
        fem = c5_context(evalmesh, 0.01, opts, a.c5_coarse)
        if kind == "quadratic":
            fem.ctx.set_predictor(sfem.PRED_QUADRATIC)
        ev = None
        if kind.startswith("learned"):
            nc = int(kind[7:])
            ev = sfem.Learner(fem.ctx, 4, seed=a.c5_seed)
            ev.set_components(nc)
            ev.set_weights(weights[nc])
            fem.ctx.set_learner(ev)
            fem.ctx.set_predictor(sfem.PRED_LEARNED)
        r = c5_run(fem.ctx, opts, steps, a.newton_tol, label=f"eval Re 100 {kind}", judged=ev is not None)
        r["soln"] = fem.ctx.get_solution()
        if ev is not None:
            r["starts"] = [int(v) for v in fem.ctx.predictor_stats()[0]]
            log(f"starts taken (linear, quadratic, learned): {r['starts']}")
        fem.ctx.close()
        if ev is not None:
            ev.close()
        runs[kind] = r
    ex, qu, le = runs["extrapolated"], runs["quadratic"], runs[f"learned{comps[0]}"]
    loss = losses[comps[0]]
    diff = float(np.max(np.abs(le["soln"] - ex["soln"])) / np.max(np.abs(ex["soln"])))
    skip = min(a.c5_record_from, steps // 2)  # after the inflow ramp and its transient
    out = {
        "metric": C5_METRIC,
        "value": le["steps_per_s"],
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * le["seconds"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 solve; MLP bf16 MFMA inference, f32 training",
        "data": "synthetic (generated mesh, seeded inlet perturbations, zero initial state)",
        "config": {"workload": f"2D cylinder D=1 in channel [-10,30]x[-10,10], O-grid 64x256 ratio 1.05 + "
                               f"structured blocks, {m0.nElem} Quad4, inflow u=1 (+1% seeded perturbation), "
                               f"slip walls, traction-free outflow, Re=100, dt=0.1, {steps} steps from rest",
                   "training": {"trajectories": len(res), "re": res, "steps_each": a.c5_train_steps,
                                "recorded_from_step": a.c5_record_from, "windows": windows,
                                "adam_steps": len(loss), "batch": a.c5_batch,
                                "lr": [a.c5_lr * 0.5 ** k for k in range(4)],
                                "seconds_generating": t_gen, "seconds_fitting": t_fit,
                                "generation": (f"one batched context, BiCGStab rtol {a.c5_train_rtol:g}, "
                                               f"{train_rates[0]:.2f} batch-steps/s") if a.c5_train_batch
                                else "one context per trajectory, the evaluation solver",
                                "loss_first50": float(loss[:50].mean()), "loss_last50": float(loss[-50:].mean())},
                   "solver": f"FP64 BiCGStab + ILU(0) (ordering {a.reorder})"
                             + (f" + multiplicative coarse pressure correction ({a.c5_coarse}^2 boxes)" if a.c5_coarse else "")
                             + f", stop at ||r|| <= max(1e-12 ||b||, "
                             f"{0.5 * a.newton_tol:g}); Newton tol {a.newton_tol:g}",
                   "parallelism": "single GPU"},
        "extrapolated": {"steps_per_s": ex["steps_per_s"], "newton_total": sum(ex["newton"]),
                         "krylov_total": sum(ex["krylov"]), "krylov_after_ramp": sum(ex["krylov"][skip:]),
                         "krylov_per_step": ex["krylov"], "r0_per_step": ex["r0"]},
        "quadratic": {"steps_per_s": qu["steps_per_s"], "newton_total": sum(qu["newton"]),
                      "krylov_total": sum(qu["krylov"]), "krylov_after_ramp": sum(qu["krylov"][skip:]),
                      "krylov_per_step": qu["krylov"], "r0_per_step": qu["r0"]},
        "learned": {"components": comps[0], "steps_per_s": le["steps_per_s"], "newton_total": sum(le["newton"]),
                    "krylov_total": sum(le["krylov"]), "krylov_after_ramp": sum(le["krylov"][skip:]),
                    "krylov_per_step": le["krylov"], "r0_per_step": le["r0"],
                    "starts_linear_quadratic_learned": le["starts"],
                    "candidate_norms_per_step_linear_quadratic_learned": le["candidates"]},
        "learned_other_components": {
            str(nc): {"steps_per_s": runs[f"learned{nc}"]["steps_per_s"],
                      "newton_total": sum(runs[f"learned{nc}"]["newton"]),
                      "krylov_total": sum(runs[f"learned{nc}"]["krylov"]),
                      "krylov_after_ramp": sum(runs[f"learned{nc}"]["krylov"][skip:]),
                      "krylov_per_step": runs[f"learned{nc}"]["krylov"],
                      "starts_linear_quadratic_learned": runs[f"learned{nc}"]["starts"],
                      "loss_first50": float(losses[nc][:50].mean()), "loss_last50": float(losses[nc][-50:].mean())}
            for nc in comps[1:]},
        "krylov_ratio_learned_vs_extrapolated": sum(le["krylov"]) / max(1, sum(ex["krylov"])),
        "krylov_ratio_learned_vs_quadratic": sum(le["krylov"]) / max(1, sum(qu["krylov"])),
        "steps_per_s_ratio_learned_vs_extrapolated": le["steps_per_s"] / ex["steps_per_s"],
        "max_rel_diff_final_state_learned_vs_extrapolated": diff,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def copy_probe(mib):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE readings give the
    factor between those counters and bytes moved."""
    import torch
    nbytes = mib << 20
    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy probe: 5 copies of {nbytes} bytes ({mib} MiB) done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (10; 20 for c4)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c3", "c2", "c4", "c5"],
                    help="c3: 115^3 hex cavity (the headline); c2: 32^3; c4: batch of 256 Tria3 cavities")
    ap.add_argument("--c5-train-re", default="60:200:12",
                    help="c5: Reynolds numbers of the training runs (a list, or lo:hi:count evenly spaced)")
    ap.add_argument("--c5-train-steps", type=int, default=200, help="c5: steps of each training run")
    ap.add_argument("--c5-record-from", type=int, default=13,
                    help="c5: first recorded step of a training run (the inflow ramp ends at step 10)")
    ap.add_argument("--c5-train-iters", type=int, default=8000, help="c5: Adam steps (four phases, lr halved)")
    ap.add_argument("--c5-batch", type=int, default=8192, help="c5: samples per Adam step")
    ap.add_argument("--c5-lr", type=float, default=1e-3)
    ap.add_argument("--c5-seed", type=int, default=0)
    ap.add_argument("--c5-train-batch", type=int, default=1,
                    help="c5: generate the training trajectories in one batched context (0: one after another)")
    ap.add_argument("--c5-train-rtol", type=float, default=1e-6,
                    help="c5: BiCGStab relative tolerance of the batched training runs")
    ap.add_argument("--c5-components", default="3,2",
                    help="c5: components the network corrects, 2 (u, v) or 3 (u, v, p); a list trains one each")
    ap.add_argument("--c5-coarse", type=int, default=16,
                    help="c5: boxes per direction of the multiplicative coarse pressure correction (0: off)")
    ap.add_argument("--c4-count", type=int, default=256)
    ap.add_argument("--c4-coarse-mult", type=int, default=1,
                    help="c4: multiplicative (1) or additive (0) form of the coarse correction")
    ap.add_argument("--c4-graph", type=int, default=1,
                    help="c4: Newton steps as HIP graphs (0: every kernel launched one by one)")
    ap.add_argument("--c4-seed", type=int, default=0)
    ap.add_argument("--c4-seq-stride", type=int, default=16,
                    help="c4: time every n-th mesh through the single-mesh path (0: skip)")
    ap.add_argument("--c4-cpu-stride", type=int, default=32, help="c4: oracle sample of every n-th mesh")
    ap.add_argument("--c4-coarse", type=int, default=8,
                    help="c4: multiplicative coarse pressure correction on n^2 boxes per mesh (0: off)")
    ap.add_argument("--c4-mixed", type=int, default=0,
                    help="c4: also run the batch with FP32 Krylov (stalls on these systems; slow)")
    ap.add_argument("--c4-store32", type=int, default=1,
                    help="c4: also run the batch with FP64 Krylov over the FP32-stored tangent and ILU(0)")
    ap.add_argument("--c4-store32-rtol", type=float, default=1e-4, help="c4: inner rtol of the FP32-stored run")
    ap.add_argument("--c4-store32-newton", type=float, default=1e-8, help="c4: Newton tol of the FP32-stored run")
    ap.add_argument("--n", type=int, default=0, help="override the mesh size n (n^3 hexes)")
    ap.add_argument("--inner-rtol", type=float, default=4e-5,
                    help="mixed path: the deepest reduction asked of one FP32 inner solve (with --inner-theta, "
                         "a solve is asked for no more than the Newton test still needs)")
    ap.add_argument("--newton-tol", type=float, default=1e-8,
                    help="FP64 path: absolute Newton tolerance on ||F|| (the reference's 1e-8)")
    ap.add_argument("--newton-tol-mixed", type=float, default=1e-9,
                    help="mixed path: absolute Newton tolerance on the FP64 residual")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-c2", action="store_true", help="skip the CPU baseline's 20-step 32^3 run")
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--reorder", type=int, default=None,
                    help="ILU(0) node ordering: 0 natural, 1 RCM, 2 multicolour, 3 tiled multicolour "
                         "(default: 3 for the hex cavities, 2 for the c4 batch)")
    ap.add_argument("--bf16-factors", type=int, default=1,
                    help="mixed path: ILU(0) factors and tile inverses stored in bfloat16 (0: FP32)")
    ap.add_argument("--predictor", choices=["previous", "extrapolate"], default="previous",
                    help="Newton initial iterate of each step (both paths)")
    ap.add_argument("--inner-theta", type=float, default=0.8,
                    help="mixed path: inexact Newton, inner rtol = min(inner-rtol-max, max(inner-rtol, "
                         "theta * newton_tol / ||F||)) per Newton iteration (0: inner-rtol throughout)")
    ap.add_argument("--inner-rtol-max", type=float, default=0.5, help="mixed path: cap of the inexact-Newton rtol")
    ap.add_argument("--tuning", choices=["timed", "fixed"], default="timed",
                    help="kernel shapes: timed autotuner (default) or fixed default shapes (reproducible FP32)")
    ap.add_argument("--pc-lag", type=int, default=6,
                    help="both paths: rebuild the ILU(0) factor on every n-th new tangent only (1: every tangent, "
                         "as the PETSc KSP of the reference does)")
    ap.add_argument("--ksp", choices=["bcgs", "gmres"], default="bcgs", help="Krylov method of both paths")
    ap.add_argument("--coarse", type=int, default=8,
                    help="fp64 and mixed paths: additive coarse pressure correction of ILU(0) on n^3 boxes (0: off)")
    ap.add_argument("--coarse-mult", type=int, default=1,
                    help="the coarse correction's multiplicative form (0: additive)")
    ap.add_argument("--plain", type=int, default=1,
                    help="also time the FP64 path with the reference's solver configuration (ILU(0) alone, "
                         "rebuilt for every tangent)")
    ap.add_argument("--restart", type=int, default=30, help="GMRES restart length")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): this build is the single-GPU path; "
              f"a multi-GPU layout is not built", file=sys.stderr)
        sys.exit(2)
    from cfdsfemmpi_amd import sfem
    sfem.load_library()
    if a.reorder is None:
        a.reorder = 2 if a.workload == "c4" else 3
    if a.workload == "c4":
        main_c4(a)
        return
    if a.workload == "c5":
        main_c5(a)
        return
    if a.steps is None:
        a.steps = 10
    n = a.n or (115 if a.workload == "c3" else 32)
    mesh, ctl = make_case(n)
    log(f"workload {n}^3 hexes: {mesh.nElem} elements, {mesh.nNode} nodes")
    o64 = sfem.solver_opts(a.ksp, "ilu", 1e-8, 5000, a.restart, reorder=a.reorder)
    omx = sfem.solver_opts(a.ksp, "ilu", a.inner_rtol, 5000, a.restart, reorder=a.reorder,
                           factor_bf16=a.bf16_factors)
    pred = sfem.PRED_EXTRAPOLATE if a.predictor == "extrapolate" else sfem.PRED_PREVIOUS
    tune_mode = sfem.TUNE_FIXED if a.tuning == "fixed" else sfem.TUNE_TIMED
    r64 = rmx = rpl = None
    if a.only is None and a.plain:
        # the reference's solver configuration: BiCGStab + ILU(0) alone, the factor rebuilt
        # for every new tangent (PCSetUp after each assembly)
        rpl = run_path(mesh, ctl, sfem.PREC_FP64, o64, a.newton_tol, a.steps, a.warmup, "fp64_plain", pred, 1,
                       tuning=tune_mode, coarse=0)
        log(f"fp64_plain: {rpl['steps_per_s']:.3f} steps/s")
    if a.only != "mixed":
        r64 = run_path(mesh, ctl, sfem.PREC_FP64, o64, a.newton_tol, a.steps, a.warmup, "fp64", pred, a.pc_lag,
                       tuning=tune_mode, coarse=a.coarse, coarse_mult=bool(a.coarse_mult))
        log(f"fp64: {r64['steps_per_s']:.3f} steps/s")
    if a.only != "fp64":
        rmx = run_path(mesh, ctl, sfem.PREC_MIXED, omx, a.newton_tol_mixed, a.steps, a.warmup, "mixed", pred,
                       a.pc_lag, (a.inner_theta, a.inner_rtol_max), tune_mode, a.coarse, bool(a.coarse_mult))
        log(f"mixed: {rmx['steps_per_s']:.3f} steps/s")
    main_res = rmx or r64
    err_u = err_p = None
    if r64 and rmx:
        s64 = r64["soln"].reshape(-1, 4)
        smx = rmx["soln"].reshape(-1, 4)
        err_u = float(np.linalg.norm(smx[:, :3] - s64[:, :3]) / np.linalg.norm(s64[:, :3]))
        err_p = float(np.linalg.norm(smx[:, 3] - s64[:, 3]) / np.linalg.norm(s64[:, 3]))
        log(f"rel L2 error mixed vs fp64: u {err_u:.3e} p {err_p:.3e}")
    err_u_pl = err_p_pl = None
    if rpl and rmx:
        spl = rpl["soln"].reshape(-1, 4)
        smx = rmx["soln"].reshape(-1, 4)
        err_u_pl = float(np.linalg.norm(smx[:, :3] - spl[:, :3]) / np.linalg.norm(spl[:, :3]))
        err_p_pl = float(np.linalg.norm(smx[:, 3] - spl[:, 3]) / np.linalg.norm(spl[:, 3]))
        log(f"rel L2 error mixed vs plain fp64: u {err_u_pl:.3e} p {err_p_pl:.3e}")
    # Roofline of the dominant single kernel of the reported run.  The SpMV is one launch
    # per product; the ILU(0) apply is a group of 2 x (colours) launches, none of which
    # outweighs the SpMV, and the element loop is one launch per colour class.
    kt = main_res["kt"]
    info = main_res["info"]
    nnz, nf = info["nnz"], info["nfree"]
    fp32 = main_res is rmx
    sm = sv = 4 if fp32 else 8   # matrix values and Krylov vectors

    def hbm_roof(kind, name, bytes_per):
        cnt, ms = kt[kind]
        avg = ms / max(cnt, 1)
        r = {"bound": "hbm", "kernel": name, "achieved": bytes_per / (avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "avg_launch_ms": avg, "launches": cnt, "algorithmic_bytes_per_launch": bytes_per}
        r["frac"] = r["achieved"] / r["peak"]
        return r

    # Algorithmic bytes of the node-block SpMV: every nonzero value once (padding of the
    # 4-wide blocks not counted), one int32 index per neighbour block, one 16-byte
    # descriptor per block row, x once and y once.  The CSR figure nnz*(s+4) + 4(n+1)
    # + 2 s n is reported beside it.
    bi = main_res["block"]
    spmv_bytes = nnz * sm + 4 * bi["npairs"] + 16 * bi["nblk"] + 2 * sv * nf
    roof = hbm_roof("spmv", f"k_bsr_spmv<{'float' if fp32 else 'double'}> (node-block SpMV)", spmv_bytes)
    roof["csr_equivalent_bytes_per_launch"] = nnz * (sm + 4) + 4 * (nf + 1) + 2 * sv * nf
    roof["traffic"] = MEASURED_TRAFFIC.get(("k_bsr_spmv", "float" if fp32 else "double", n))
    roof["traffic_source"] = TRAFFIC_SOURCE if roof["traffic"] else None
    # ILU(0) apply: both sweeps together read every factor value, block index and
    # descriptor once; r read, y written, y read and written, plus the gathered y once
    # per sweep (5 vector passes).  With the tiled ordering the values inside the tiles are
    # replaced by the tiles' packed inverse triangles (2 x 528 values per tile, stored in
    # the factor's type).
    lu_s = 2 if (fp32 and a.bf16_factors) else sm
    ti = main_res["tiles"]
    pc_bytes = nnz * lu_s + 4 * bi["npairs"] + 32 * bi["nblk"] + 5 * sv * nf
    pc_name = "ILU(0) apply (2 x colours launches)"
    if ti["ntiles"]:
        inv_s = lu_s  # the tile inverses are stored like the factor
        pc_bytes += ti["ntiles"] * 2 * 528 * inv_s - ti["tile_vals"] * lu_s
        pc_name = (f"ILU(0) apply, tiled: k_bsr_tile_lsolve/usolve, {2 * ti['ncolors'] - 1} launches, "
                   f"{ti['ntiles']} tiles")
    co = main_res["coarse"]
    if co["built"]:
        # coarse correction: r read, z read and written at the pressure slots, their slot
        # indices read twice
        pc_bytes += co["pressure_dofs"] * (3 * sv + 8)
        if a.coarse_mult:
            # multiplicative: the product's pressure slots read once more, and A Z y added to
            # the product (its index and value entries, and the product, read and written)
            pc_bytes += co["pressure_dofs"] * sv + 2 * sv * nf
        pc_name += (f", + {'multiplicative' if a.coarse_mult else 'additive'} coarse pressure correction "
                    f"({co['unknowns']} unknowns, {4 if a.coarse_mult else 3} launches)")
    roof_pc = hbm_roof("pc", pc_name, pc_bytes)
    roof_pc["tiles"] = ti
    n_k = len(main_res["hist"])
    n_f = sum(h[0] for h in main_res["hist"]) - n_k
    cnt, ms = kt["assembly"]
    flops = mesh.nElem * (n_k * HEXA8_FLOPS_KF + n_f * HEXA8_FLOPS_F)
    # one tangent + residual pass: the tangent arithmetic (Ke) runs in the tangent's type
    # (FP32 on the mixed path, against the FP32 peak), the residual (Fe) in FP64; the peak
    # quoted is the rate at which that mix would run with both units at their peaks.  The
    # published FP32 peak counts packed FP32 instructions, which this build does not emit
    # (its device code is compiled without packed FP32), so the unpacked FP32 rate is the
    # FP64 rate; frac_unpacked is the fraction of that.
    ap_ = main_res["asm_pass"]
    fl_k, fl_f = mesh.nElem * (HEXA8_FLOPS_KF - HEXA8_FLOPS_F), mesh.nElem * HEXA8_FLOPS_F
    t_bound = fl_k / ((FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS) * 1e12) + fl_f / (FP64_PEAK_TFLOPS * 1e12)
    roof_asm = {"bound": "vector", "kernel": "k_hex8_batch<%d> (one tangent + residual pass, %d colour launches)"
                % (2 if fp32 else 0, info["ncolors"]),
                "achieved": (fl_k + fl_f) / (ap_["K_F_ms"] * 1e-3) / 1e12,
                "peak": (fl_k + fl_f) / t_bound / 1e12, "unit": "TFLOP/s",
                "tangent_type": "fp32" if fp32 else "fp64", "pass_ms": ap_["K_F_ms"],
                "flops_per_element": HEXA8_FLOPS_KF,
                "residual_only_pass": {"kernel": "k_hex8_batch<1>", "pass_ms": ap_["F_ms"],
                                       "achieved": fl_f / (ap_["F_ms"] * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                                       "flops_per_element": HEXA8_FLOPS_F},
                "run": {"passes_K_F": n_k, "passes_F": n_f, "achieved": flops / (ms * 1e-3) / 1e12}}
    roof_asm["frac"] = roof_asm["achieved"] / roof_asm["peak"]
    roof_asm["frac_unpacked"] = roof_asm["achieved"] / FP64_PEAK_TFLOPS
    roof_asm["residual_only_pass"]["frac"] = roof_asm["residual_only_pass"]["achieved"] / FP64_PEAK_TFLOPS
    cpu = None
    if not a.no_cpu and r64 is not None:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(n, c2=not a.no_cpu_c2 and n != 32)
    out = {
        "metric": METRIC,
        "value": main_res["steps_per_s"],
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * main_res["seconds"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 tangent+Krylov" + (" (bf16 ILU factors)" if a.bf16_factors else "") +
                  (", fp64 coarse step" if a.coarse else "") + " / fp64 residual") if fp32 else "f64",
        "data": "synthetic (generated mesh, zero initial state)",
        "config": {"workload": f"3D lid-driven cavity, {n}^3 graded Hexa8, Re=1000, gen-alpha rho_inf=0, "
                               f"dt=0.1, {a.steps} steps", "n_elements": mesh.nElem, "nfree": nf, "nnz": nnz,
                   "solver": ("BiCGStab" if a.ksp == "bcgs" else f"GMRES({a.restart})") + " + ILU(0)" +
                             (f" + coarse pressure correction ({a.coarse}^3 boxes)" if a.coarse else ""),
                   "ilu_ordering": ["natural", "rcm", "multicolour", "tiled multicolour"][a.reorder],
                   "newton_tol": a.newton_tol, "newton_tol_mixed": a.newton_tol_mixed,
                   "inner_rtol_mixed": a.inner_rtol, "ilu_factors_mixed": "bf16" if a.bf16_factors else "fp32",
                   "newton_initial_iterate": a.predictor, "pc_lag": a.pc_lag,
                   "inner_theta_mixed": a.inner_theta, "inner_rtol_max_mixed": a.inner_rtol_max,
                   "kernel_shapes": a.tuning, "coarse_pressure_boxes": a.coarse,
                   "coarse_form": "multiplicative" if a.coarse_mult else "additive",
                   "coarse_unknowns": main_res["coarse"]["unknowns"], "parallelism": "single GPU"},
        # fp64: the FP64 path with the same solver options as the mixed run (ILU(0), coarse
        # pressure correction, preconditioner lag); fp64_plain: the reference's solver
        # configuration (ILU(0) alone, rebuilt for every tangent)
        "fp64_steps_per_s": r64["steps_per_s"] if r64 else None,
        "fp64_plain_steps_per_s": rpl["steps_per_s"] if rpl else None,
        "mixed_steps_per_s": rmx["steps_per_s"] if rmx else None,
        "speedup_mixed_vs_fp64": (rmx["steps_per_s"] / r64["steps_per_s"]) if (r64 and rmx) else None,
        "speedup_mixed_vs_fp64_plain": (rmx["steps_per_s"] / rpl["steps_per_s"]) if (rpl and rmx) else None,
        # the gate is on velocity (rel. L2 of u vs the FP64 run); pressure is reported beside it
        "error_gate_1e-6_passed": (err_u <= 1e-6) if err_u is not None else None,
        "rel_l2_u_mixed_vs_fp64": err_u,
        "rel_l2_p_mixed_vs_fp64": err_p,
        "rel_l2_u_mixed_vs_fp64_plain": err_u_pl,
        "rel_l2_p_mixed_vs_fp64_plain": err_p_pl,
        "krylov_iters": {r["label"]: [h[1] for h in r["hist"]] for r in (rpl, r64, rmx) if r},
        "newton_iters": {r["label"]: [h[0] for h in r["hist"]] for r in (rpl, r64, rmx) if r},
        "kernel_ms": {r["label"]: {k: {"launches": v[0], "total_ms": v[1]} for k, v in r["kt"].items()}
                      for r in (rpl, r64, rmx) if r},
        "autotuned": {r["label"]: r["tune"] for r in (rpl, r64, rmx) if r},
        "fp32_gmres_fallbacks": rmx["stats"][0] if rmx else None,
        "fp32_partial_solves": rmx["stats"][1] if rmx else None,
        "block_form": main_res["block"],
        "roofline": roof,
        "roofline_ilu_apply": roof_pc,
        "roofline_assembly": roof_asm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
