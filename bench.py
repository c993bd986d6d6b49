#!/usr/bin/env python3
"""Benchmark of the fluid_mc element assembly (R + J) on one MI355X.

A step is one fluid_assembly::execute() (sim/fluid_mc/P0/fluid_assembly.hpp:39-68)
over the 2D jittered channel of 2048x1024 quads (4,194,304 linear triangles,
2,100,225 nodes, 2 components, 5 dofs/node), plain FP64: every interior and
boundary element, Dirichlet rule and deterministic scatter into the device BSR
matrix and rhs.  State and mesh are resident in HBM before timing starts.

Prints ONE JSON line with the headline value (element assemblies/s), the
roofline of the dominant kernel (the interior element kernel, HIP-event timed
by the library on its own stream around that kernel alone), the FP32 tier on
the same workload and a CPU baseline (the FP64 oracle on the host cores,
bounded sample).  roofline.traffic is the kernel's measured HBM bytes per
launch from the FETCH_SIZE / WRITE_SIZE counter runs committed in
profiles/k_patch_channel2d_4m_memory.md (null if that file is absent); roofline.compute
is the same kernel against the FP64 vector peak from the instruction counters
in profiles/k_patch_channel2d_4m_valu.md.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "gales-multi-fluid_amd"), ROOT]

import numpy as np  # noqa: E402

from gales_mc import abi, cases  # noqa: E402
from gales_mc.lib import Context  # noqa: E402

# the headline metric's name as the benchmark definition states it; what the
# FP32 tier computes on is config.tier_fp32_element_matrix
METRIC = "element assemblies/s (R+J) + HBM GB/s; FP32-MFMA vs FP64 at \u22641e-10 Newton resid"
TIER_MATRIX_UNIT = "VALU (FP32 FMA)"
TIER_MATRIX_UNIT_MFMA = "least squares on MFMA (v_mfma_f32_16x16x4f32), other terms VALU (FP32 FMA)"
HBM_PEAK_GBS = 8000.0
REFINE_PASSES = 4
COUNTERS = os.path.join(ROOT, "profiles", "k_patch_channel2d_4m_memory.md")
VALU_COUNTERS = os.path.join(ROOT, "profiles", "k_patch_channel2d_4m_valu.md")
FP64_PEAK_TFLOPS = 78.6


def measured_traffic(precision):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary (line 'traffic_bytes_per_launch <precision>: <bytes>')."""
    if not os.path.exists(COUNTERS):
        return None
    key = f"traffic_bytes_per_launch {precision}:"
    for line in open(COUNTERS):
        if line.strip().startswith(key):
            return float(line.split(":", 1)[1])
    return None


def measured_valu(precision, kernel_s):
    """The dominant kernel's arithmetic against the FP64 vector peak, from the
    committed instruction-counter summary (lines 'valu_per_launch <precision>
    <name>: <value>'): issued VALU instructions, FP64 FMA instructions and MFMA
    instructions per launch.  flops = 2 x 64 lanes x FMA instructions (the
    FMAs alone; adds and multiplies not counted).  None if the file is absent."""
    if not os.path.exists(VALU_COUNTERS):
        return None
    vals = {}
    for line in open(VALU_COUNTERS):
        line = line.strip()
        if line.startswith(f"valu_per_launch {precision} "):
            name, v = line[len(f"valu_per_launch {precision} "):].split(":", 1)
            vals[name.strip()] = float(v)
    if "fma_f64" not in vals:
        return None
    # FP64 and unpacked FP32 FMAs issue at the same rate per lane (packed FP32
    # is not used), so both count against the 78.6 TFLOP/s vector peak
    tflops = 2 * 64 * (vals["fma_f64"] + vals.get("fma_f32", 0.0)) / kernel_s / 1e12
    return {"bound": "valu", "fma_tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS,
            "frac": tflops / FP64_PEAK_TFLOPS, "valu_busy": vals.get("valu_busy"),
            "mfma_insts": vals.get("mfma"), "source": os.path.relpath(VALU_COUNTERS, ROOT)}


def algorithmic_bytes(mesh, info):
    """Compulsory HBM bytes of one assembly: connectivity (int32), coordinates,
    I and P read, rhs written, J values written (FP64)."""
    E = info.n_elems
    N = info.n_nodes
    return (E * info.nen * 4 + N * info.dim * 8 + 2 * N * info.ndof * 8 + N * info.ndof * 8
            + info.nnzb * info.ndof * info.ndof * 8)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_cflags():
    """CC and CFLAGS of the oracle library (oracle/Makefile)."""
    cc, flags = "gcc", ""
    for line in open(os.path.join(ROOT, "oracle", "Makefile")):
        if line.startswith("CC ?="):
            cc = line.split("=", 1)[1].strip()
        if line.startswith("CFLAGS ?="):
            flags = line.split("=", 1)[1].strip()
    return f"{cc} {flags}"


# host cores the baseline may use: the CPU share of one GPU on the
# measurement host is 16 cores, whatever the machine's own count
CPU_SHARE = 16


def cpu_baseline(threads=None, scale=4, reps=5):
    """The FP64 oracle's assembly (element partition per thread into a private
    BSR, then a reduction -- the per-rank CSR + GlobalAssemble structure of the
    reference) on a bounded sample: the c1 mesh family at 1/scale resolution,
    same physics and per-element work.  Median of reps timed runs after one
    untimed run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    mesh, P, I, Pst = cases.config(1, scale=scale)
    if threads is None:
        threads = min(len(os.sched_getaffinity(0)), CPU_SHARE)
    A = orc.Assembler(P, mesh)
    E = mesh["conn"].shape[0]
    bounds = np.linspace(0, E, threads + 1).astype(np.int64)
    parts = [A.new_system() for _ in range(threads)]

    def run():
        for v, r in parts:
            v[...] = 0.0
            r[...] = 0.0
        ths = [threading.Thread(target=A.assemble, args=(I, Pst, int(bounds[t]), int(bounds[t + 1]),
                                                           parts[t][0], parts[t][1]))
               for t in range(threads)]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        vals, rhs = parts[0]
        for v, r in parts[1:]:
            vals += v
            rhs += r
        return time.perf_counter() - t0

    run()
    ts = sorted(run() for _ in range(reps))
    t = ts[len(ts) // 2]
    # one core: the first 1/16 of the same elements on one thread
    e1 = E // 16
    v1, r1 = A.new_system()
    t0 = time.perf_counter()
    A.assemble(I, Pst, 0, e1, v1, r1)
    t1 = time.perf_counter() - t0
    return {"value": E / t, "unit": "elem/s", "cores": threads, "kind": "port",
            "value_1core": e1 / t1,
            "sample": f"FP64 oracle assembly of the 2D channel at {mesh['shape'][0]}x{mesh['shape'][1]} "
                      f"quads ({E} triangles, same physics as the timed workload), "
                      f"{threads} threads with private BSR + reduction, median of {reps} "
                      f"(min {E / ts[-1]:.3e}, max {E / ts[0]:.3e} elem/s); "
                      f"value_1core: {e1} of those elements on one thread; "
                      f"CPU {cpu_model()}, {len(os.sched_getaffinity(0))} cores visible, "
                      f"{CPU_SHARE} the share of one GPU; oracle built with {oracle_cflags()}"}


def tier_eval(cid, scale, max_iters, passes=3, inner_tol=1e-6, fp64_solve=True, precond=2, restart=100):
    """Errors and times of the precision tier against the plain FP64 path on
    one state: rhs (FP64 in both), J x (FP32-element matrix vs FP64), the
    matrix-free FP64 product vs the assembled FP64 matrix, the Newton residual
    ||rhs - J64 x|| / ||rhs|| after each refinement pass (inner solves:
    GMRES(100) with block Jacobi (precond 2) or block ILU(0) (3) on the
    FP32-tier matrix to inner_tol, at most max_iters iterations), and the
    update's difference from the FP64 path's GMRES solution (skipped without
    fp64_solve)."""
    mesh, P, I, Pst = cases.config(cid, scale=scale)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    f64 = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS
    f32 = f64 | abi.GMC_ASM_FP32

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    out = {"config": cid, "scale": scale, "elements": int(mesh["conn"].shape[0]), "dofs": int(I.size)}
    ctx.assemble(f64)
    t0 = time.perf_counter()
    ctx.assemble(f64)
    out["assemble_fp64_ms"] = (time.perf_counter() - t0) * 1e3
    _, rhs64 = ctx.system()
    x = np.random.default_rng(3).standard_normal(I.size)
    y64 = ctx.spmv(x)
    d64 = None
    if fp64_solve:
        t0 = time.perf_counter()
        s64 = ctx.solve(restart=restart, max_iters=max_iters, rel_tol=1e-10, precond=precond)
        out["solve_fp64_s"] = time.perf_counter() - t0
        out["solve_fp64_iterations"] = s64.iterations
        out["solve_fp64_rel_residual"] = s64.rel_residual
        d64 = ctx.solution()
    ctx.assemble(f32)
    t0 = time.perf_counter()
    ctx.assemble(f32)
    out["assemble_fp32_ms"] = (time.perf_counter() - t0) * 1e3
    _, rhs32 = ctx.system()
    out["rhs_rel_err"] = rel(rhs32, rhs64)
    out["Jx_rel_err_fp32_matrix"] = rel(ctx.spmv(x), y64)
    out["Jx_rel_err_matrix_free_fp64"] = rel(ctx.jvp(x), y64)
    t0 = time.perf_counter()
    res = ctx.refine(n_refine=passes, restart=restart, max_iters=max_iters, rel_tol=inner_tol, precond=precond)
    out["refine_s"] = time.perf_counter() - t0
    out["refine_passes"] = passes
    out["refine_precond"] = {1: "point Jacobi", 2: "block Jacobi", 3: "block ILU(0)"}[precond]
    out["refine_inner_rel_tol"] = inner_tol
    out["newton_rel_residual_per_pass"] = [float(r / res[0]) for r in res]
    out["newton_rel_residual"] = float(res[-1] / res[0])
    if d64 is not None:
        out["update_rel_diff_vs_fp64_path"] = rel(ctx.solution(), d64)
    ctx.close()
    return out


def vjp_eval(cid, steps):
    """Times of the exact VJP (v^T dR/dI, v^T dR/dP, v^T dR/dtheta) and the
    frozen-coefficient J^T v on device vectors, and of an FP64 assembly on the
    same state, for config cid."""
    import torch
    mesh, P, I, Pst = cases.config(cid)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    n = I.size
    dev = torch.device("cuda")
    v = torch.from_numpy(np.random.default_rng(5).standard_normal(n)).to(dev)
    gs = torch.empty(n, dtype=torch.float64, device=dev)
    gp = torch.empty_like(gs)
    gt = torch.empty(6, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    out = {"config": cid, "elements": int(mesh["conn"].shape[0]), "dofs": int(n)}
    for name, mode, args in (("exact", abi.GMC_VJP_EXACT, (gs.data_ptr(), gp.data_ptr(), gt.data_ptr())),
                             ("frozen", abi.GMC_VJP_FROZEN, (gs.data_ptr(), 0, 0))):
        ctx.vjp_device(v.data_ptr(), *args, mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.vjp_device(v.data_ptr(), *args, mode)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[f"vjp_{name}_ms"] = dt * 1e3
        out[f"vjp_{name}_elem_per_s"] = mesh["conn"].shape[0] / dt
    ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS)
    out["assemble_fp64_ms"] = (time.perf_counter() - t0) / steps * 1e3
    ctx.close()
    return out


def solve_eval(cid, iters):
    """The Krylov consumer of the assembled system on config cid: GMRES(50),
    right preconditioning, x0 = 0, capped at `iters` iterations, with point
    Jacobi, block Jacobi and block ILU(0).  Reports iterations, the true relative residual,
    time per iteration and the HBM bytes one iteration streams at least (the
    BSR values and indices once for the SpMV, and CGS2's two passes over the
    basis so far, on average half the restart length)."""
    mesh, P, I, Pst = cases.config(cid)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    info = ctx.info
    ctx.assemble(abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS)
    n = info.n_nodes * info.ndof
    restart = 50
    spmv_bytes = info.nnzb * (info.ndof * info.ndof * 8 + 4) + (info.n_nodes + 1) * 4 + 2 * n * 8
    orth_bytes = 2 * 2 * (restart / 2) * n * 8
    out = {"config": cid, "dofs": n, "restart": restart, "iteration_cap": iters,
           "bytes_per_iteration_min": spmv_bytes + orth_bytes}
    for name, pc in (("point_jacobi", 1), ("block_jacobi", 2), ("block_ilu0", 3)):
        ctx.solve(restart=restart, max_iters=2, rel_tol=1e-30, precond=pc)  # warm-up, allocations
        si = ctx.solve(restart=restart, max_iters=iters, rel_tol=1e-8, precond=pc)
        ms_it = si.seconds / max(si.iterations, 1) * 1e3
        out[name] = {"iterations": si.iterations, "rel_residual": si.rel_residual,
                     "seconds": si.seconds, "ms_per_iteration": ms_it,
                     "GBps_min": out["bytes_per_iteration_min"] / (ms_it * 1e-3) / 1e9}
    ctx.close()
    return out


def full_eval(adj_scale, n_steps, ckpt):
    """Config 4 (3D Kuhn box 192x192x96 on [0,2]x[0,2]x[0,1], 21,233,664
    tetrahedra, 4 components, 8 dofs/node).  At full size: FP64 and FP32-tier
    assembly, the exact VJP (state, previous state, 6 material parameters) and
    the frozen J^T v on device vectors, and the device memory in use.  Then an
    n_steps-step adjoint of the damped block-Jacobi integrator of
    gales_mc.adjoint at 1/adj_scale resolution, with checkpoints every ckpt
    steps (recompute in between) and without: forward and backward times and
    torch's peak device memory above its baseline.  The objective is a fixed
    seeded linear functional of the final state.  (Unmasked, that integrator
    is stable on this box up to 1/8 resolution; finer meshes need the smooth-
    element mask the 2D tests use.)"""
    import torch
    from gales_mc import adjoint, material
    f64 = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS
    f32 = f64 | abi.GMC_ASM_FP32
    out = {"config": 4}
    dev = torch.device("cuda")

    def say(msg):
        print(f"# full-eval: {msg}", file=sys.stderr, flush=True)

    say("building the 192x192x96 Kuhn mesh")
    mesh, P, I, Pst = cases.config(4)
    say("creating the context")
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    info = ctx.info
    E = info.n_elems
    out.update({"elements": E, "nodes": info.n_nodes, "ndof": info.ndof, "nnzb": info.nnzb,
                "scatter": ["colour", "element buffer", "patch"][info.scatter],
                "context_gib": info.device_bytes / 2**30})
    for name, fl in (("fp64", f64), ("fp32", f32)):
        say(f"timing {name} assembly")
        el, _, _ = time_assembly(ctx, fl, 3, 1)
        out[f"assemble_{name}_ms"] = el / 3 * 1e3
        out[f"assemble_{name}_elem_per_s"] = E / (el / 3)
        # one interior launch: element buffer; one per colour: colour scatter
        out[f"assemble_{name}_interior_launches"] = ctx.timing().interior_launches
    out["fp32_speedup"] = out["assemble_fp64_ms"] / out["assemble_fp32_ms"]
    n = I.size
    v = torch.from_numpy(np.random.default_rng(5).standard_normal(n)).to(dev)
    gs = torch.empty(n, dtype=torch.float64, device=dev)
    gp = torch.empty_like(gs)
    gt = torch.empty(6, dtype=torch.float64, device=dev)
    for name, mode, args, reps in (("frozen", abi.GMC_VJP_FROZEN, (gs.data_ptr(), 0, 0), 3),
                                   ("exact", abi.GMC_VJP_EXACT,
                                    (gs.data_ptr(), gp.data_ptr(), gt.data_ptr()), 1)):
        say(f"timing the {name} VJP")
        ctx.vjp_device(v.data_ptr(), *args, mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.vjp_device(v.data_ptr(), *args, mode)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[f"vjp_{name}_ms"] = dt * 1e3
        out[f"vjp_{name}_elem_per_s"] = E / dt
    free, total = torch.cuda.mem_get_info()
    out["device_used_gib_full_size"] = (total - free) / 2**30
    ctx.close()
    del mesh, I, Pst, v, gs, gp
    say(f"full size done: {json.dumps(out)}")

    say(f"adjoint at 1/{adj_scale} resolution: mesh, context, block-diagonal inverse")
    mesh, P, I, Pst = cases.config(4, scale=adj_scale)
    ctx = Context(mesh, P)
    E = ctx.info.n_elems
    Binv = adjoint.block_diag_inverse(ctx, I, Pst)
    st = adjoint.Stepper(ctx, Binv, omega=0.25)
    w = torch.from_numpy(np.random.default_rng(7).standard_normal(I.size)).to(dev)

    def obj(Iv, th):
        return (w * Iv).sum()

    def run(k, ck):
        It = torch.from_numpy(I.reshape(-1)).to(dev).requires_grad_(True)
        Pt = torch.from_numpy(Pst.reshape(-1)).to(dev)
        th = torch.tensor(material.theta_of(P), dtype=torch.float64, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = adjoint.rollout(st, It, Pt, th, k, obj, checkpoint_every=ck)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        return t1 - t0, time.perf_counter() - t1, th.grad, It.grad

    say("warm-up step")
    run(1, None)  # first use allocates the library and allocator workspaces
    res = {"scale": adj_scale, "elements": E, "steps": n_steps}
    grads = {}
    for tag, ck in (("checkpointed", ckpt), ("all_states_kept", None)):
        say(f"{n_steps}-step rollout and backward, {tag}")
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        t_f, t_b, gth, gI = run(n_steps, ck)
        res[tag] = {"checkpoint_every": ck, "forward_s": t_f, "backward_s": t_b,
                    "peak_torch_mib_above_baseline": (torch.cuda.max_memory_allocated() - base) / 2**20}
        grads[tag] = (gth.numpy().copy(), gI.cpu().numpy())
    res["adjoint_elem_steps_per_s"] = E * n_steps / res["checkpointed"]["backward_s"]
    res["gradients_equal_bitwise"] = bool(np.array_equal(grads["checkpointed"][0], grads["all_states_kept"][0])
                                          and np.array_equal(grads["checkpointed"][1],
                                                             grads["all_states_kept"][1]))
    res["dL_dtheta"] = [float(x) for x in grads["checkpointed"][0]]
    res["note"] = ("backward = recompute of each checkpoint segment + one exact VJP (state, previous "
                   "state, 6 material parameters) per step")
    out["adjoint_rollout"] = res
    ctx.close()
    return out


def time_assembly(ctx, flags, steps, warmup):
    """Wall time of `steps` assemblies (each synchronised, as the caller of
    fluid_assembly::execute sees it) and the HIP-event time of the interior
    element kernel summed over them."""
    for _ in range(warmup):
        ctx.assemble(flags)
    t_int = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble(flags)
        tm = ctx.timing()
        t_int += tm.interior_s
        launches += tm.interior_launches
    return time.perf_counter() - t0, t_int, launches


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1, help="workload: 1 (2D FP64, default) or 2 (3D)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="element-matrix precision of the timed assembly (fp32: the precision tier)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-refine", action="store_true",
                    help="skip the bounded refinement leg (the tier's Newton residual)")
    ap.add_argument("--tier-eval", nargs=3, type=int, metavar=("CONFIG", "SCALE", "MAX_ITERS"),
                    help="precision-tier evaluation instead of the benchmark: FP32-element assembly + "
                         "FP64 refinement vs the plain FP64 path on config CONFIG at 1/SCALE resolution "
                         "(block-Jacobi GMRES(100), at most MAX_ITERS iterations per solve); prints one "
                         "JSON line")
    ap.add_argument("--tier-passes", type=int, default=3, help="refinement passes of --tier-eval")
    ap.add_argument("--tier-inner-tol", type=float, default=1e-6,
                    help="relative tolerance of each inner solve of --tier-eval")
    ap.add_argument("--tier-precond", type=int, default=2, choices=[1, 2, 3],
                    help="preconditioner of --tier-eval's solves: 1 point / 2 block Jacobi, 3 block ILU(0)")
    ap.add_argument("--tier-restart", type=int, default=100,
                    help="GMRES restart length of --tier-eval's solves (at most 200)")
    ap.add_argument("--tier-no-fp64-solve", action="store_true",
                    help="--tier-eval without the plain FP64 reference solve")
    ap.add_argument("--vjp-eval", nargs=2, type=int, metavar=("CONFIG", "STEPS"),
                    help="adjoint timing instead of the benchmark: exact and frozen-coefficient "
                         "vector-Jacobian products of the residual on config CONFIG (device vectors, "
                         "STEPS timed calls each after one warm-up); prints one JSON line")
    ap.add_argument("--solve-eval", nargs=2, type=int, metavar=("CONFIG", "ITERS"),
                    help="GMRES(50) with point and block Jacobi on the assembled system of CONFIG, "
                         "at most ITERS iterations, instead of the benchmark; prints one JSON line")
    ap.add_argument("--full-eval", nargs=3, type=int, metavar=("ADJ_SCALE", "STEPS", "CKPT"),
                    help="config 4 (3D, 4 components, 21.2M tets) instead of the benchmark: FP64 and "
                         "FP32-tier assembly at full size, then a STEPS-step checkpointed adjoint "
                         "(checkpoint every CKPT) at 1/ADJ_SCALE resolution; prints one JSON line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (torch) to calibrate "
                         "the FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 2**20 // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 8, "launches": 3}))
        return
    if args.tier_eval:
        print(json.dumps(tier_eval(*args.tier_eval, passes=args.tier_passes, inner_tol=args.tier_inner_tol,
                                   fp64_solve=not args.tier_no_fp64_solve, precond=args.tier_precond,
                                   restart=args.tier_restart)),
              flush=True)
        return
    if args.vjp_eval:
        print(json.dumps(vjp_eval(*args.vjp_eval)), flush=True)
        return
    if args.solve_eval:
        print(json.dumps(solve_eval(*args.solve_eval)), flush=True)
        return
    if args.full_eval:
        print(json.dumps(full_eval(*args.full_eval)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding of this path is design-only", file=sys.stderr)
        sys.exit(2)

    mesh, P, I, Pst = cases.config(args.config)
    ctx = Context(mesh, P)
    ctx.set_state(I, Pst)
    info = ctx.info
    print(f"# workload: {info.n_elems} elements, {info.n_nodes} nodes, nnzb {info.nnzb}, "
          f"{info.n_colors} colours, device {info.device_bytes / 2**30:.2f} GiB", file=sys.stderr, flush=True)
    fp64 = abi.GMC_ASM_MATRIX | abi.GMC_ASM_RHS
    fp32 = fp64 | abi.GMC_ASM_FP32
    main_flags = fp32 if args.precision == "fp32" else fp64
    other_flags = fp64 if args.precision == "fp32" else fp32
    elapsed, t_int, launches = time_assembly(ctx, main_flags, args.steps, args.warmup)
    ms = elapsed / args.steps * 1e3
    value = info.n_elems / (elapsed / args.steps)
    B = algorithmic_bytes(mesh, info)
    per_elem = B / info.n_elems
    # dominant kernel: interior element kernel (one launch over all elements)
    achieved = per_elem * info.n_elems * args.steps / t_int / 1e9
    print(f"# {args.precision}: step {ms:.3f} ms, interior kernel {t_int / args.steps * 1e3:.3f} ms/step "
          f"({launches // args.steps} launches), {per_elem:.1f} B/elem", file=sys.stderr, flush=True)
    # the other precision on the same workload, for the tier comparison
    e2, t2, _ = time_assembly(ctx, other_flags, args.steps, args.warmup)
    ms2 = e2 / args.steps * 1e3
    ms64, ms32 = (ms, ms2) if args.precision == "fp64" else (ms2, ms)
    print(f"# tiers: fp64 {ms64:.3f} ms/step, fp32 {ms32:.3f} ms/step, speedup {ms64 / ms32:.2f}x",
          file=sys.stderr, flush=True)
    # the tier's Newton residual on this workload at full size: mixed-precision
    # refinement (GMRES(100) + block Jacobi on the FP32-tier matrix to 1e-2 per
    # pass, FP64 matrix-free residual), a bounded number of passes
    refine = None
    if not args.no_refine:
        t0 = time.perf_counter()
        res = ctx.refine(n_refine=REFINE_PASSES, restart=100, max_iters=150, rel_tol=1e-2, precond=2)
        refine = {"passes": REFINE_PASSES, "inner": "GMRES(100) + block Jacobi to 1e-2, at most 150 iterations",
                  "newton_rel_residual_per_pass": [float(r / res[0]) for r in res],
                  "newton_rel_residual": float(res[-1] / res[0]), "seconds": time.perf_counter() - t0}
        print(f"# refinement: ||rhs - J64 x|| / ||rhs|| per pass " +
              " ".join(f"{x:.2e}" for x in refine["newton_rel_residual_per_pass"]), file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 matrix / f64 residual",
        "data": "synthetic",
        "config": {"workload": (f"{'2D channel' if info.dim == 2 else '3D box'}: {info.n_elems} P1 "
                                f"{'triangles' if info.dim == 2 else 'tetrahedra'}, {info.n_nodes} nodes, "
                                f"{info.nc} components, {info.ndof} dofs/node, "
                                f"{args.precision.upper()} R+J assembly into device BSR"),
                   "elements": info.n_elems, "nodes": info.n_nodes, "nnzb": info.nnzb,
                   "precision": args.precision,
                   # the precision tier on the same workload, as scalars
                   "tier_fp64_ms_per_step": ms64, "tier_fp32_ms_per_step": ms32,
                   "tier_fp32_speedup": ms64 / ms32,
                   "tier_fp32_element_matrix": TIER_MATRIX_UNIT_MFMA if info.fp32_mfma else TIER_MATRIX_UNIT,
                   "tier_newton_rel_residual": refine["newton_rel_residual"] if refine else None},
        "tiers": {"fp64_ms_per_step": ms64, "fp32_ms_per_step": ms32, "fp32_speedup": ms64 / ms32,
                  "refinement": refine},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     # the committed counter summaries cover the 2D channel's
                     # default (patch-local) scatter only
                     "traffic": (measured_traffic(args.precision)
                                 if args.config == 1 and info.scatter == 2 else None),
                     "kernel": ("interior element kernel: k_patch (patch-local LDS assembly)"
                                if info.scatter == 2 else "interior element kernel: k_interior"),
                     "kernel_ms": t_int / args.steps * 1e3, "bytes_per_element": per_elem,
                     "compute": (measured_valu(args.precision, t_int / args.steps)
                                 if args.config == 1 and info.scatter == 2 else None)},
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
