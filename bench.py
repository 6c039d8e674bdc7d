"""Benchmark of the octree-AMR lattice-Boltzmann step on one MI355X.

Workload (one "step" = one coarse time step, S_Advance of the root level with recursive
sub-cycling): D3Q19 BGK lid-driven cavity, root grid 64^3 in 4x4x4 cell blocks, 3 refinement
levels marked by the velocity-gradient-magnitude criterion (|grad u| above 2^-2 on the root level,
halved per level) and re-evaluated every 100 coarse steps, Re = 1000, lid
velocity 0.05, FP32 populations in structure-of-arrays layout.  The warm-up runs the real driver
(M_AdvanceLoop: steps plus solution-based rebuilds) from rest so the mesh is the one the flow
produces; the timed window then runs K coarse steps.

value: MLUPS counting every cell update of every active block (2^L sub-steps on level L) divided
by the step time of the window, rebuild time excluded (reported separately, and folded in at one
measured rebuild per 100 coarse steps as mlups_rebuilds_amortized).
roofline: the fused collide+boundary+stream kernel, algorithmic bytes 2*Q*4 per cell update
divided by its average launch time from HIP events on the context stream.
cpu_baseline: the CPU restatement (OpenMP) stepping a copy of the same mesh for a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + %HBM roofline, D3Q19 AMR; FP16-store vs FP32 at ≤1e-3 vel err"
T_START = time.time()


def log(msg: str):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.time() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=99)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--capacity", type=int, default=400000, help="block pool size")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-storage-study", action="store_true",
                    help="skip the FP16/BF16/FP32/FP64 comparison on one mesh sequence")
    ap.add_argument("--no-sphere-study", action="store_true",
                    help="skip the flow past the noisy icosphere (D3Q27 MRT, FP64/FP32/FP16)")
    ap.add_argument("--sphere-steps", type=int, default=1500, help="coarse steps of the sphere flow")
    ap.add_argument("--no-learned-study", action="store_true",
                    help="skip the learned refinement predictor on the cavity workload")
    ap.add_argument("--roofline-steps", type=int, default=20)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying recorded step graphs (same kernels "
                         "and results; for runs under a kernel tracer)")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (torch) to calibrate HBM byte counters")
    return ap.parse_args()


def main():
    args = parse()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB read + 1 GiB written
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU partitioning is "
              "a design, not built.", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from agal_amd import capi
    from agal_amd.mesh import LbmMesh

    Q, b = 19, 4
    N, U, Re, levels = 64, 0.05, 1000.0, 3
    p = capi.default_params(velocity_set=capi.D3Q19, precision=capi.FP32, Nx=N, MAX_LEVELS_WALL=1,
                            MAX_LEVELS_INTERIOR=levels, v0=(U * N / Re) / N, bc_preset=capi.BC_LDC3D,
                            bc_velocity=U, P_REFINE=100, N_REFINE_START=-2.0, N_REFINE_INC=1.0,
                            N_REFINE_MAX=1.0, ic_rho=1.0, S_CRITERION=2)
    m = LbmMesh(p, args.capacity)
    if args.no_graphs:
        m.set_option(capi.OPT_GRAPHS, 0)
    m.S_SetIC(0)
    m.M_FreezeRefinedCells(0)
    log(f"cavity: {args.warmup} warm-up coarse steps with rebuilds")
    t0 = time.time()
    st_w, _, reb_w = m.advance_loop(0, args.warmup)
    m.sync()
    warm_s = time.time() - t0
    log(f"cavity: warm-up done in {warm_s:.1f} s; roofline pass and timed window")
    s = m.refresh_sizes()
    n_ids = [int(s.n_ids[L]) for L in range(levels)]

    # Roofline pass on the same mesh: HIP events around every fused launch (direct launches; the
    # stepping state advances but no rebuild is triggered).
    import ctypes as C
    m._f["timing_reset"](m.h, 1)
    for _ in range(args.roofline_steps):
        m.S_Advance(0)
    m.sync()
    fms, nl, fcells = C.c_double(), C.c_longlong(), C.c_double()
    m._f["timing_get"](m.h, C.byref(fms), C.byref(nl), C.byref(fcells))
    m._f["timing_reset"](m.h, 0)

    # Timed window: K coarse steps of the driver (each step replayed from a recorded HIP graph),
    # bracketed by HIP events on the context stream.
    m.sync()
    st, step_ms, rebuild_ms = m.advance_loop(args.warmup, args.steps)
    m.sync()
    s = m.refresh_sizes()
    n_ids_end = [int(s.n_ids[L]) for L in range(levels)]

    # Cell updates per coarse step (the mesh is fixed between rebuilds).
    cells = 64
    upd_actual = sum(n_ids_end[L] * cells * (1 << L) for L in range(levels))
    leaves = [n_ids_end[L] - (n_ids_end[L + 1] // 8 if L + 1 < levels else 0) for L in range(levels)]
    upd_reference = sum(leaves[L] * cells * (1 << L) for L in range(levels))
    upd_fine_equiv = (N << (levels - 1)) ** 3 * (1 << (levels - 1))
    K = args.steps
    step_s = step_ms * 1e-3
    mlups = upd_actual * K / step_s / 1e6
    mlups_ref = upd_reference * K / step_s / 1e6
    mlups_fine = upd_fine_equiv * K / step_s / 1e6
    mlups_e2e = upd_actual * K / ((step_ms + rebuild_ms) * 1e-3) / 1e6
    fused_ms = fms.value
    launches = nl.value
    alg_bytes = 2 * Q * b * fcells.value
    achieved = alg_bytes / (fused_ms * 1e-3) / 1e9 if fused_ms > 0 else 0.0

    # Rebuild cost on this mesh (one solution-based rebuild after the window).
    t1 = time.time()
    m.M_Advance_RefineWithSolution()
    m.sync()
    rebuild_one_ms = (time.time() - t1) * 1e3
    # The window holds no rebuild; with one rebuild per P_REFINE coarse steps at this cost (host
    # wall clock, so an upper bound) the rate over a long run is:
    mlups_amortized = upd_actual * K / ((step_ms + rebuild_one_ms * K / p.P_REFINE) * 1e-3) / 1e6

    traffic = None
    tp = os.path.join(ROOT, "profiles", "k_collide_stream_d3q19_f32_ldc64_3lev_traffic.json")
    if os.path.exists(tp):
        try:
            traffic = json.load(open(tp)).get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample")
        cpu = cpu_baseline(m, p, args.capacity, upd_actual, args.cpu_seconds)
    del m
    storage = None
    if not args.no_storage_study:
        log("storage study: FP64 / FP32 / FP16 / BF16 on one mesh sequence")
        storage = storage_study(p, args)
    learned = None
    if not args.no_learned_study:
        log("learned refinement study: 2000 training + 2000 inference coarse steps, with and without predictions")
        try:
            learned = learned_study(p, args)
        except Exception as e:
            log(f"learned study failed: {e}")
            learned = {"error": str(e)}
    sphere = None
    if not args.no_sphere_study:
        log("sphere study: D3Q27 MRT flow past the noisy icosphere")
        try:
            sphere = sphere_study(args)
        except Exception as e:  # reported in the line; the headline measurement stands on its own
            log(f"sphere study failed: {e}")
            sphere = {"error": str(e)}

    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: lid-driven cavity from rest, mesh produced by the solver's own refinement",
        "config": {
            "workload": "D3Q19 BGK lid-driven cavity, root 64^3 (4x4x4 blocks), 3 levels by the |grad u| "
                        "criterion, Re 1000, U 0.05, rebuild every 100 coarse steps, FP32 SoA",
            "blocks_per_level": n_ids_end,
            "cell_updates_per_step": upd_actual,
            "parallelism": "single GPU",
            "step_launch": "direct launches" if args.no_graphs else "one recorded HIP graph per coarse step",
        },
        "mlups_reference_formula": round(mlups_ref, 1),
        "mlups_fine_equivalent": round(mlups_fine, 1),
        "mlups_end_to_end": round(mlups_e2e, 1),
        "rebuild_ms_in_window": round(rebuild_ms, 3),
        "rebuild_ms_one": round(rebuild_one_ms, 3),
        "mlups_rebuilds_amortized": round(mlups_amortized, 1),
        "warmup_seconds": round(warm_s, 3),
        "step_hbm_gbs_algorithmic": round(2 * Q * b * upd_actual * K / step_s / 1e9, 1),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_collide_stream<float, float, 19, false>",
            "launches": launches,
            "avg_launch_ms": round(fused_ms / max(1, launches), 5),
            "bytes_per_launch_algorithmic": round(alg_bytes / max(1, launches)),
        },
        "cpu_baseline": cpu,
        "storage16": storage,
        "learned_refinement": learned,
        "sphere": sphere,
    }
    txt = json.dumps(line)
    print(txt, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")


def storage_study(p, args):
    """FP16 / BF16 pools of f - w_q against FP32 and FP64 on one mesh sequence.

    An FP64 run from rest leads; FP32, FP16 and BF16 runs follow it, taking its refinement marks
    at every rebuild, so all four share every mesh.  After the warm-up the velocity of the leaf
    cells is compared with FP64, then each follower times K coarse steps (graph replays) on the
    shared mesh, and the fused kernel of each is timed per launch."""
    import ctypes as C

    from agal_amd import capi
    from agal_amd.mesh import LbmMesh
    from agal_amd.twins import advance_twins, velocity_error

    precs = {"fp64": capi.FP64, "fp32": capi.FP32, "fp16": capi.FP16, "bf16": capi.BF16}
    ctx = {}
    for k, v in precs.items():
        q = capi.Params.from_buffer_copy(p)
        q.precision = v
        ctx[k] = LbmMesh(q, args.capacity)
        if args.no_graphs:
            ctx[k].set_option(capi.OPT_GRAPHS, 0)
        ctx[k].S_SetIC(0)
        ctx[k].M_FreezeRefinedCells(0)
    t0 = time.time()
    advance_twins(ctx["fp64"], [ctx["fp32"], ctx["fp16"], ctx["bf16"]], 0, args.warmup)
    warm_s = time.time() - t0
    errs = {k: velocity_error(ctx["fp64"], ctx[k]) for k in ("fp32", "fp16", "bf16")}
    s = ctx["fp64"].refresh_sizes()
    levels = s.max_levels
    n_ids = [int(s.n_ids[L]) for L in range(levels)]
    upd = sum(n_ids[L] * 64 * (1 << L) for L in range(levels))
    out = {"warmup_steps": args.warmup, "warmup_seconds": round(warm_s, 2), "blocks_per_level": n_ids,
           "cell_updates_per_step": upd, "vel_err_vs_fp64": {}, "mlups": {}, "fused_kernel": {}}
    bpp = {"fp64": 8, "fp32": 4, "fp16": 2, "bf16": 2}
    for k in ("fp64", "fp32", "fp16", "bf16"):
        m = ctx[k]
        log(f"storage study: timing {k}")
        m.sync()
        _, step_ms, _ = m.advance_loop(args.warmup, args.steps)
        m.sync()
        out["mlups"][k] = round(upd * args.steps / (step_ms * 1e-3) / 1e6, 1)
        m._f["timing_reset"](m.h, 1)
        for _ in range(max(1, args.roofline_steps // 2)):
            m.S_Advance(0)
        m.sync()
        fms, nl, fc = C.c_double(), C.c_longlong(), C.c_double()
        m._f["timing_get"](m.h, C.byref(fms), C.byref(nl), C.byref(fc))
        m._f["timing_reset"](m.h, 0)
        alg = 2 * 19 * bpp[k] * fc.value
        out["fused_kernel"][k] = {"avg_launch_ms": round(fms.value / max(1, nl.value), 5),
                                  "achieved_gbs": round(alg / (fms.value * 1e-3) / 1e9, 1) if fms.value > 0 else None,
                                  "bytes_per_cell_update": 2 * 19 * bpp[k]}
        if k != "fp64":
            out["vel_err_vs_fp64"][k] = {"rel_l2": float(f"{errs[k]['rel_l2_velocity']:.4g}"),
                                         "max_abs": float(f"{errs[k]['max_abs_velocity']:.4g}")}
    out["leaf_cells_compared"] = errs["fp16"]["leaf_cells"]
    out["speedup_fp16_vs_fp32"] = round(out["mlups"]["fp16"] / out["mlups"]["fp32"], 3)
    out["fp16_within_1e-3"] = out["vel_err_vs_fp64"]["fp16"]["rel_l2"] <= 1e-3
    return out


def learned_study(p, args):
    """The cavity workload with the learned refinement predictor (12-64-64-1 MLP, FP16 weights,
    Adam lr 1e-3, batches of 4096 blocks): features snapshotted every 25 coarse steps and trained on
    the plain criterion's labels at each rebuild over the first 2000 steps, then inference every 25
    steps.  The same run without applying predictions gives the plain method's numbers; the referee is
    the plain criterion at every rebuild (blocks it flags that are still unrefined = missed)."""
    from agal_amd import capi
    from agal_amd.mesh import LbmMesh

    out = {"train_steps": 2000, "inference_steps": 2000, "infer_every": 25, "batch": 4096}
    res = {}
    for use in (0, 1):
        m = LbmMesh(p, args.capacity)
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        opts = capi.learn_opts(train_steps=2000, infer_every=25, batch=4096, adam_steps=8, lr=1e-3, threshold=0.5,
                               pos_weight=4.0, seed=11, use_predictions=use)
        t0 = time.time()
        res[use] = m.advance_loop_learned(0, 4000, opts)
        res[use]["wall_s"] = round(time.time() - t0, 3)
        s = m.refresh_sizes()
        res[use]["blocks_per_level"] = [int(s.n_ids[L]) for L in range(s.max_levels)]
        m.close()
        log(f"learned study: use_predictions={use} done in {res[use]['wall_s']} s")
    for key, r in (("plain", res[0]), ("learned", res[1])):
        out[key] = {
            "flagged": r["flagged"], "missed": r["missed"],
            "missed_rate": round(r["missed"] / r["flagged"], 5) if r["flagged"] else None,
            "predicted_refinements": r["predicted_refinements"], "early_rebuilds": r["early_rebuilds"],
            "step_ms": round(r["step_ms"], 2), "rebuild_ms": round(r["rebuild_ms"], 2),
            "predictor_ms": round(r["predictor_ms"], 2), "blocks_per_level_end": r["blocks_per_level"],
        }
    out["training"] = {"adam_steps": res[1]["adam_steps_done"], "last_loss": round(res[1]["last_loss"], 5),
                       "plain_missed_rate_during_training": round(res[1]["missed_train"] / max(1, res[1]["flagged_train"]), 5)}
    out["learned_within_1pct"] = out["learned"]["missed_rate"] is not None and out["learned"]["missed_rate"] <= 0.01
    return out


def sphere_study(args):
    """Flow past the seeded noisy icosphere: 5,120 triangles, radius 0.08 L with vertex radii
    perturbed by +-5% (uniform, seeded), centred at 0.3 L in the cubic domain; D3Q27 MRT, root 64^3,
    4 levels refined near the surface, inflow 0.04, Re 200 on the diameter.  FP64 (vector MRT), FP32
    (matrix-core MRT and vector MRT) and FP16 storage of f - w (matrix-core MRT) run the same static
    mesh from the same uniform start in lockstep; leaf velocities are compared with FP64 every 250
    coarse steps and the drag (momentum exchange, FP64 reduction) is averaged over the last 500."""
    import ctypes as C
    import math

    from agal_amd import capi
    from agal_amd.mesh import LbmMesh
    from agal_amd.twins import velocity_error

    N, U, Re, D, levels = 64, 0.04, 200.0, 0.16, 4
    tri = capi.icosphere((0.3, 0.5, 0.5), D / 2, 4, 0.05, 20240)
    out = {"workload": "D3Q27 MRT flow past the noisy icosphere (5120 triangles, r 0.08 L +-5%, seed 20240), "
                       "root 64^3, 4 wall levels, inflow 0.04, Re 200, MRT rates s1..s5 parser defaults, s6 1.2 "
                       "(1.98 diverges from the impulsive start)",
           "coarse_steps": args.sphere_steps, "mlups": {}, "cd": {}, "fused_kernel": {},
           "mrt_transform": {"fp64": "vector", "fp32": "matrix cores", "fp32_vector": "vector",
                             "fp16": "matrix cores"}}
    ctx = {}
    for k, prec in (("fp64", capi.FP64), ("fp32", capi.FP32), ("fp32_vector", capi.FP32), ("fp16", capi.FP16)):
        p = capi.default_params(velocity_set=capi.D3Q27, precision=prec, collision=capi.MRT, Nx=N,
                                MAX_LEVELS_WALL=levels, MAX_LEVELS_INTERIOR=1, v0=U * D / Re,
                                bc_preset=capi.BC_FPSC3D, bc_velocity=U, ic_u=U, P_REFINE=1 << 30, s6=1.2)
        m = LbmMesh(p, args.capacity)
        if k in ("fp32", "fp16"):
            m.set_option(capi.OPT_MRT_MFMA, 1)
        m.set_geometry(tri)
        t0 = time.time()
        m.M_Advance_RefineNearWall()
        m.sync()
        out.setdefault("near_wall_setup_s", {})[k] = round(time.time() - t0, 3)
        ctx[k] = m
        log(f"sphere {k}: near-wall hierarchy {[int(m.sizes.n_ids[L]) for L in range(levels)]}")
    s = ctx["fp64"].refresh_sizes()
    n_ids = [int(s.n_ids[L]) for L in range(levels)]
    upd = sum(n_ids[L] * 64 * (1 << L) for L in range(levels))
    out["blocks_per_level"] = n_ids
    out["cell_updates_per_step"] = upd
    K = args.sphere_steps
    avg_from = max(0, K - 500)
    # All runs advance in lockstep chunks; after each chunk the leaf velocities are compared with
    # FP64 (the error series shows how fast the runs separate in this unsteady flow).
    ms = {k: 0.0 for k in ctx}
    cds = {k: [] for k in ctx}
    series = []
    done = 0
    while done < K:
        step = min(10 if done >= avg_from else 250, K - done, (avg_from - done) if done < avg_from else K)
        for k, m in ctx.items():
            _, t, _ = m.advance_loop(done, step)
            ms[k] += t
        done += step
        if done > avg_from:
            for k, m in ctx.items():
                F = m.S_ComputeForces()
                cds[k].append(8.0 * F[0] / (U * U * math.pi * D * D))
        if done % 250 == 0 or done == K:
            row = {"step": done}
            for k in ctx:
                if k != "fp64":
                    row[k] = float(f"{velocity_error(ctx['fp64'], ctx[k])['rel_l2_velocity']:.3g}")
            series.append(row)
            log(f"sphere: {done} coarse steps, velocity error vs fp64 {row}")
    for k, m in ctx.items():
        out["mlups"][k] = round(upd * K / (ms[k] * 1e-3) / 1e6, 1)
        out["cd"][k] = round(sum(cds[k]) / len(cds[k]), 5)
        m._f["timing_reset"](m.h, 1)
        for _ in range(5):
            m.S_Advance(0)
        m.sync()
        fms, nl, fc = C.c_double(), C.c_longlong(), C.c_double()
        m._f["timing_get"](m.h, C.byref(fms), C.byref(nl), C.byref(fc))
        m._f["timing_reset"](m.h, 0)
        b = {"fp64": 8, "fp32": 4, "fp32_vector": 4, "fp16": 2}[k]
        out["fused_kernel"][k] = {"avg_launch_ms": round(fms.value / max(1, nl.value), 5),
                                  "achieved_gbs": round(2 * 27 * b * fc.value / (fms.value * 1e-3) / 1e9, 1)
                                  if fms.value > 0 else None}
    out["cd_window"] = f"mean over the last {K - avg_from} coarse steps, sampled every 10"
    out["vel_err_vs_fp64_series"] = series
    out["cd_rel_err_vs_fp64"] = {k: float(f"{abs(out['cd'][k] / out['cd']['fp64'] - 1):.3g}") for k in ctx if k != "fp64"}
    out["cd_schiller_naumann_unbounded"] = round(24.0 / Re * (1 + 0.15 * Re ** 0.687), 4)
    for m in ctx.values():
        m.close()
    return out


def cpu_baseline(m, p, capacity, upd_per_step, budget_s):
    """The CPU restatement stepping the same mesh (state copied from the GPU) on the host cores."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import OracleMesh, set_threads
    except Exception as e:  # the checker is not built
        return {"value": None, "unit": "MLUPS", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = set_threads(max(1, min(ncpu, 16)))
    o = OracleMesh(p, capacity)
    o.import_state(m.state())
    o.S_Advance(0)  # first touch
    steps, t = 0, 0.0
    while t < budget_s and steps < 200:
        t0 = time.time()
        o.S_Advance(0)
        t += time.time() - t0
        steps += 1
    o.close()
    return {
        "value": round(upd_per_step * steps / t / 1e6, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} coarse steps of the same 3-level mesh (FP32, OpenMP over blocks), {t:.1f} s",
    }


if __name__ == "__main__":
    main()
