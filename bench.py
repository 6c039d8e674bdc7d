"""Benchmark of the octree-AMR lattice-Boltzmann step on one MI355X.

Headline workload (one "step" = one coarse time step, S_Advance of the root level with recursive
sub-cycling): D3Q19 BGK lid-driven cavity, root grid 64^3 in 4x4x4 cell blocks, 3 refinement levels
marked by the velocity-gradient-magnitude criterion (|grad u| above 2^-2 on the root level, halved
per level) and re-evaluated every 100 coarse steps, Re 1000, lid velocity 0.05, FP32 populations in
structure-of-arrays layout.

Phases (the mesh is the one the flow produces, independent of --warmup):
  1. development: `--develop` coarse steps of the advance loop (M_AdvanceLoop: steps plus the
     solution-based rebuilds every 100 steps) from rest; the run stops with an error if the finer
     levels stay empty;
  2. warm-up: `--warmup` coarse steps on the developed mesh, untimed;
  3. roofline pass: `--roofline-steps` coarse steps with HIP events around every launch of the
     fused collide + boundary + stream kernel (direct launches on the context stream);
  4. timed window: `--steps` coarse steps of the advance loop (recorded HIP graphs), HIP events on the
     context stream; a rebuild that falls inside the window is timed separately and the cell
     updates are counted on the mesh of each stretch between rebuilds.

value: actual cell updates per second (every active block, 2^L sub-steps on level L) in MLUPS.
roofline: the fused kernel, algorithmic bytes 2*Q*4 per cell update / its average launch time.
--copy-calibration runs only a plain torch device copy (its own process: torch brings its own HIP
runtime): three 1 GiB copies for the byte-counter calibration, then the copy rate (bytes read +
written per second) for 256 MiB, 1 GiB and 4 GiB sources, the practical HBM ceiling to set beside
the fused kernel's rate.
cpu_baseline: the CPU restatement (the oracle, OpenMP over blocks) stepping a copy of the
developed mesh on all host cores (up to 16), plus a serial leg.

A second leg runs the large cavity (root 256^3, 5 levels, Re 5000, rebuild every 50 coarse
steps) in FP32 and FP16 storage; studies of the storage precision, the learned refinement
predictor and the D3Q27 MRT sphere flow follow.
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
HBM_PEAK_GBS = 8000.0
U_LID = 0.05


def log(msg: str):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.time() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=99)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--develop", type=int, default=1000,
                    help="coarse steps from rest with rebuilds before the warm-up (builds the adaptive mesh)")
    ap.add_argument("--capacity", type=int, default=400000, help="block pool size of the 64^3 cavity")
    ap.add_argument("--storage-develop", type=int, default=10000,
                    help="coarse steps of the FP64/FP32/FP16/BF16 runs before the velocity comparison (configuration 2 runs 10 000)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample (both legs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3 five-level cavity leg")
    ap.add_argument("--large-develop", type=int, default=300)
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--large-capacity", type=int, default=18000000)
    ap.add_argument("--no-storage-study", action="store_true",
                    help="skip the FP16/BF16/FP32/FP64 comparison on one mesh sequence")
    ap.add_argument("--no-sphere-study", action="store_true",
                    help="skip the flow past the noisy icosphere (D3Q27 MRT, FP64/FP32/FP16)")
    ap.add_argument("--sphere-steps", type=int, default=1000, help="coarse steps of the sphere flow")
    ap.add_argument("--no-learned-study", action="store_true",
                    help="skip the learned refinement predictor on the cavity workload")
    ap.add_argument("--roofline-steps", type=int, default=20)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly instead of replaying recorded step graphs (same kernels "
                         "and results; for runs under a kernel tracer)")
    ap.add_argument("--only-headline", action="store_true", help="headline leg only (no CPU, large leg or studies)")
    ap.add_argument("--cavity-precision", choices=("fp32", "fp16", "bf16"), default="fp32",
                    help="population storage of the cavity leg (fp32 is the headline; the others are for profiles)")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy (torch): 1 GiB copies to calibrate HBM byte counters, then "
                         "the copy rate of 256 MiB, 1 GiB and 4 GiB sources")
    return ap.parse_args()


def cavity_params(N, levels, Re, P_REFINE, precision=None):
    from agal_amd import capi
    return capi.default_params(velocity_set=capi.D3Q19, precision=capi.FP32 if precision is None else precision,
                               Nx=N, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=levels, v0=U_LID / Re,
                               bc_preset=capi.BC_LDC3D, bc_velocity=U_LID, P_REFINE=P_REFINE, N_REFINE_START=-2.0,
                               N_REFINE_INC=1.0, N_REFINE_MAX=1.0, ic_rho=1.0, S_CRITERION=2)


def blocks(m):
    s = m.refresh_sizes()
    return [int(s.n_ids[L]) for L in range(s.max_levels)]


def updates_per_step(n_ids):
    """Actual cell updates of one coarse step: every active block, 2^L sub-steps on level L."""
    return sum(n * 64 * (1 << L) for L, n in enumerate(n_ids))


def leaf_updates_per_step(n_ids):
    """The reference's MLUPS count (solver_lbm_advance.cu:81-87): leaf blocks weighted 2^L."""
    ML = len(n_ids)
    return sum((n_ids[L] - (n_ids[L + 1] // 8 if L + 1 < ML else 0)) * 64 * (1 << L) for L in range(ML))


def develop(m, n, what):
    """n coarse steps of the advance loop from rest (rebuilds included); returns the next step index."""
    t0 = time.time()
    done = 0
    while done < n:
        k = min(n - done, 250)
        m.advance_loop(done, k)
        done += k
        log(f"{what}: developed {done}/{n} coarse steps, blocks per level {blocks(m)}")
    m.sync()
    return n, time.time() - t0


def timed_window(m, i0, K):
    """K coarse steps of the advance loop from step index i0, split at rebuilds; returns step and rebuild
    times and the cell updates counted per stretch on its own mesh."""
    P = m.params.P_REFINE
    i, left = i0, K
    step_ms = rebuild_ms = 0.0
    upd = upd_ref = 0
    while left > 0:
        chunk = min(left, P - (i % P))  # the last step of a chunk may trigger a rebuild
        nid = blocks(m)
        st, s_ms, r_ms = m.advance_loop(i, chunk)
        step_ms += s_ms
        rebuild_ms += r_ms
        upd += updates_per_step(nid) * chunk
        upd_ref += leaf_updates_per_step(nid) * chunk
        i += chunk
        left -= chunk
    m.sync()
    return {"step_ms": step_ms, "rebuild_ms": rebuild_ms, "updates": upd, "updates_reference": upd_ref, "next": i}


def roofline_pass(m, steps):
    """Average launch time of the fused kernel over `steps` coarse steps (direct launches, HIP
    events recorded by the dispatches themselves) and the cells those launches updated."""
    import ctypes as C
    m._f["timing_reset"](m.h, 1)
    for _ in range(steps):
        m.S_Advance(0)
    m.sync()
    fms, nl, fcells = C.c_double(), C.c_longlong(), C.c_double()
    m._f["timing_get"](m.h, C.byref(fms), C.byref(nl), C.byref(fcells))
    m._f["timing_reset"](m.h, 0)
    return fms.value, int(nl.value), fcells.value


def profiled_traffic(kernel_file, n_ids):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when it
    was measured on this exact mesh (same blocks per level); otherwise None."""
    tp = os.path.join(ROOT, "profiles", kernel_file)
    try:
        d = json.load(open(tp))
    except Exception:
        return None, None
    if list(d.get("blocks_per_level", [])) != list(n_ids):
        return None, d.get("blocks_per_level")
    return d.get("bytes_per_launch"), d.get("blocks_per_level")


def cavity_leg(args, N, levels, Re, P_REFINE, capacity, develop_steps, steps, what, precision=None):
    from agal_amd import capi
    from agal_amd.mesh import LbmMesh

    p = cavity_params(N, levels, Re, P_REFINE, precision)
    m = LbmMesh(p, capacity)
    if args.no_graphs:
        m.set_option(capi.OPT_GRAPHS, 0)
    m.S_SetIC(0)
    m.M_FreezeRefinedCells(0)
    i, dev_s = develop(m, develop_steps, what)
    nid = blocks(m)
    if any(n == 0 for n in nid[1:]):
        raise SystemExit(f"{what}: the mesh did not develop its finer levels after {develop_steps} coarse steps "
                         f"({nid}); nothing adaptive to time")
    log(f"{what}: warm-up {args.warmup} coarse steps")
    if args.warmup > 0:
        w = timed_window(m, i, args.warmup)
        i = w["next"]
    nid_roof = blocks(m)
    log(f"{what}: roofline pass ({args.roofline_steps} coarse steps, events per fused launch)")
    fused_ms, launches, fcells = roofline_pass(m, args.roofline_steps)
    log(f"{what}: timed window of {steps} coarse steps")
    w = timed_window(m, i, steps)
    nid_end = blocks(m)
    step_s = w["step_ms"] * 1e-3
    Q = 19
    b = 2 if p.precision in (capi.FP16, capi.BF16) else (8 if p.precision == capi.FP64 else 4)
    alg_bytes = 2 * Q * b * fcells
    achieved = alg_bytes / (fused_ms * 1e-3) / 1e9 if fused_ms > 0 else 0.0
    fine_equiv = (N << (levels - 1)) ** 3 * (1 << (levels - 1))
    out = {
        "m": m,
        "p": p,
        "develop_steps": develop_steps,
        "develop_seconds": round(dev_s, 2),
        "blocks_per_level": nid_roof,
        "blocks_per_level_end": nid_end,
        "cell_updates_per_step": updates_per_step(nid_roof),
        "mlups": w["updates"] / step_s / 1e6,
        "mlups_reference_formula": w["updates_reference"] / step_s / 1e6,
        "mlups_fine_equivalent": fine_equiv * steps / step_s / 1e6,
        "mlups_end_to_end": w["updates"] / ((w["step_ms"] + w["rebuild_ms"]) * 1e-3) / 1e6,
        "ms_per_step": w["step_ms"] / steps,
        "rebuild_ms_in_window": w["rebuild_ms"],
        "step_hbm_gbs_algorithmic": 2 * Q * b * w["updates"] / step_s / 1e9,
        "fused": {"ms": fused_ms, "launches": launches, "cells": fcells, "alg_bytes": alg_bytes,
                  "achieved": achieved, "bytes_per_cell_update": 2 * Q * b},
        "next_step": w["next"],
    }
    return out


def headline(args):
    from agal_amd import capi

    N, levels, Re = 64, 3, 1000.0
    prec = {"fp32": capi.FP32, "fp16": capi.FP16, "bf16": capi.BF16}[args.cavity_precision]
    leg = cavity_leg(args, N, levels, Re, 100, args.capacity, args.develop, args.steps, "cavity 64^3", precision=prec)
    m, p = leg["m"], leg["p"]
    # Rebuild cost on this mesh (one solution-based rebuild after the window).
    t1 = time.time()
    m.M_Advance_RefineWithSolution()
    m.sync()
    rebuild_one_ms = (time.time() - t1) * 1e3
    step_ms = leg["ms_per_step"]
    # One rebuild per P_REFINE coarse steps at this cost (host wall clock, an upper bound):
    mlups_amortized = leg["cell_updates_per_step"] / ((step_ms + rebuild_one_ms / p.P_REFINE) * 1e-3) / 1e6
    fu = leg["fused"]
    traffic, prof_mesh = profiled_traffic("k_collide_stream_d3q19_f32_ldc64_3lev_traffic.json",
                                          leg["blocks_per_level"]) if args.cavity_precision == "fp32" else (None, None)
    roof = {
        "bound": "hbm",
        "achieved": round(fu["achieved"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(fu["achieved"] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": "k_collide_stream<float, %s, 19, false>" % {"fp32": "float", "fp16": "_Float16",
                                                               "bf16": "bf16s"}[args.cavity_precision],
        "launches": fu["launches"],
        "avg_launch_ms": round(fu["ms"] / max(1, fu["launches"]), 5),
        "bytes_per_launch_algorithmic": round(fu["alg_bytes"] / max(1, fu["launches"])),
        "bytes_per_cell_update_algorithmic": fu["bytes_per_cell_update"],
        "blocks_per_level": leg["blocks_per_level"],
        "traffic_source": ("profiles/k_collide_stream_d3q19_f32_ldc64_3lev_traffic.json (rocprofv3 FETCH_SIZE / "
                           "WRITE_SIZE on this mesh)") if traffic else
                          f"null: the committed counter profile is for mesh {prof_mesh}",
    }
    return leg, roof, rebuild_one_ms, mlups_amortized


def main():
    args = parse()
    if args.copy_calibration:
        # Three 1 GiB copies (the counter calibration: 1 GiB read + 1 GiB written each), then timed
        # copies (events on torch's stream) for the copy rate.
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB
        a = torch.ones(n, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        rates = {}
        for label, nbytes in (("256MiB", 1 << 28), ("1GiB", 1 << 30), ("4GiB", 1 << 32)):
            a = torch.ones(nbytes // 4, device="cuda:0")
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                b.copy_(a)
            e1.record()
            torch.cuda.synchronize()
            rates[label] = round(2 * nbytes / (e0.elapsed_time(e1) / 20 * 1e-3) / 1e9, 1)
            del a, b
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                          "copy_gbs": rates}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU partitioning is "
              "a design, not built.", file=sys.stderr)
        sys.exit(2)
    if args.only_headline:
        args.no_cpu = args.no_large = args.no_storage_study = args.no_sphere_study = args.no_learned_study = True

    leg, roof, rebuild_one_ms, mlups_amortized = headline(args)
    log(f"cavity 64^3: {leg['mlups']:.0f} MLUPS, fused kernel {roof['achieved']:.0f} GB/s ({roof['frac']:.3f})")
    cpu = None
    if not args.no_cpu:
        log("CPU baseline sample (serial and all host cores)")
        cpu = cpu_baseline(leg["m"], leg["p"], args.capacity, leg["cell_updates_per_step"], leg["blocks_per_level"],
                           args.cpu_seconds)
    leg["m"].close()
    large = None
    if not args.no_large:
        try:
            large = large_leg(args)
        except Exception as e:  # reported in the line; the headline stands on its own
            log(f"large cavity failed: {e}")
            large = {"error": str(e)}
    storage = None
    if not args.no_storage_study:
        log("storage study: FP64 / FP32 / FP16 / BF16 on one mesh sequence")
        storage = storage_study(leg["p"], args)
    learned = None
    if not args.no_learned_study:
        log("learned refinement study: 2000 training + 2000 inference coarse steps, with and without predictions")
        try:
            learned = learned_study(leg["p"], args)
        except Exception as e:
            log(f"learned study failed: {e}")
            learned = {"error": str(e)}
    sphere = None
    if not args.no_sphere_study:
        log("sphere study: D3Q27 MRT flow past the noisy icosphere")
        try:
            sphere = sphere_study(args)
        except Exception as e:
            log(f"sphere study failed: {e}")
            sphere = {"error": str(e)}

    K = args.steps
    line = {
        "metric": METRIC,
        "value": round(leg["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(leg["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "f16 storage, f32 compute", "bf16": "bf16 storage, f32 compute"}[args.cavity_precision],
        "data": "synthetic: lid-driven cavity from rest, mesh produced by the solver's own refinement",
        "config": {
            "workload": "D3Q19 BGK lid-driven cavity, root 64^3 (4x4x4 blocks), 3 levels by the |grad u| "
                        "criterion, Re 1000, U 0.05, rebuild every 100 coarse steps, FP32 SoA; "
                        f"mesh developed over {leg['develop_steps']} coarse steps from rest",
            "blocks_per_level": leg["blocks_per_level"],
            "blocks_per_level_end": leg["blocks_per_level_end"],
            "cell_updates_per_step": leg["cell_updates_per_step"],
            "parallelism": "single GPU",
            "step_launch": "direct launches" if args.no_graphs else "recorded HIP graphs, up to 4 coarse steps each",
        },
        "mlups_reference_formula": round(leg["mlups_reference_formula"], 1),
        "mlups_fine_equivalent": round(leg["mlups_fine_equivalent"], 1),
        "mlups_end_to_end": round(leg["mlups_end_to_end"], 1),
        "rebuild_ms_in_window": round(leg["rebuild_ms_in_window"], 3),
        "rebuild_ms_one": round(rebuild_one_ms, 3),
        "mlups_rebuilds_amortized": round(mlups_amortized, 1),
        "develop_seconds": leg["develop_seconds"],
        "step_hbm_gbs_algorithmic": round(leg["step_hbm_gbs_algorithmic"], 1),
        "step_roofline_frac": round(leg["step_hbm_gbs_algorithmic"] / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "large_cavity": large,
        "storage16": storage,
        "learned_refinement": learned,
        "sphere": sphere,
    }
    txt = json.dumps(line)
    print(txt, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")


def large_leg(args):
    """The large cavity: root 256^3 (262,144 blocks), 5 levels, Re 5000, rebuild every 50 coarse
    steps, timed in FP32 and in FP16 storage of f - w_q (each developed from rest)."""
    from agal_amd import capi

    out = {"workload": "D3Q19 BGK lid-driven cavity, root 256^3 (4x4x4 blocks), 5 levels by the |grad u| criterion "
                       "(2^-2 on the root level, halved per level), Re 5000, U 0.05, rebuild every 50 coarse steps",
           "develop_steps": args.large_develop, "steps": args.large_steps}
    for key, prec in (("fp32", capi.FP32), ("fp16", capi.FP16)):
        log(f"large cavity {key}: develop {args.large_develop} coarse steps")
        leg = cavity_leg(args, 256, 5, 5000.0, 50, args.large_capacity, args.large_develop, args.large_steps,
                         f"large cavity {key}", precision=prec)
        fu = leg["fused"]
        out[key] = {
            "mlups": round(leg["mlups"], 1),
            "mlups_reference_formula": round(leg["mlups_reference_formula"], 1),
            "mlups_fine_equivalent": round(leg["mlups_fine_equivalent"], 1),
            "mlups_end_to_end": round(leg["mlups_end_to_end"], 1),
            "ms_per_step": round(leg["ms_per_step"], 3),
            "rebuild_ms_in_window": round(leg["rebuild_ms_in_window"], 2),
            "develop_seconds": leg["develop_seconds"],
            "blocks_per_level": leg["blocks_per_level"],
            "blocks_per_level_end": leg["blocks_per_level_end"],
            "cell_updates_per_step": leg["cell_updates_per_step"],
            "step_roofline_frac": round(leg["step_hbm_gbs_algorithmic"] / HBM_PEAK_GBS, 4),
            "fused_kernel": {"avg_launch_ms": round(fu["ms"] / max(1, fu["launches"]), 4),
                             "achieved_gbs": round(fu["achieved"], 1), "frac": round(fu["achieved"] / HBM_PEAK_GBS, 4),
                             "bytes_per_cell_update": fu["bytes_per_cell_update"]},
        }
        leg["m"].close()
        log(f"large cavity {key}: {out[key]['mlups']:.0f} MLUPS on {leg['blocks_per_level']}")
    out["speedup_fp16_vs_fp32"] = round(out["fp16"]["mlups"] / out["fp32"]["mlups"], 3)
    return out


def storage_study(p, args):
    """FP16 / BF16 pools of f - w_q against FP32 and FP64 on one mesh sequence.

    An FP64 run from rest leads; FP32, FP16 and BF16 runs follow it, taking its refinement marks
    at every rebuild, so all four share every mesh.  After the development the velocity of the leaf
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
    advance_twins(ctx["fp64"], [ctx["fp32"], ctx["fp16"], ctx["bf16"]], 0, args.storage_develop)
    warm_s = time.time() - t0
    errs = {k: velocity_error(ctx["fp64"], ctx[k]) for k in ("fp32", "fp16", "bf16")}
    n_ids = blocks(ctx["fp64"])
    upd = updates_per_step(n_ids)
    out = {"develop_steps": args.storage_develop, "develop_seconds": round(warm_s, 2), "blocks_per_level": n_ids,
           "cell_updates_per_step": upd, "vel_err_vs_fp64": {}, "mlups": {}, "fused_kernel": {}}
    bpp = {"fp64": 8, "fp32": 4, "fp16": 2, "bf16": 2}
    K = min(max(args.steps, 20), p.P_REFINE - 1)  # no rebuild inside the window
    for k in ("fp64", "fp32", "fp16", "bf16"):
        m = ctx[k]
        log(f"storage study: timing {k}")
        m.sync()
        _, step_ms, _ = m.advance_loop(args.storage_develop, K)
        m.sync()
        out["mlups"][k] = round(upd * K / (step_ms * 1e-3) / 1e6, 1)
        fms, nl, fc = roofline_pass(m, max(1, args.roofline_steps // 2))
        alg = 2 * 19 * bpp[k] * fc
        out["fused_kernel"][k] = {"avg_launch_ms": round(fms / max(1, nl), 5),
                                  "achieved_gbs": round(alg / (fms * 1e-3) / 1e9, 1) if fms > 0 else None,
                                  "bytes_per_cell_update": 2 * 19 * bpp[k]}
        if k != "fp64":
            out["vel_err_vs_fp64"][k] = {"rel_l2": float(f"{errs[k]['rel_l2_velocity']:.4g}"),
                                         "max_abs": float(f"{errs[k]['max_abs_velocity']:.4g}")}
    out["timed_steps"] = K
    out["leaf_cells_compared"] = errs["fp16"]["leaf_cells"]
    out["speedup_fp16_vs_fp32"] = round(out["mlups"]["fp16"] / out["mlups"]["fp32"], 3)
    out["fp16_within_1e-3"] = out["vel_err_vs_fp64"]["fp16"]["rel_l2"] <= 1e-3
    for m in ctx.values():
        m.close()
    return out


def learned_study(p, args):
    """The cavity workload with the learned refinement predictor (12-64-64-1 MLP, FP16 weights,
    Adam lr 1e-3, batches of 4096 blocks): features snapshotted every 25 coarse steps and trained on
    the plain criterion's labels at each rebuild over the first 2000 steps, then inference every 25
    steps.  The same run without applying predictions gives the plain method's numbers; the referee is
    the plain criterion at every rebuild (blocks it flags that are still unrefined = missed)."""
    from agal_amd import capi
    from agal_amd.mesh import LbmMesh

    out = {"train_steps": 2000, "inference_steps": 2000, "infer_every": 25, "batch": 4096,
           "controls": "plain_every25: the plain criterion's refinement marks (no coarsening) at the same 25-step "
                       "cadence as the network; plain_dilated: the plain criterion with its refinement marks grown "
                       "by one same-level block at the rebuilds"}
    res = {}
    for use in (0, 1, 2, 3):
        m = LbmMesh(p, args.capacity)
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        opts = capi.learn_opts(train_steps=2000, infer_every=25, batch=4096, adam_steps=8, lr=1e-3, threshold=0.5,
                               pos_weight=4.0, seed=11, use_predictions=use)
        t0 = time.time()
        res[use] = m.advance_loop_learned(0, 4000, opts)
        res[use]["wall_s"] = round(time.time() - t0, 3)
        res[use]["blocks_per_level"] = blocks(m)
        m.close()
        log(f"learned study: use_predictions={use} done in {res[use]['wall_s']} s")
    for key, r in (("plain", res[0]), ("learned", res[1]), ("plain_every25", res[2]), ("plain_dilated", res[3])):
        out[key] = {
            "flagged": r["flagged"], "missed": r["missed"],
            "missed_rate": round(r["missed"] / r["flagged"], 5) if r["flagged"] else None,
            "predicted_refinements": r["predicted_refinements"], "early_rebuilds": r["early_rebuilds"],
            "finest_blocks_end": r["blocks_per_level"][-1],
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
    4 levels refined near the surface, inflow 0.04, Re 200 on the diameter.  Plain bounce-back on the
    body and the first-order momentum exchange on face-cell links that agree with the cell masks
    (AGAL_OPT_LINK_MASKS); C_D = 8 F_x / (U^2 pi D^2) with F the momentum-exchange force reported as
    the reference reports it (block sums x dx^3, total / dx of the finest level).

    First the force on the body in a fluid at rest (one step), with the reference's link rules and
    with links agreeing with the masks.  Then FP64 (vector MRT), FP32 (vector and matrix-core MRT)
    and FP16 storage of f - w (vector and matrix-core MRT) run the same static mesh from the same
    uniform start in lockstep, with a BGK pair (FP64, FP16) as the control; leaf velocities are
    compared with the FP64 run of the same collision every 250 coarse steps, and C_D is averaged
    over coarse steps 250-500 and over the second half of the run (sampled every 10 coarse steps).
    The run ends at 1000 coarse steps (6.25 diameters of travel): the wake reaches the outflow face
    about 1100 coarse steps after the start.  With these MRT rates small differences between runs
    grow by about 10^3 per 500 coarse steps from step 500 on (BGK: about 10 per 500), so the early
    window is where the precisions are compared on equal flows."""
    import math

    from agal_amd import capi
    from agal_amd.mesh import LbmMesh
    from agal_amd.twins import velocity_error

    N, U, Re, D, levels = 64, 0.04, 200.0, 0.16, 4
    tri = capi.icosphere((0.3, 0.5, 0.5), D / 2, 4, 0.05, 20240)
    rates = dict(s6=1.2)
    out = {"workload": "D3Q27 MRT flow past the noisy icosphere (5120 triangles, r 0.08 L +-5%, seed 20240), "
                       "root 64^3, 4 wall levels, inflow 0.04, Re 200, MRT rates s1..s5 parser defaults, s6 1.2 "
                       "(1.98 diverges from the uniform start), plain bounce-back, first-order momentum exchange "
                       "on links agreeing with the cell masks",
           "cd_formula": "8 F_x / (U^2 pi D^2), F = agal_forces (S_ReportForces scaling)",
           "coarse_steps": args.sphere_steps, "mlups": {}, "cd": {}, "fused_kernel": {},
           "mrt_transform": {"fp64": "vector", "fp32": "vector", "fp32_mfma": "matrix cores (bf16 two-term)",
                             "fp16": "vector", "fp16_mfma": "matrix cores (binary16 one term)"},
           "bgk_control": "bgk_fp64 / bgk_fp16: the same flow with BGK collision, compared with each other",
           "control_volume": "cv_fp64: FP64 MRT with the control-volume force (S_FORCE_TYPE 1) over the box "
                             "0.125-0.875 L; its schedule re-averages and re-interpolates the hierarchy after every "
                             "coarse step, which rewrites the GHOST cells (cv_fp64 vs fp64 velocity errors are given "
                             "over all leaf cells and without GHOST cells); its C_D against fp64's momentum exchange"}

    # The control volume of the force type 1 run: the box 0.125 L .. 0.875 L around the body (its edges
    # are root-block edges, so the reference's snapping keeps them).
    cv_box = dict(S_FORCEVOLUME_Xm=0.125, S_FORCEVOLUME_XM=0.875, S_FORCEVOLUME_Ym=0.125,
                  S_FORCEVOLUME_YM=0.875, S_FORCEVOLUME_Zm=0.125, S_FORCEVOLUME_ZM=0.875)

    def params(prec, u, coll=capi.MRT, ftype=0):
        return capi.default_params(velocity_set=capi.D3Q27, precision=prec, collision=coll, Nx=N,
                                   MAX_LEVELS_WALL=levels, MAX_LEVELS_INTERIOR=1, v0=U * D / Re,
                                   bc_preset=capi.BC_FPSC3D, bc_velocity=u, ic_u=u, P_REFINE=1 << 30,
                                   G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=16, G_BIN_LEVELS=10, G_BIN_SPEC=7,
                                   S_BC_TYPE=0, S_FORCE_TYPE=ftype, S_FORCE_ORDER=1, **rates, **cv_box)

    cd_scale = 8.0 / (U * U * math.pi * D * D)
    rest = {}
    for key, fix in (("reference_link_rules", 0), ("links_agree_with_masks", 1)):
        m = LbmMesh(params(capi.FP64, 0.0), args.capacity)
        m.set_option(capi.OPT_LINK_MASKS, fix)
        m.set_geometry(tri)
        m.M_Advance_RefineNearWall()
        m.S_Advance(0)
        F = m.S_ComputeForces()
        rest[key] = [float(f"{cd_scale * x:.3g}") for x in F]
        m.close()
    out["rest_force_in_cd_units"] = rest
    log(f"sphere: force at rest in C_D units {rest}")
    ctx = {}
    runs = (("fp64", capi.FP64, capi.MRT, 0, 0), ("fp32", capi.FP32, capi.MRT, 0, 0),
            ("fp32_mfma", capi.FP32, capi.MRT, 1, 0), ("fp16", capi.FP16, capi.MRT, 0, 0),
            ("fp16_mfma", capi.FP16, capi.MRT, 1, 0), ("bgk_fp64", capi.FP64, capi.BGK, 0, 0),
            ("bgk_fp16", capi.FP16, capi.BGK, 0, 0), ("cv_fp64", capi.FP64, capi.MRT, 0, 1))
    for k, prec, coll, mf, ftype in runs:
        m = LbmMesh(params(prec, U, coll, ftype), args.capacity)
        if mf:
            m.set_option(capi.OPT_MRT_MFMA, 1)
        m.set_option(capi.OPT_LINK_MASKS, 1)
        m.set_geometry(tri)
        t0 = time.time()
        m.M_Advance_RefineNearWall()
        m.sync()
        out.setdefault("near_wall_setup_s", {})[k] = round(time.time() - t0, 3)
        ctx[k] = m
        log(f"sphere {k}: near-wall hierarchy {blocks(m)}")
    n_ids = blocks(ctx["fp64"])
    upd = updates_per_step(n_ids)
    out["blocks_per_level"] = n_ids
    out["cell_updates_per_step"] = upd
    K = args.sphere_steps
    avg_from = K // 2
    ms = {k: 0.0 for k in ctx}
    cds = {k: [] for k in ctx}
    series, cd_series = [], []
    done = 0
    while done < K:
        step = min(10, K - done)
        for k, m in ctx.items():
            _, t, _ = m.advance_loop(done, step)
            ms[k] += t
            F = m.S_ComputeForces()
            cds[k].append(cd_scale * F[0])
        done += step
        if done % 250 == 0 or done == K:
            row = {"step": done}
            for k in ctx:
                if k not in ("fp64", "bgk_fp64"):
                    ref = "bgk_fp64" if k.startswith("bgk") else "fp64"
                    row[k] = float(f"{velocity_error(ctx[ref], ctx[k])['rel_l2_velocity']:.3g}")
                    if k == "cv_fp64":
                        row["cv_fp64_non_ghost"] = float(
                            f"{velocity_error(ctx[ref], ctx[k], exclude_ghost=True)['rel_l2_velocity']:.3g}")
            series.append(row)
            cd_series.append({"step": done, **{k: round(float(sum(cds[k][-25:]) / len(cds[k][-25:])), 4) for k in ctx}})
            log(f"sphere: {done} coarse steps, velocity error vs fp64 {row}, C_D {cd_series[-1]}")
    nwin = (K - avg_from) // 10
    out["cd_early"] = {}
    for k, m in ctx.items():
        out["mlups"][k] = round(upd * K / (ms[k] * 1e-3) / 1e6, 1)
        out["cd"][k] = round(sum(cds[k][-nwin:]) / nwin, 5)
        out["cd_early"][k] = round(sum(cds[k][25:50]) / 25, 5) if len(cds[k]) >= 50 else None
        fms, nl, fc = roofline_pass(m, 5)
        b = 8 if "fp64" in k else (2 if "fp16" in k else 4)
        out["fused_kernel"][k] = {"avg_launch_ms": round(fms / max(1, nl), 5),
                                  "achieved_gbs": round(2 * 27 * b * fc / (fms * 1e-3) / 1e9, 1)
                                  if fms > 0 else None}
    out["cd_window"] = f"mean over coarse steps {avg_from}-{K}, sampled every 10"
    out["cd_early_window"] = "mean over coarse steps 250-500, sampled every 10"
    out["vel_err_vs_fp64_series"] = series
    out["cd_series_250_step_means"] = cd_series
    ref_of = lambda k: "bgk_fp64" if k.startswith("bgk") else "fp64"
    out["cd_rel_err_vs_fp64"] = {k: float(f"{abs(out['cd'][k] / out['cd'][ref_of(k)] - 1):.3g}") for k in ctx
                                 if k not in ("fp64", "bgk_fp64")}
    out["cd_early_rel_err_vs_fp64"] = {k: float(f"{abs(out['cd_early'][k] / out['cd_early'][ref_of(k)] - 1):.3g}")
                                       for k in ctx if k not in ("fp64", "bgk_fp64") and out["cd_early"][k]}
    out["cd_schiller_naumann_unbounded"] = round(24.0 / Re * (1 + 0.15 * Re ** 0.687), 4)
    for m in ctx.values():
        m.close()
    return out


def cpu_baseline(m, p, capacity, upd_per_step, n_ids, budget_s):
    """The CPU restatement stepping the same developed mesh (state copied from the GPU): a serial
    leg and a leg on all host cores (capped at 16, the GPU's share of the box)."""
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import OracleMesh, set_threads
    except Exception as e:  # the checker is not built
        return {"value": None, "unit": "MLUPS", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    nproc = os.cpu_count()
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    o = OracleMesh(p, capacity)
    o.import_state(m.state())

    def leg(threads, budget, max_steps):
        used = set_threads(threads)
        o.S_Advance(0)  # first touch
        steps, t = 0, 0.0
        while t < budget and steps < max_steps:
            t0 = time.time()
            o.S_Advance(0)
            t += time.time() - t0
            steps += 1
        return used, steps, t

    threads, steps, t = leg(max(1, min(ncpu, 16)), 0.6 * budget_s, 400)
    _, s_steps, s_t = leg(1, 0.4 * budget_s, 100)
    o.close()
    return {
        "value": round(upd_per_step * steps / t / 1e6, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} coarse steps of the developed {len(n_ids)}-level mesh {n_ids} (FP32, OpenMP over "
                  f"blocks, {threads} threads), {t:.1f} s; host nproc {nproc}, affinity {ncpu}",
        "serial": {"value": round(upd_per_step * s_steps / s_t / 1e6, 3), "cores": 1,
                   "sample": f"{s_steps} coarse steps of the same mesh on one thread, {s_t:.1f} s"},
    }


if __name__ == "__main__":
    main()
