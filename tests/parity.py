"""Helpers shared by the parity tests: build the same case on the HIP library and on the CPU
restatement, and compare their full states (integer state bit-exactly, populations bit-exactly
under the shared FP contraction rule)."""
from __future__ import annotations

import numpy as np

from agal_amd import capi


def ldc_params(vs=capi.D3Q19, N=16, Re=50.0, U=0.05, levels=1, prec=capi.FP32, **kw):
    """Lid-driven cavity: tau from nu = U*N/Re, lid on the top face, fluid at rest."""
    nu = U * N / Re
    base = dict(velocity_set=vs, precision=prec, Nx=N, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=levels,
                v0=nu / N, bc_preset=capi.BC_LDC2D if vs == capi.D2Q9 else capi.BC_LDC3D, bc_velocity=U,
                P_REFINE=20, N_REFINE_START=-2.0, N_REFINE_INC=1.0, N_REFINE_MAX=1.0)
    base.update(kw)
    return capi.default_params(**base)


def sphere_params(vs=capi.D3Q27, N=16, levels=3, Re=200.0, U=0.04, D=0.16, prec=capi.FP32,
                  collision=capi.MRT, **kw):
    """External flow past a body of diameter D (domain units): inflow / moving side walls at U and
    an outflow face (the FPSC3D hook), uniform initial flow, nu = U*D/Re, the wall levels refined
    near the body."""
    base = dict(velocity_set=vs, precision=prec, collision=collision, Nx=N, MAX_LEVELS_WALL=levels,
                MAX_LEVELS_INTERIOR=1, v0=U * D / Re, bc_preset=capi.BC_FPSC3D, bc_velocity=U, ic_u=U,
                P_REFINE=1000)
    base.update(kw)
    return capi.default_params(**base)


def body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.05, seed=1):
    """The seeded noisy icosphere, (n, 3, 3) float32."""
    return capi.icosphere(center, radius, subdivisions, noise, seed)


def active_mask(state):
    return state["ref"] != capi.REF_INACTIVE


def compare_states(a: dict, b: dict, *, check_f=True, rtol=None) -> list:
    """Returns a list of human-readable mismatches (empty when the states agree)."""
    bad = []
    for k in ("n_ids", "id_max", "n_gaps"):
        if a[k] != b[k]:
            bad.append(f"{k}: {a[k]} != {b[k]}")
    if bad:
        return bad
    for L, (x, y) in enumerate(zip(a["id_sets"], b["id_sets"])):
        if not np.array_equal(x, y):
            bad.append(f"id_set[{L}] differs")
    if not np.array_equal(a["gap_set"], b["gap_set"]):
        bad.append("gap_set differs")
    act = active_mask(a) & active_mask(b)
    if not np.array_equal(active_mask(a), active_mask(b)):
        bad.append("active block sets differ")
    for k in ("ref", "level", "onb", "blk_mask"):
        if not np.array_equal(a[k][act], b[k][act]):
            idx = np.nonzero((a[k] != b[k]) & act)[0][:5]
            bad.append(f"{k} differs at blocks {idx.tolist()}: {a[k][idx].tolist()} vs {b[k][idx].tolist()}")
    for k in ("nbr", "nbr_child"):
        if not np.array_equal(a[k][:, act], b[k][:, act]):
            d = np.nonzero((a[k] != b[k]) & act[None, :])
            bad.append(f"{k} differs (p, block) {list(zip(d[0][:5].tolist(), d[1][:5].tolist()))}")
    if not np.array_equal(a["coords"][:, act], b["coords"][:, act]):
        bad.append("coords differ")
    cpb = a["cell_mask"].size // max(1, act.size)
    cact = np.repeat(act, cpb)
    if not np.array_equal(a["cell_mask"][cact], b["cell_mask"][cact]):
        bad.append("cell masks differ")
    if check_f:
        fa, fb = a["f"][:, cact], b["f"][:, cact]
        if rtol is None:
            if not np.array_equal(fa, fb, equal_nan=True):
                d = np.nonzero(~((fa == fb) | (np.isnan(fa) & np.isnan(fb))))
                cells = np.unique(d[1])
                bad.append(f"populations differ in {cells.size} cells; first (q,cell) "
                           f"{list(zip(d[0][:4].tolist(), d[1][:4].tolist()))} "
                           f"max|diff|={np.nanmax(np.abs(fa - fb)):.3e}")
        else:
            err = np.max(np.abs(fa - fb) / np.maximum(np.abs(fb), 1e-30))
            if not err <= rtol:
                bad.append(f"populations differ: max rel err {err:.3e} > {rtol}")
    return bad
