"""Pins the restatement's multi-level advance (Solver_LBM::S_Advance, solver_lbm_advance.cu:30-140)
and the swap-stream guard (Cu_Stream, solver_lbm_kernels.cu:597-629) against closed forms and
brute force that do not go through the restatement's own code paths.

* Stream guard: one in-place stream of a field with non-positive values, zeros and solid cells,
  checked against a dense numpy evaluation of the rule "swap slot p of x with slot pb(p) of
  x + c_p iff x is not SOLID and the partner's value is > 0" (`kernels.cu:624-628`), with links
  found from cell coordinates.
* Multi-level plane Couette flow: walls on z (moving at z = Lz), periodic x and y, a refined band
  in the interior on two nested levels.  The exact steady state is f_q = feq_q(u(z)) +
  g_q, u = U z / Lz, with the non-equilibrium part g_q = -3 tau_L w_q c_qx c_qz (du/dz)_L constant
  per level (Chapman-Enskog is exact for a linear velocity field under BGK; half-way bounce-back
  with the moving-wall correction is exact for Couette).  Interpolation rescales g by
  tau_{L+1}/tau_L in the reference's units (tau_L = 3 v0 + dx_L / 2), which is exactly the ratio of
  g between the levels; linear interpolation and 2^d-averaging reproduce linear fields.  So the
  exact state is kept step after step up to the quadratic term of feq (O((U dx / Lz)^2)).  A
  missing or wrong rescale, a dropped average or interpolation, or wrong stencil weights move
  interface cells by O(g) and fail.
* The recursion of S_Advance against the call sequence of solver_lbm_advance.cu:30-140 written
  out flat for two and three levels, through the restatement's separate collide / boundary /
  stream / interpolate / average entries: bit-identical populations.
* Mistakes planted in the restatement (Oracle::plant: dropped rescales, shifted stencils, a
  dropped child cell, skipped or reordered interface operations, an unguarded swap) each fail at
  least one of these pins or the interpolate-then-average round trip on a linear field.
"""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
import oracle_lib as O
from oracle_lib import OracleMesh
from parity import ldc_params
from test_oracle_physics import cell_positions, feq_np, interpolate_average_error


# ---- stream guard: dense brute force ------------------------------------------------------------
def guard_case(Q, N, periodic, plant=0, seed=0):
    """One stream of a signed field with solid cells; returns (restatement result, brute force,
    branch counts)."""
    dim = 2 if Q == 9 else 3
    p = ldc_params(vs=Q, N=N, prec=capi.FP64, bc_preset=capi.BC_NONE,
                   PERIODIC_X=periodic, PERIODIC_Y=periodic, PERIODIC_Z=periodic)
    m = OracleMesh(p, 64)
    m.plant(plant)
    m.S_SetIC(0)
    nc = m.populations().shape[1]
    c, _, pb, _ = O.velset(Q)
    rng = np.random.default_rng(100 + Q + periodic + seed)
    # Distinct magnitudes per (slot, cell); about 30% negative, a few exact zeros.
    mag = 1.0 + np.arange(Q * nc, dtype=np.float64).reshape(Q, nc) / (Q * nc)
    slots = mag * np.where(rng.random((Q, nc)) < 0.3, -1.0, 1.0)
    slots[rng.random((Q, nc)) < 0.03] = 0.0
    # About 10% SOLID cells; their own slots keep arbitrary signs, so a solid partner with a
    # positive value still swaps with a fluid initiator (the rule reads only the initiator's mask).
    solid = rng.random(nc) < 0.1
    masks = m.cell_masks()
    masks[solid] = capi.CELL_SOLID
    m.set_cell_masks(masks)
    m.set_populations(slots[pb[:Q]])  # download/upload row q is slot pb(q)
    m.stream(0)
    after = m.populations()[pb[:Q]]

    pos = np.rint(cell_positions(m) - 0.5).astype(int)
    n = np.array([N, N, N if dim == 3 else 1])
    lin = pos[0] + n[0] * (pos[1] + n[1] * pos[2])
    where = np.full(n.prod(), -1)
    where[lin] = np.arange(nc)
    expect = slots.copy()
    half = [s for s in range(1, Q) if O.in_half_set(dim, s)]
    assert len(half) == (Q - 1) // 2
    counts = {"swapped": 0, "kept": 0, "kept_nonpositive_partner": 0, "kept_solid": 0, "solid_partner_swapped": 0}
    for s in half:
        tgt = pos + c[:, s][:, None]
        inside = np.all((tgt >= 0) & (tgt < n[:, None]), axis=0)
        if periodic:
            tgt = tgt % n[:, None]
            inside[:] = True
        partner = np.full(nc, -1)
        tl = tgt[0, inside] + n[0] * (tgt[1, inside] + n[1] * tgt[2, inside])
        partner[inside] = where[tl]
        has = partner >= 0
        f_pb = np.where(has, slots[pb[s], np.maximum(partner, 0)], -1.0)
        swap = (~solid) & (f_pb > 0) & has
        x = np.nonzero(swap)[0]
        y = partner[x]
        expect[s, x] = slots[pb[s], y]
        expect[pb[s], y] = slots[s, x]
        counts["swapped"] += x.size
        counts["kept"] += int(np.sum(~swap))
        counts["kept_nonpositive_partner"] += int(np.sum(has & ~solid & (f_pb <= 0)))
        counts["kept_solid"] += int(np.sum(has & solid))
        counts["solid_partner_swapped"] += int(np.sum(swap & solid[np.maximum(partner, 0)]))
    return after, expect, counts


@pytest.mark.parametrize("Q,N", [(9, 16), (19, 8), (27, 8)])
@pytest.mark.parametrize("periodic", [1, 0])
def test_stream_guard_brute_force(Q, N, periodic):
    after, expect, counts = guard_case(Q, N, periodic)
    # the field exercises every branch of the rule
    assert all(v > 0 for v in counts.values()), counts
    assert np.array_equal(after, expect)


# ---- multi-level plane Couette flow -------------------------------------------------------------
def couette_mesh(Q, interp, U, tau0=0.8, plant=0):
    """Walls normal to the last axis (moving lid at its upper face), periodic otherwise, aspect
    ratio 4 along the walls' normal; the root block layers 3 and 4 refined once and the middle two
    level-1 layers refined again."""
    dim = 2 if Q == 9 else 3
    N = 8
    nu_lat = (tau0 - 0.5) / 3.0
    kw = dict(PERIODIC_X=1, interp_order=interp, P_REFINE=1 << 30, v0=nu_lat / N, bc_velocity=U)
    if dim == 3:
        kw.update(PERIODIC_Y=1, L_fz=4.0)
    else:
        kw.update(L_fy=4.0)
    p = ldc_params(vs=Q, N=N, levels=3, prec=capi.FP64, **kw)
    m = OracleMesh(p, 4096)
    m.S_SetIC(0)
    ax = dim - 1
    X = m.coords()
    lev = m.block_ints(capi.BLK_LEVEL)
    ids = np.nonzero((lev == 0) & ((np.isclose(X[ax], 1.5)) | np.isclose(X[ax], 2.0)))[0]
    m.mark_blocks(ids, capi.REF_MARK_REFINE)
    assert m.M_RefineAndCoarsenBlocks()[0] == capi.AGAL_OK
    X = m.coords()
    lev = m.block_ints(capi.BLK_LEVEL)
    ref = m.block_ints(capi.BLK_REF)
    ids = np.nonzero((lev == 1) & (ref != capi.REF_INACTIVE) & (np.isclose(X[ax], 1.75) | np.isclose(X[ax], 2.0)))[0]
    m.mark_blocks(ids, capi.REF_MARK_REFINE)
    assert m.M_RefineAndCoarsenBlocks()[0] == capi.AGAL_OK
    m.plant(plant)
    s = m.refresh_sizes()
    return m, p, [int(s.n_ids[L]) for L in range(3)]


def couette_exact(m, p, Q, U):
    """Exact populations (physical order) and velocity of every cell up to id_max."""
    dim = 2 if Q == 9 else 3
    ax = dim - 1
    Lz = 4.0  # domain length along the wall normal (L_c = 1, aspect 4)
    pos = cell_positions(m) / p.Nx  # cell centres, domain units
    nc = pos.shape[1]
    lev = np.repeat(m.block_ints(capi.BLK_LEVEL), 64 if dim == 3 else 16)
    u = np.zeros((3, nc))
    u[0] = U * pos[ax] / Lz
    f = feq_np(Q, np.ones(nc), u)
    c, w, _, _ = O.velset(Q)
    dx0 = 1.0 / p.Nx
    tau_ref = 3 * p.v0 + 0.5 * dx0 / 2.0 ** lev  # tau_L in the reference's units = tau_lat,L * dx_L
    dudz = U / Lz
    g = -3.0 * (w[:Q, None] * c[0, :Q, None] * c[ax, :Q, None]) * (dudz * tau_ref)[None, :]
    return f + g, u


def couette_errors(Q, interp, plant=0, U=1e-4, steps=200):
    """Per level: max |u_x - U z / Lz| / U and max |u_y, u_z| / U over all active cells after
    `steps` coarse steps from the exact state (ghost cells refilled by the interpolation that
    opens the next coarse step)."""
    m, p, n_ids = couette_mesh(Q, interp, U, plant=plant)
    assert all(n > 0 for n in n_ids), n_ids
    cpb = 64 if Q != 9 else 16
    f0, u_exact = couette_exact(m, p, Q, U)
    m.set_populations(f0)
    active = np.repeat(m.block_ints(capi.BLK_REF) != capi.REF_INACTIVE, cpb)
    lev = np.repeat(m.block_ints(capi.BLK_LEVEL), cpb)
    m.advance_loop(0, steps)
    # Ghost cells of the refined levels stream without their coarse-side inflow during the
    # sub-steps; the interpolation that opens the next coarse step refills them.
    for L in range(2):
        m.S_Interpolate(L, capi.INTERP_INTERFACE)
    for L in range(3):
        m.S_RefreshVariables(L)
    mac = m.macros()
    out = []
    for L in range(3):
        sel = active & (lev == L)
        assert sel.any()
        out.append((float(np.max(np.abs(mac[1][sel] - u_exact[0][sel]))) / U,
                    float(np.max(np.abs(mac[2:4][:, sel]))) / U))
    return out


@pytest.mark.parametrize("Q", [9, 19])
@pytest.mark.parametrize("interp", [capi.INTERP_LINEAR, capi.INTERP_CUBIC])
def test_multilevel_couette_keeps_exact_state(Q, interp):
    # The quadratic feq terms are interpolated / averaged with an error below (U dx0 / Lz)^2,
    # under 1e-8 of U here (about 1e-11 is measured); a wrong interface operation moves cells by
    # O(g), 1e-4 to 1e-1 of U.  The reference's 3D cubic table carries round-off terms up to
    # 1.7e-11 (inc/util_interp.h), up to about 1e-8 of U on the interpolated cells.
    tol = 1e-9 if (interp == capi.INTERP_LINEAR or Q == 9) else 2e-8
    for L, (e_par, e_perp) in enumerate(couette_errors(Q, interp)):
        assert e_par < tol and e_perp < tol, (L, e_par, e_perp)


# ---- the recursion of S_Advance against its flat call sequence -----------------------------------
def flat_schedule(levels):
    """The kernel sequence of one coarse step from level 0 as solver_lbm_advance.cu:30-140 issues
    it (forces off), written out flat: ("I", L) S_Interpolate(V_INTERP_ADVANCE), ("A", L) S_Average
    (V_AVERAGE_ADVANCE), ("C", L) S_Collide + S_ImposeBC + S_Stream."""
    if levels == 2:
        return [("I", 0), ("C", 0), ("C", 1), ("C", 1), ("A", 0)]
    if levels == 3:
        fine = [("C", 2), ("C", 2)]
        return ([("I", 0), ("C", 0),
                 ("I", 1), ("C", 1)] + fine + [("A", 1), ("I", 1), ("C", 1)] + fine + [("A", 1),
                ("A", 0)])
    raise ValueError(levels)


def schedule_case(Q, levels, interp, plant=0):
    """(populations after 3 coarse steps of S_Advance, the same after the flat sequence)."""
    dim = 2 if Q == 9 else 3
    N = 16 if dim == 2 else 8
    p = ldc_params(vs=Q, N=N, levels=levels, prec=capi.FP64, Re=100.0, interp_order=interp, P_REFINE=1 << 30)
    avg_mode = capi.AVERAGE_BLOCK if interp == capi.INTERP_CUBIC else capi.AVERAGE_INTERFACE
    ms = []
    for _ in range(2):
        m = OracleMesh(p, 4096)
        m.S_SetIC(0)
        # Refine a corner region twice (nested), so both interface kinds exist on every level.
        for L in range(levels - 1):
            X = m.coords()
            lev = m.block_ints(capi.BLK_LEVEL)
            ref = m.block_ints(capi.BLK_REF)
            h = 4.0 / N / 2 ** L
            sel = (lev == L) & (ref == capi.REF_UNREFINED) & np.all(X[:dim] >= 0.25 - 1e-9, axis=0) & \
                np.all(X[:dim] < 0.25 + 2 * h * 2 ** (levels - 1 - L) - 1e-9, axis=0)
            ids = np.nonzero(sel)[0]
            assert ids.size > 0
            m.mark_blocks(ids, capi.REF_MARK_REFINE)
            assert m.M_RefineAndCoarsenBlocks()[0] == capi.AGAL_OK
        ms.append(m)
    n_ids = [int(ms[0].sizes.n_ids[L]) for L in range(levels)]
    assert all(n > 0 for n in n_ids), n_ids
    # A developed, non-trivial state on both copies.
    ms[0].advance_loop(0, 5)
    f = ms[0].populations()
    rng = np.random.default_rng(3)
    f = f * (1 + 1e-3 * rng.uniform(-1, 1, f.shape))
    for m in ms:
        m.set_populations(f)
        m.plant(plant)
    for _ in range(3):
        ms[0].S_Advance(0)
        for op, L in flat_schedule(levels):
            if op == "I":
                ms[1].S_Interpolate(L, capi.INTERP_INTERFACE)
            elif op == "A":
                ms[1].S_Average(L, avg_mode)
            else:
                ms[1].collide(L)
                ms[1].impose_bc(L)
                ms[1].stream(L)
    return ms[0].populations(), ms[1].populations()


@pytest.mark.parametrize("Q,levels,interp", [(19, 2, capi.INTERP_LINEAR), (19, 3, capi.INTERP_LINEAR),
                                             (9, 3, capi.INTERP_LINEAR), (19, 3, capi.INTERP_CUBIC)])
def test_advance_recursion_matches_flat_schedule(Q, levels, interp):
    a, b = schedule_case(Q, levels, interp)
    assert np.array_equal(a, b)


# ---- each pin above catches a planted mistake -----------------------------------------------------
PLANTS = {
    1: "interpolation without the tau rescale",
    2: "averaging without the tau rescale",
    3: "linear interpolation at offsets +1/4",
    4: "averaging drops a child cell",
    5: "advance skips the mid-cycle interpolation",
    6: "advance skips the average after the second sub-cycle",
    7: "root averages before the finer levels advance",
    8: "stream without the f_pb > 0 guard",
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pins_catch_planted_mistakes(plant):
    """With a known mistake planted in the restatement, at least one of the closed-form /
    brute-force pins of this file fails (and which ones is recorded in the assertion text)."""
    caught = []
    errs = couette_errors(19, capi.INTERP_LINEAR, plant=plant)
    if max(max(e) for e in errs) >= 1e-9:
        caught.append("couette")
    a, b = schedule_case(19, 3, capi.INTERP_LINEAR, plant=plant)
    if not np.array_equal(a, b):
        caught.append("flat schedule")
    after, expect, _ = guard_case(19, 8, 1, plant=plant)
    if not np.array_equal(after, expect):
        caught.append("stream brute force")
    if interpolate_average_error(plant) >= 1e-14:
        caught.append("interpolate-average round trip")
    assert caught, f"planted mistake {plant} ({PLANTS[plant]}) passed every pin"
