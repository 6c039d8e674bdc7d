"""Pins the restatement's control-volume force (S_FORCE_TYPE 1, Cu_ComputeForcesCV,
solver_lbm_kernels_forces.cu:13-246, scheduled by solver_lbm_advance.cu:45-68 / :139-149).

The reference ships no force outputs for it, so the pins are the balance it states:
* without a body the momentum in the box changes only by what crosses its faces, so on one level the
  reported force is zero to rounding for any flow (a wrong face test, direction, sign, cell volume or
  collision leaves a flux-sized residue; two such mistakes are planted and caught);
* with a body at rest under plain bounce-back on links that agree with the cell masks, the momentum
  the fluid in the box loses per step is exactly what the links exchange, so the control-volume
  force equals the momentum-exchange force of the same run to rounding, step by step, which pins
  the sign and the 1/dx factor as well.
"""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from oracle_lib import OracleMesh, set_threads
from test_oracle_physics import feq_np

BOX = dict(S_FORCEVOLUME_Xm=0.3, S_FORCEVOLUME_XM=0.8, S_FORCEVOLUME_Ym=0.2, S_FORCEVOLUME_YM=0.7,
           S_FORCEVOLUME_Zm=0.25, S_FORCEVOLUME_ZM=0.9)


def periodic_case(Q, plant=0, seed=3):
    """One level, periodic in every direction, no boundary hook; a random smooth flow with random
    non-equilibrium parts."""
    dim = 2 if Q == 9 else 3
    N = 32 if dim == 2 else 16
    p = capi.default_params(velocity_set=Q, precision=capi.FP64, Nx=N, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=1,
                            v0=0.02 / N, bc_preset=capi.BC_NONE, PERIODIC_X=1, PERIODIC_Y=1,
                            PERIODIC_Z=1 if dim == 3 else 0, P_REFINE=1 << 30, S_FORCE_TYPE=1, **BOX)
    m = OracleMesh(p, 4096)
    if plant:
        m.plant(plant)
    m.S_SetIC(0)
    nc = m.populations().shape[1]
    rng = np.random.default_rng(seed)
    rho = 1 + 0.05 * rng.uniform(-1, 1, nc)
    u = np.zeros((3, nc))
    u[:dim] = 0.05 * rng.uniform(-1, 1, (dim, nc))
    f = feq_np(Q, rho, u)
    f *= 1 + 0.05 * rng.uniform(-1, 1, f.shape)
    m.set_populations(f)
    return m, N


# The face fluxes of these flows: about rho c_s^2 per unit face area, times the box faces (~1 in
# domain units) -- the force of a flux-sized mistake is 1e-4 or more, rounding leaves 1e-17.
RESIDUE_TOL, MISTAKE_MIN = 1e-13, 1e-5


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_balance_closes_without_a_body(Q):
    set_threads(8)
    m, N = periodic_case(Q)
    for _ in range(3):
        m.S_Advance(0)
        F = m.S_ComputeForces()
        assert np.all(np.isfinite(F))
        assert np.abs(F).max() < RESIDUE_TOL, F
    m.close()


@pytest.mark.parametrize("plant", [9, 10])
def test_planted_mistakes_leave_a_residue(plant):
    set_threads(8)
    m, N = periodic_case(19, plant=plant)
    m.S_Advance(0)
    F = m.S_ComputeForces()
    assert np.abs(F).max() > MISTAKE_MIN, F
    m.close()


KEYS = ["S_FORCEVOLUME_" + k for k in ("Xm", "XM", "Ym", "YM", "Zm", "ZM")]
WIDE = (0.125, 0.875, 0.125, 0.875, 0.125, 0.875)


def sphere_forces(ftype, box, steps=6):
    """The noisy icosphere on one 32^3 level, D3Q19 FP64, uniform inflow U = 0.04, plain bounce-back
    on links made to agree with the cell masks; the force after each coarse step."""
    tri = capi.icosphere((0.45, 0.5, 0.5), 0.1, 3, 0.05, 7)
    U = 0.04
    p = capi.default_params(velocity_set=capi.D3Q19, precision=capi.FP64, collision=capi.BGK, Nx=32,
                            MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=1, v0=U * 0.2 / 100.0,
                            bc_preset=capi.BC_FPSC3D, bc_velocity=U, ic_u=U, P_REFINE=1 << 30,
                            G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=4, G_BIN_LEVELS=3, G_BIN_SPEC=1,
                            S_BC_TYPE=0, S_FORCE_TYPE=ftype, S_FORCE_ORDER=1, **dict(zip(KEYS, box)))
    m = OracleMesh(p, 4096)
    m.set_option(capi.OPT_LINK_MASKS, 1)
    m.set_geometry(tri)
    assert m.M_Advance_RefineNearWall() == capi.AGAL_OK
    fs = []
    for _ in range(steps):
        m.S_Advance(0)
        fs.append(m.S_ComputeForces())
    m.close()
    return np.array(fs)


def test_equals_the_momentum_exchange_around_a_body():
    """Two runs that differ only in S_FORCE_TYPE follow the same populations; with the box around
    the body their forces agree to rounding at every reported step."""
    set_threads(8)
    mea, cv = sphere_forces(0, WIDE), sphere_forces(1, WIDE)
    assert np.abs(mea[:, 0]).min() > 1e-6  # drag of the impulsively started flow
    np.testing.assert_allclose(cv, mea, rtol=1e-8, atol=1e-8 * np.abs(mea).max())


def test_box_is_snapped_to_root_block_edges():
    """Bounds are 4 int(X Nx / 4) dx0 (solver_lbm_init.cu:36-41).  A box whose lower x face cuts the
    body (at x = 0.5, a root-block edge) gives the same force, bit for bit, from the fractions 0.5
    and 0.6 (three cells further in, still inside the same block row), and a force other than the
    body's."""
    set_threads(8)
    a = sphere_forces(1, (0.5,) + WIDE[1:], steps=3)
    b = sphere_forces(1, (0.6, 0.99) + WIDE[2:], steps=3)
    np.testing.assert_array_equal(a, b)
    whole = sphere_forces(1, WIDE, steps=3)
    assert np.abs(a - whole).max() > 1e-3 * np.abs(whole).max()


def test_schedule_leaves_the_flow_outside_ghost_cells_unchanged():
    """On three levels the second half follows a full average (GRID mode) and the interface
    interpolations of every level (solver_lbm_advance.cu:60-68).  Those rewrite GHOST cells and
    cells of blocks with children; the next coarse step re-interpolates the GHOST cells before any
    sub-step reads them, so on the leaf cells that are not GHOST the flow of a force-type-1 run is
    bit for bit the flow of the same run with force type 0."""
    set_threads(8)
    runs = []
    for ftype in (0, 1):
        p = capi.default_params(velocity_set=capi.D3Q19, precision=capi.FP64, Nx=16, MAX_LEVELS_WALL=1,
                                MAX_LEVELS_INTERIOR=3, v0=0.05 / 100.0, bc_preset=capi.BC_LDC3D, bc_velocity=0.05,
                                P_REFINE=1 << 30, S_FORCE_TYPE=ftype, **BOX)
        m = OracleMesh(p, 8192)
        m.S_SetIC(0)
        for _ in range(20):
            m.S_Advance(0)
        m.mark_blocks([5, 6, 9, 10, 21, 22, 25, 26, 42], capi.REF_MARK_REFINE)
        m.M_RefineAndCoarsenBlocks()
        for _ in range(20):
            m.S_Advance(0)
        runs.append(m)
    a, b = runs
    s = a.refresh_sizes()
    nb = int(s.id_max[s.max_levels])
    for m in runs:
        for L in range(s.max_levels):
            m.S_RefreshVariables(L)
    leaf = (a.block_ints(capi.BLK_REF, nb) != capi.REF_INACTIVE) & (a.block_ints(capi.BLK_NBR_CHILD, nb)[0] < 0)
    cm = a.cell_masks(nb)
    cells = np.repeat(leaf, s.cells_per_block) & (cm != capi.CELL_SOLID)
    ghost = cells & (cm == capi.CELL_GHOST)
    ua, ub = a.macros(nb)[1:4], b.macros(nb)[1:4]
    assert s.n_ids[1] > 0 and ghost.any()
    np.testing.assert_array_equal(ua[:, cells & ~ghost], ub[:, cells & ~ghost])
    assert np.abs(ua[:, ghost] - ub[:, ghost]).max() > 0  # the GHOST cells were rewritten
    for m in runs:
        m.close()
