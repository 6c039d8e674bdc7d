"""Several contexts stepped on one mesh sequence: storage-precision studies.

The solution-based rebuild decides refinement from each context's own velocity field, so two runs
in different precisions may refine different blocks and can no longer be compared cell by cell.
`advance_twins` runs the body of M_AdvanceLoop (src/mesh/mesh_advance.cu:421-463) on a lead
context and on followers that take the lead's refinement marks at every rebuild: the lead
evaluates the criterion (solver_lbm_criterion.cu:594-613), its marks are copied into the
followers' ref tables, and every context then runs the same M_RefineAndCoarsenBlocks.  The
rebuild is integer bookkeeping driven only by the marks, so the block pools stay identical.
"""
from __future__ import annotations

import numpy as np

from . import capi
from .mesh import LbmMesh


def _prepare_rebuild(m: LbmMesh):
    """The field updates M_Advance_RefineWithSolution makes before the criterion
    (src/mesh/mesh_advance.cu:204-217)."""
    s = m.sizes
    ML, L0 = s.max_levels, m.params.N_LEVEL_START
    for L in range(ML - 2, L0 - 1, -1):
        m.S_RefreshVariables(L + 1)
        m.S_Average(L, capi.AVERAGE_GRID)
    for L in range(L0, ML - 1):
        m.S_RefreshVariables(L)
        m.S_Interpolate(L, capi.INTERP_INTERFACE)


def advance_twins(lead: LbmMesh, followers: list, i0: int, n: int) -> list:
    """Coarse steps [i0, i0+n) on every context; returns the rebuild statuses of the lead."""
    ctxs = [lead, *followers]
    ML, L0 = lead.sizes.max_levels, lead.params.N_LEVEL_START
    P = lead.params.P_REFINE
    out = []
    for i in range(i0, i0 + n):
        for m in ctxs:
            for _ in range(1 << L0):
                m.S_Advance(L0)
        if ML > 1 and ML != L0 + 1 and (i + 1) % P == 0 and i > 0:
            for m in ctxs:
                _prepare_rebuild(m)
            lead.S_ComputeRefCriteria()
            marks = lead.block_ints(capi.BLK_REF)
            for m in followers:
                m.set_block_ints(capi.BLK_REF, marks)
            res = [m.M_RefineAndCoarsenBlocks() for m in ctxs]
            if any(r[0] != res[0][0] for r in res):
                raise RuntimeError(f"rebuild statuses differ between contexts: {res}")
            out.append(res[0][0])
    return out


def velocity_error(ref: LbmMesh, other: LbmMesh, exclude_ghost: bool = False) -> dict:
    """Relative L2 velocity difference over the active, non-solid leaf cells of two contexts on one
    mesh; with `exclude_ghost` the GHOST cells (coarse-to-fine interpolation targets, whose values
    between steps depend on when they were last interpolated) are left out as well."""
    s = ref.refresh_sizes()
    ML = s.max_levels
    for m in (ref, other):
        for L in range(ML):
            m.S_RefreshVariables(L)
    nb = int(s.id_max[ML])
    cpb = s.cells_per_block
    ref_ids = ref.block_ints(capi.BLK_REF, nb)
    if not np.array_equal(ref_ids, other.block_ints(capi.BLK_REF, nb)):
        raise RuntimeError("the contexts do not share one mesh")
    first_child = ref.block_ints(capi.BLK_NBR_CHILD, nb)[0]
    leaf = (ref_ids != capi.REF_INACTIVE) & (first_child < 0)
    cm = ref.cell_masks(nb)
    cells = np.repeat(leaf, cpb) & (cm != capi.CELL_SOLID)
    if exclude_ghost:
        cells &= cm != capi.CELL_GHOST
    a = ref.macros(nb)[1:1 + s.n_dim][:, cells].astype(np.float64)
    b = other.macros(nb)[1:1 + s.n_dim][:, cells].astype(np.float64)
    err = float(np.sqrt(np.sum((a - b) ** 2) / np.sum(a ** 2)))
    return {"rel_l2_velocity": err, "max_abs_velocity": float(np.max(np.abs(a - b))),
            "leaf_cells": int(cells.sum())}
