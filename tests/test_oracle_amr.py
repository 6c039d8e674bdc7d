"""Pins the restatement's block-pool bookkeeping: root grid and boundary IDs by brute force, and
the refine/coarsen rebuild through invariants, the reference's validators and a brute-force
reconstruction of every neighbour table from block levels and coordinates."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from amr_checks import check_state
from oracle_lib import OracleMesh
from parity import ldc_params


def n_root(p, dim):
    ny = int(p.Nx * p.L_fy)
    nz = int(p.Nx * p.L_fz) if dim == 3 else 4
    return (p.Nx // 4, ny // 4, nz // 4 if dim == 3 else 1)


@pytest.mark.parametrize("Q", [9, 19])
@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_root_grid_brute_force(Q, periodic):
    dim = 2 if Q == 9 else 3
    p = ldc_params(vs=Q, N=16, L_fy=0.5, L_fz=0.75, PERIODIC_X=periodic[0], PERIODIC_Y=periodic[1],
                   PERIODIC_Z=periodic[2])
    m = OracleMesh(p, 512)
    st = m.state()
    bad = check_state(st, dim, n_root(p, dim), periodic=periodic[:dim] + (0,) * (3 - dim), n_max=512)
    assert not bad, bad
    onb = st["onb"]
    if periodic == (1, 1, 1):
        assert not onb.any()


@pytest.mark.parametrize("Q", [9, 19])
def test_solution_adaptive_rebuilds_keep_invariants(Q):
    dim = 2 if Q == 9 else 3
    N = 32 if dim == 2 else 16
    p = ldc_params(vs=Q, N=N, Re=60.0 if dim == 3 else 200.0, levels=3, P_REFINE=20)
    cap = 6000
    m = OracleMesh(p, cap)
    m.S_SetIC(0)
    m.M_FreezeRefinedCells(0)
    seen_levels = set()
    for k in range(6):
        m.advance_loop(20 * k, 20)
        st = m.state()
        bad = check_state(st, dim, n_root(p, dim), n_max=cap)
        assert not bad, (k, bad[:5])
        seen_levels |= {L for L, n in enumerate(st["n_ids"]) if n}
    assert seen_levels == {0, 1, 2}


def test_random_refine_coarsen_sequences():
    p = ldc_params(vs=19, N=16, levels=4)
    cap = 20000
    m = OracleMesh(p, cap)
    m.S_SetIC(0)
    rng = np.random.default_rng(11)
    for it in range(12):
        st = m.state()
        ref = st["ref"]
        act = np.nonzero(ref != capi.REF_INACTIVE)[0]
        # eligible as in the criterion (solver_lbm_criterion.cu:575-584): no missing neighbour
        eligible = ~np.any(st["nbr"][:, act] == capi.N_SKIPID, axis=0) & (st["level"][act] < 3)
        leaves = act[np.isin(ref[act], [capi.REF_UNREFINED]) & eligible]
        parents = act[ref[act] == capi.REF_REFINED]
        if leaves.size:
            m.mark_blocks(rng.choice(leaves, min(6, leaves.size), replace=False), capi.REF_MARK_REFINE)
        if parents.size and it % 2:
            m.mark_blocks(rng.choice(parents, min(4, parents.size), replace=False), capi.REF_MARK_COARSEN)
        status, counts = m.M_RefineAndCoarsenBlocks()
        st = m.state()
        bad = check_state(st, 3, n_root(p, 3), n_max=cap)
        assert not bad, (it, counts, bad[:5])


def test_ids_follow_gap_stack_rule():
    """The k-th block marked for refinement (in ID order) takes children
    gap[n_gaps-1-8k] .. +7 from the reversed tail of the gap stack (mesh_amr.cu:2334-2352)."""
    p = ldc_params(vs=19, N=16, levels=2)
    m = OracleMesh(p, 1000)
    m.S_SetIC(0)
    gaps = m.gap_set()
    marked = [40, 3, 17]
    m.mark_blocks(marked, capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    nbc = m.block_ints(capi.BLK_NBR_CHILD)
    for k, b in enumerate(sorted(marked)):
        assert nbc[0, b] == gaps[len(gaps) - 1 - 8 * k]
    assert nbc[0, 3] == 64  # first child of the lowest marked ID is the first free ID
    assert np.array_equal(m.id_set(1), np.arange(64, 64 + 24))


def test_coarsened_children_return_to_gap_stack_reversed_by_level():
    p = ldc_params(vs=19, N=16, levels=3)
    m = OracleMesh(p, 3000)
    m.S_SetIC(0)
    m.mark_blocks([0, 5], capi.REF_MARK_REFINE)
    m.M_RefineAndCoarsenBlocks()
    before = m.gap_set()
    kids0 = m.block_ints(capi.BLK_NBR_CHILD)[0, 0]
    m.mark_blocks([0], capi.REF_MARK_COARSEN)
    m.M_RefineAndCoarsenBlocks()
    after = m.gap_set()
    assert np.array_equal(after[:before.size], before)
    assert np.array_equal(after[before.size:], np.arange(kids0 + 7, kids0 - 1, -1))


def test_pool_exhaustion_cancels_and_reverts_marks():
    p = ldc_params(vs=19, N=8, levels=2)
    m = OracleMesh(p, 8 + 10)
    m.S_SetIC(0)
    m.mark_blocks([0, 1], capi.REF_MARK_REFINE)
    status, counts = m.M_RefineAndCoarsenBlocks()
    assert status == capi.AGAL_ERR_POOL_EXHAUSTED
    st = m.state()
    assert np.all(st["ref"][:8] == capi.REF_UNREFINED)
    assert st["n_ids"] == [8, 0]
    assert not check_state(st, 3, (2, 2, 2), n_max=18)


def test_root_grid_larger_than_pool_is_rejected():
    p = ldc_params(vs=19, N=16)
    with pytest.raises(capi.AgalError) as e:
        OracleMesh(p, 64)
    assert e.value.status == capi.AGAL_ERR_CAPACITY
