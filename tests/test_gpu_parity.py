"""GPU parity: the HIP library against the CPU restatement, same inputs, full state compared.

Populations are compared bit for bit: both sides evaluate the reference's expressions in the same
order with FMA contraction off.  Integer state (IDs, gap stack, neighbour tables, refinement
states, masks) must be identical."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from oracle_lib import OracleMesh
from parity import compare_states, ldc_params

pytestmark = pytest.mark.gpu


def pair(p, cap):
    return LbmMesh(p, cap), OracleMesh(p, cap)


def both(ms, fn):
    for m in ms:
        fn(m)


def assert_same(g, o, **kw):
    bad = compare_states(g.state(), o.state(), **kw)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_d2q9_uniform_ldc(prec):
    p = ldc_params(vs=capi.D2Q9, N=32, Re=100.0, levels=1, prec=prec)
    g, o = pair(p, 256)
    both((g, o), lambda m: m.S_SetIC(0))
    for _ in range(150):
        both((g, o), lambda m: m.S_Advance(0))
    assert_same(g, o)


@pytest.mark.parametrize("vs,N", [(capi.D3Q19, 16), (capi.D3Q27, 8)])
@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_d3_uniform_ldc(vs, N, prec):
    p = ldc_params(vs=vs, N=N, Re=20.0, levels=1, prec=prec)
    g, o = pair(p, 512)
    both((g, o), lambda m: m.S_SetIC(0))
    for _ in range(40):
        both((g, o), lambda m: m.S_Advance(0))
    assert_same(g, o)


def test_d3q19_forced_refinement_two_levels():
    p = ldc_params(vs=capi.D3Q19, N=16, Re=30.0, levels=2)
    g, o = pair(p, 2048)
    both((g, o), lambda m: m.S_SetIC(0))
    for _ in range(10):
        both((g, o), lambda m: m.S_Advance(0))
    ids = [5, 6, 9, 10, 21, 22, 25, 26, 42]
    both((g, o), lambda m: m.mark_blocks(ids, capi.REF_MARK_REFINE))
    rg = g.M_RefineAndCoarsenBlocks()
    ro = o.M_RefineAndCoarsenBlocks()
    assert rg == ro
    assert_same(g, o)
    for _ in range(15):
        both((g, o), lambda m: m.S_Advance(0))
    assert_same(g, o)


@pytest.mark.parametrize("vs", [capi.D3Q19, capi.D3Q27])
def test_solution_adaptive_three_levels(vs):
    p = ldc_params(vs=vs, N=16, Re=40.0, levels=3, P_REFINE=15)
    g, o = pair(p, 8192)
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.M_FreezeRefinedCells(0))
    for k in range(4):
        rg = g.advance_loop(15 * k, 15)
        ro = o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[2] > 0


@pytest.mark.parametrize("vs,N", [(capi.D3Q19, 8), (capi.D2Q9, 16)])
def test_solution_adaptive_four_levels(vs, N):
    """Four levels: deeper sub-cycling, where held-back second sub-steps of two coarse levels share
    launches with finer levels' first sub-steps and their interpolations."""
    p = ldc_params(vs=vs, N=N, Re=40.0, levels=4, P_REFINE=10, N_REFINE_START=-3.0)
    g, o = pair(p, 16384)
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.M_FreezeRefinedCells(0))
    for k in range(4):
        rg = g.advance_loop(10 * k, 10)
        ro = o.advance_loop(10 * k, 10)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[3] > 0


@pytest.mark.parametrize("vs,N,prec", [(capi.D3Q19, 16, capi.FP32), (capi.D3Q19, 16, capi.FP64),
                                       (capi.D2Q9, 32, capi.FP32), (capi.D3Q27, 16, capi.FP16)])
def test_cubic_interpolation_adaptive(vs, N, prec):
    """Cubic coarse-to-fine interpolation with BLOCK averaging in the advance, through solution-
    based rebuilds (ADDED children are interpolated cubically too)."""
    p = ldc_params(vs=vs, N=N, Re=40.0, levels=3, P_REFINE=15, prec=prec, interp_order=capi.INTERP_CUBIC)
    g, o = pair(p, 8192)
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.M_FreezeRefinedCells(0))
    for k in range(3):
        rg = g.advance_loop(15 * k, 15)
        ro = o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[1] > 0


@pytest.mark.parametrize("vs,N", [(capi.D2Q9, 32), (capi.D3Q19, 16), (capi.D3Q27, 8)])
@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_mrt_uniform_ldc(vs, N, prec):
    p = ldc_params(vs=vs, N=N, Re=50.0, levels=1, prec=prec, collision=capi.MRT)
    g, o = pair(p, 512)
    both((g, o), lambda m: m.S_SetIC(0))
    for _ in range(60):
        both((g, o), lambda m: m.S_Advance(0))
    assert_same(g, o)


@pytest.mark.parametrize("vs,prec", [(capi.D3Q19, capi.FP32), (capi.D3Q27, capi.FP16)])
def test_mrt_adaptive(vs, prec):
    p = ldc_params(vs=vs, N=16, Re=40.0, levels=3, P_REFINE=15, prec=prec, collision=capi.MRT)
    g, o = pair(p, 8192)
    both((g, o), lambda m: m.S_SetIC(0))
    for k in range(3):
        rg = g.advance_loop(15 * k, 15)
        ro = o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[1] > 0


@pytest.mark.parametrize("crit", [1, 2])
def test_solution_adaptive_other_criteria(crit):
    """The Q-like invariant and the |grad u| criterion drive the same rebuild bit for bit."""
    p = ldc_params(vs=capi.D3Q19, N=16, Re=40.0, levels=3, P_REFINE=15, S_CRITERION=crit,
                   N_REFINE_START=-4.0 if crit == 1 else -2.0)
    g, o = pair(p, 8192)
    both((g, o), lambda m: m.S_SetIC(0))
    for k in range(3):
        rg = g.advance_loop(15 * k, 15)
        ro = o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[1] > 0


def test_d2q9_solution_adaptive():
    p = ldc_params(vs=capi.D2Q9, N=32, Re=100.0, levels=3, P_REFINE=25)
    g, o = pair(p, 2048)
    both((g, o), lambda m: m.S_SetIC(0))
    for k in range(4):
        g.advance_loop(25 * k, 25)
        o.advance_loop(25 * k, 25)
        assert_same(g, o)


def test_coarsening_path():
    p = ldc_params(vs=capi.D3Q19, N=16, Re=30.0, levels=3)
    g, o = pair(p, 4096)
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.mark_uniform(0))
    both((g, o), lambda m: m.M_RefineAndCoarsenBlocks())
    assert_same(g, o)
    for _ in range(5):
        both((g, o), lambda m: m.S_Advance(0))
    # Coarsen a few refined root blocks (their children are leaves).
    ids = [0, 1, 7, 20, 63]
    both((g, o), lambda m: m.mark_blocks(ids, capi.REF_MARK_COARSEN))
    both((g, o), lambda m: m.mark_blocks([30], capi.REF_MARK_COARSEN))
    rg = g.M_RefineAndCoarsenBlocks()
    ro = o.M_RefineAndCoarsenBlocks()
    assert rg == ro, (rg, ro)
    assert_same(g, o)
    for _ in range(5):
        both((g, o), lambda m: m.S_Advance(0))
    assert_same(g, o)


def test_stream_guard_exceptions():
    """Negative post-collision populations disable the swap of their link in the reference
    (solver_lbm_kernels.cu:624); the push kernel must reproduce that through its exception list."""
    p = ldc_params(vs=capi.D3Q19, N=8, Re=10.0, levels=1)
    g, o = pair(p, 64)
    both((g, o), lambda m: m.S_SetIC(0))
    f = o.populations()
    rng = np.random.default_rng(3)
    f = f * (1.0 + 1.5 * rng.standard_normal(f.shape).astype(f.dtype))
    both((g, o), lambda m: m.set_populations(f))
    for _ in range(3):
        both((g, o), lambda m: m.S_Advance(0))
    g.sync()
    assert_same(g, o)


def test_exception_list_overflow_is_reported_and_stops_the_context():
    """With the exception list bounded to 4 entries (host option), the perturbed state of the guard
    test overflows it: the list keeps its bound (kernels only skip the extra entries), the next sync
    reports AGAL_ERR_OVERFLOW, and the context then refuses further steps."""
    p = ldc_params(vs=capi.D3Q19, N=8, Re=10.0, levels=1)
    g = LbmMesh(p, 64)
    g.S_SetIC(0)
    f = g.populations()
    rng = np.random.default_rng(3)
    g.set_populations(f * (1.0 + 1.5 * rng.standard_normal(f.shape).astype(f.dtype)))
    g.set_option(capi.OPT_EXC_CAP, 4)
    g.S_Advance(0)
    with pytest.raises(capi.AgalError) as e:
        g.sync()
    assert e.value.status == capi.AGAL_ERR_OVERFLOW
    with pytest.raises(capi.AgalError) as e:
        g.S_Advance(0)
    assert e.value.status == capi.AGAL_ERR_STATE
    with pytest.raises(capi.AgalError):
        g.set_option(capi.OPT_EXC_CAP, 1 << 40)  # above the allocated capacity
    g.close()


@pytest.mark.parametrize("vs,prec", [(capi.D3Q19, capi.FP32), (capi.D3Q19, capi.FP16), (capi.D3Q19, capi.BF16),
                                     (capi.D3Q27, capi.FP32)])
def test_stream_guard_exceptions_across_levels(vs, prec):
    """Guard exceptions on several levels at once: the coarse levels' fused sub-steps share the
    finest level's launches and the fixup serves each level's own list, while the first of each
    pair of finest sub-steps leaves its list for the second to apply as it loads.  A developed
    three-level mesh gets perturbed populations (many non-positive post-collision values on every
    level), then coarse steps as recorded graphs and as direct launches, both against the
    restatement."""
    p = ldc_params(vs=vs, N=16, Re=40.0, levels=3, P_REFINE=15, prec=prec)
    g, o = pair(p, 8192)
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.M_FreezeRefinedCells(0))
    both((g, o), lambda m: m.advance_loop(0, 30))
    assert g.sizes.n_ids[2] > 0
    f = o.populations()
    rng = np.random.default_rng(5)
    # A milder perturbation for FP16 pools: the perturbed flow grows without bound, past what a
    # binary16 deviation holds (65504) within these steps at larger amplitudes, and the two sides may
    # differ in which non-finite value they carry (0 * inf terms).  BF16 has FP32's range.
    amp = 0.35 if prec == capi.FP16 else 1.5
    f = f * (1.0 + amp * rng.standard_normal(f.shape).astype(f.dtype))
    both((g, o), lambda m: m.set_populations(f))
    for k in range(4):
        if k == 2:
            g.set_option(capi.OPT_GRAPHS, 0)
        both((g, o), lambda m: m.S_Advance(0))
    g.sync()
    assert np.isfinite(o.populations()).all()
    assert_same(g, o)


def test_graph_replays_match_direct_launches():
    """Each coarse step is replayed from a recorded HIP graph; interleaved with rebuilds, direct
    calls and host synchronisations, every replay must equal the directly launched run."""
    p = ldc_params(vs=capi.D3Q19, N=16, Re=40.0, levels=3, P_REFINE=7)
    a, b = LbmMesh(p, 8192), LbmMesh(p, 8192)
    b.set_option(capi.OPT_GRAPHS, 0)
    for m in (a, b):
        m.S_SetIC(0)
    for k in range(30):
        a.S_Advance(0)
        b.S_Advance(0)
        if k % 7 == 6:
            for m in (a, b):
                m.M_Advance_RefineWithSolution()
        if k % 3 == 0:
            a.sync()
            bad = compare_states(a.state(), b.state())
            assert not bad, (k, bad)
        if k == 11:
            for m in (a, b):  # direct launches between replays
                m.collide_stream(1)
                m.collide_stream(1)
    assert not compare_states(a.state(), b.state())


def test_pool_exhausted_cancels_refinement():
    p = ldc_params(vs=capi.D3Q19, N=8, levels=2)
    g, o = pair(p, 8 + 12)  # room for one refined block only
    both((g, o), lambda m: m.S_SetIC(0))
    both((g, o), lambda m: m.mark_blocks([0, 3], capi.REF_MARK_REFINE))
    sg, cg = g.M_RefineAndCoarsenBlocks()
    so, co = o.M_RefineAndCoarsenBlocks()
    assert sg == so == capi.AGAL_ERR_POOL_EXHAUSTED
    assert cg == co
    assert_same(g, o)


@pytest.mark.parametrize("vs,N,prec", [(capi.D3Q27, 8, capi.FP32), (capi.D3Q19, 16, capi.FP32),
                                       (capi.D3Q27, 8, capi.FP16)])
def test_mrt_matrix_cores_agree_to_fp32_rounding(vs, N, prec):
    """The MRT moment transform on matrix cores (bf16 two-term splits, FP32 sums in the hardware's
    order) against the restatement's fixed-order FP32 sums on a uniform cavity.  Integer
    state identical; populations to 5e-6 after 40 steps (1.5e-5 of their scale: rounding
    differences of the two summation orders, grown over the steps; FP16 pools: to half a binary16
    step of the stored deviation, 2e-5)."""
    p = ldc_params(vs=vs, N=N, Re=50.0, levels=1, prec=prec, collision=capi.MRT)
    g, o = pair(p, 512)
    g.set_option(capi.OPT_MRT_MFMA, 1)
    both((g, o), lambda m: m.S_SetIC(0))
    for _ in range(40):
        both((g, o), lambda m: m.S_Advance(0))
    g.sync()
    sg, so = g.state(), o.state()
    assert not compare_states(sg, so, check_f=False)
    tol = 2e-5 if prec == capi.FP16 else 5e-6
    np.testing.assert_allclose(sg["f"], so["f"], rtol=0, atol=tol)
    assert np.abs(sg["f"] - so["f"]).max() > 0 or prec == capi.FP16  # a different summation order


def test_mrt_matrix_cores_on_the_sphere_hierarchy():
    """Static near-wall hierarchy (three levels, solid cells, inflow / outflow hook): the matrix-core
    MRT run stays within FP32 rounding of the restatement over six coarse steps."""
    from parity import body, sphere_params
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.05, seed=1)
    p = sphere_params(vs=capi.D3Q27, N=16, levels=3, prec=capi.FP32, collision=capi.MRT, D=0.16)
    g, o = pair(p, 20000)
    g.set_option(capi.OPT_MRT_MFMA, 1)
    for m in (g, o):
        m.set_geometry(tri)
        m.M_Advance_RefineNearWall()
    for _ in range(6):
        both((g, o), lambda m: m.S_Advance(0))
    g.sync()
    sg, so = g.state(), o.state()
    assert not compare_states(sg, so, check_f=False)
    fluid = np.repeat(sg["ref"] != capi.REF_INACTIVE, 64) & (sg["cell_mask"] != capi.CELL_SOLID)
    np.testing.assert_allclose(sg["f"][:, fluid], so["f"][:, fluid], rtol=0, atol=2e-6)
