"""GPU parity of the control-volume force (S_FORCE_TYPE 1): the HIP library against the CPU
restatement on the same inputs -- populations bit for bit (the force's second half on several
levels re-averages and re-interpolates the whole hierarchy every coarse step, so the state carries
the schedule too) and the reported force to the rounding of the FP64 sums' order."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from oracle_lib import OracleMesh
from parity import body, compare_states, ldc_params, sphere_params
from test_oracle_physics import feq_np

pytestmark = pytest.mark.gpu

BOX = dict(S_FORCEVOLUME_Xm=0.3, S_FORCEVOLUME_XM=0.8, S_FORCEVOLUME_Ym=0.2, S_FORCEVOLUME_YM=0.7,
           S_FORCEVOLUME_Zm=0.25, S_FORCEVOLUME_ZM=0.9)


def assert_same(g, o):
    g.sync()
    bad = compare_states(g.state(), o.state())
    assert not bad, "\n".join(bad)


def assert_forces(g, o, atol):
    fg, fo = g.S_ComputeForces(), o.S_ComputeForces()
    assert np.all(np.isfinite(fg))
    np.testing.assert_allclose(fg, fo, rtol=1e-9, atol=atol)
    return fg


@pytest.mark.parametrize("vs", [capi.D2Q9, capi.D3Q19, capi.D3Q27])
@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_cv_periodic_single_level(vs, prec):
    """Random flow on one periodic level: both halves on the root level, the balance closing."""
    print(f"cv periodic single level: Q {vs} precision {prec}", flush=True)
    dim = 2 if vs == capi.D2Q9 else 3
    N = 32 if dim == 2 else 16
    p = capi.default_params(velocity_set=vs, precision=prec, Nx=N, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=1,
                            v0=0.02 / N, bc_preset=capi.BC_NONE, PERIODIC_X=1, PERIODIC_Y=1,
                            PERIODIC_Z=1 if dim == 3 else 0, P_REFINE=1 << 30, S_FORCE_TYPE=1, **BOX)
    g, o = LbmMesh(p, 4096), OracleMesh(p, 4096)
    for m in (g, o):
        m.S_SetIC(0)
    nc = o.populations().shape[1]
    rng = np.random.default_rng(vs)
    rho = 1 + 0.05 * rng.uniform(-1, 1, nc)
    u = np.zeros((3, nc))
    u[:dim] = 0.05 * rng.uniform(-1, 1, (dim, nc))
    f = feq_np(vs, rho, u) * (1 + 0.05 * rng.uniform(-1, 1, (vs, nc)))
    for m in (g, o):
        m.set_populations(f)
    for _ in range(3):
        for m in (g, o):
            m.S_Advance(0)
        assert_same(g, o)
        fg = assert_forces(g, o, atol=1e-12)
        assert np.abs(fg).max() < (1e-13 if prec == capi.FP64 else 1e-5)
    g.close()
    o.close()


@pytest.mark.parametrize("vs,N,prec,graphs", [(capi.D3Q19, 16, capi.FP32, 1), (capi.D3Q19, 16, capi.FP64, 1),
                                              (capi.D2Q9, 32, capi.FP32, 1), (capi.D3Q27, 16, capi.FP16, 1),
                                              (capi.D3Q19, 16, capi.FP32, 0)])
def test_cv_adaptive_three_levels(vs, N, prec, graphs):
    """Lid-driven cavity, three levels rebuilt from the solution every 15 coarse steps, the box
    holding the lid rows: the first half on every level, the full average, the interface
    interpolations and the second half after each coarse step, through recorded step graphs (and
    once with direct launches)."""
    print(f"cv adaptive three levels: Q {vs} N {N} precision {prec} graphs {graphs}", flush=True)
    box = dict(BOX, S_FORCEVOLUME_YM=1.0) if vs == capi.D2Q9 else dict(BOX, S_FORCEVOLUME_ZM=1.0)
    p = ldc_params(vs=vs, N=N, Re=40.0, levels=3, P_REFINE=15, prec=prec, S_FORCE_TYPE=1, **box)
    g, o = LbmMesh(p, 8192), OracleMesh(p, 8192)
    g.set_option(capi.OPT_GRAPHS, graphs)
    for m in (g, o):
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
    for k in range(3):
        rg = g.advance_loop(15 * k, 15)
        ro = o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
        fg = assert_forces(g, o, atol=1e-12)
    assert g.sizes.n_ids[1] > 0
    assert np.abs(fg).max() > 0
    g.close()
    o.close()


def test_cv_sphere_three_levels_with_links():
    """The noisy icosphere, D3Q27 MRT, three wall levels, plain bounce-back on links agreeing with
    the masks, the box around the body: step by step, then through the advance loop."""
    print("cv sphere three levels", flush=True)
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.05, seed=1)
    p = sphere_params(vs=capi.D3Q27, N=16, levels=3, collision=capi.MRT, D=0.16, G_BIN_MAKE_LINKS=1,
                      G_BIN_DENSITY=4, G_BIN_LEVELS=3, G_BIN_SPEC=1, S_BC_TYPE=0, S_FORCE_TYPE=1,
                      S_FORCEVOLUME_Xm=0.125, S_FORCEVOLUME_XM=0.75, S_FORCEVOLUME_Ym=0.25,
                      S_FORCEVOLUME_YM=0.75, S_FORCEVOLUME_Zm=0.25, S_FORCEVOLUME_ZM=0.75)
    g, o = LbmMesh(p, 20000), OracleMesh(p, 20000)
    for m in (g, o):
        m.set_option(capi.OPT_LINK_MASKS, 1)
        m.set_geometry(tri)
        assert m.M_Advance_RefineNearWall() == capi.AGAL_OK
    for m in (g, o):
        m.S_Advance(0)
    assert_same(g, o)
    assert_forces(g, o, atol=1e-12)
    for m in (g, o):
        m.advance_loop(1, 8)
    assert_same(g, o)
    fg = assert_forces(g, o, atol=1e-12)
    assert np.abs(fg[0]) > 0
    g.close()
    o.close()
