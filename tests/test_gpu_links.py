"""GPU parity of the face-cell links against the CPU restatement: the face bins of the link search
(ray-indicator filter, triangle / box overlap, order inside a bin), the ranks of the boundary
blocks and the cut-link fractions bit for bit; then coarse steps with the interpolated bounce-back
(both forms) and the momentum-exchange force of order 1 and 2, populations bit for bit and the
force to a relative 1e-9 (its FP64 total is summed in another order).  Also the plane Couette
flow over an off-lattice slab (one level, periodic sides), where the reference's validation
corrections fire on blocks that are not the ranked ones."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from oracle_lib import OracleMesh
from parity import body, compare_states, sphere_params

pytestmark = pytest.mark.gpu


def assert_same(g, o):
    bad = compare_states(g.state(), o.state())
    assert not bad, "\n".join(bad)


def assert_links_same(g, o):
    lg, lo = g.links(), o.links()
    assert lg["present"] and lo["present"]
    assert lg["n_solidb"] == lo["n_solidb"] and lg["corrections"] == lo["corrections"]
    np.testing.assert_array_equal(lg["onb_solid"], lo["onb_solid"])
    np.testing.assert_array_equal(lg["X_b"].view(np.uint32), lo["X_b"].view(np.uint32))
    bg, bo = g.link_bins(), o.link_bins()
    assert bg["G"] == bo["G"] and bg["kept"] == bo["kept"]
    np.testing.assert_array_equal(bg["start"], bo["start"])
    np.testing.assert_array_equal(bg["faces"], bo["faces"])
    return lg


# noise: the +-5% vertex noise of the workload's body.  On it, at this resolution, a few boundary
# cells have a solid cell behind a cut link; the reference's linear form then reads the solid
# value and the run diverges within a few steps -- identically on both sides, which is what this
# case checks.  The smooth sphere stays finite.
@pytest.mark.parametrize("vs,prec,coll,order,noise", [(capi.D3Q19, capi.FP32, capi.BGK, 2, 0.0),
                                                      (capi.D3Q27, capi.FP32, capi.MRT, 1, 0.0),
                                                      (capi.D3Q19, capi.FP64, capi.BGK, 2, 0.05),
                                                      (capi.D3Q27, capi.FP16, capi.MRT, 2, 0.0)])
def test_sphere_links_bouzidi_and_forces(vs, prec, coll, order, noise):
    print(f"links: sphere (noise {noise}), velocity set {vs}, precision {prec}, collision {coll}, force order {order}",
          flush=True)
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=noise, seed=1)
    p = sphere_params(vs=vs, N=16, levels=3, prec=prec, collision=coll, D=0.16, G_BIN_MAKE_LINKS=1,
                      G_BIN_DENSITY=4, G_BIN_LEVELS=3, G_BIN_SPEC=1, S_BC_TYPE=1, S_FORCE_TYPE=0, S_FORCE_ORDER=order)
    g, o = LbmMesh(p, 20000), OracleMesh(p, 20000)
    for m in (g, o):
        m.set_geometry(tri)
    assert g.M_Advance_RefineNearWall() == o.M_Advance_RefineNearWall() == capi.AGAL_OK
    g.sync()
    assert_same(g, o)
    L = assert_links_same(g, o)
    assert (L["X_b"] > 0).sum() > 1000 and ((L["X_b"] > 0) & (L["X_b"] < 0.5)).any() and (L["X_b"] >= 0.5).any()
    for k in range(3):
        for m in (g, o):
            for _ in range(2):
                m.S_Advance(0)
        g.sync()
        assert_same(g, o)
        fg, fo = g.S_ComputeForces(), o.S_ComputeForces()
        np.testing.assert_allclose(fg, fo, rtol=1e-9, atol=1e-14)
    if noise == 0.0:
        assert np.all(np.isfinite(fg)) and np.abs(fg).max() > 0
    g.close()
    o.close()


def test_links_survive_graph_replays_and_loop():
    """The advance loop (recorded step graphs) gives the same populations and forces."""
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.0, seed=1)
    p = sphere_params(vs=capi.D3Q27, N=16, levels=3, collision=capi.MRT, D=0.16, G_BIN_MAKE_LINKS=1,
                      G_BIN_DENSITY=4, G_BIN_LEVELS=3, G_BIN_SPEC=1, S_BC_TYPE=1, S_FORCE_TYPE=0, S_FORCE_ORDER=2)
    g, o = LbmMesh(p, 20000), OracleMesh(p, 20000)
    for m in (g, o):
        m.set_geometry(tri)
        m.M_Advance_RefineNearWall()
        m.advance_loop(0, 9)
    g.sync()
    assert_same(g, o)
    np.testing.assert_allclose(g.S_ComputeForces(), o.S_ComputeForces(), rtol=1e-9, atol=1e-14)


def box_triangles(x0, x1, y0, y1, z0, z1):
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return np.array([[v[a], v[b], v[c]] for a, b, c in f], np.float32)


@pytest.mark.parametrize("zw", [0.19, 0.245])
def test_slab_couette_links(zw):
    N = 16
    p = capi.default_params(velocity_set=capi.D3Q19, Nx=N, L_fy=0.25, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=1,
                            v0=0.1 / N, bc_preset=capi.BC_LDC3D, bc_velocity=0.01, PERIODIC_X=1, PERIODIC_Y=1,
                            P_REFINE=1 << 30, G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=4, G_BIN_LEVELS=1, G_BIN_SPEC=1,
                            S_BC_TYPE=1, S_FORCE_TYPE=0, S_FORCE_ORDER=2)
    g, o = LbmMesh(p, 4096), OracleMesh(p, 4096)
    for m in (g, o):
        m.set_geometry(box_triangles(-0.1, 1.1, -0.1, 1.1, -0.1, zw))
        m.M_Advance_RefineNearWall()
    g.sync()
    assert_links_same(g, o)
    for m in (g, o):
        m.advance_loop(0, 200)
    g.sync()
    assert_same(g, o)
    np.testing.assert_allclose(g.S_ComputeForces(), o.S_ComputeForces(), rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("U", [0.0, 0.04])
def test_links_agreeing_with_masks(U):
    """AGAL_OPT_LINK_MASKS on the workload's noisy icosphere (32^3 root, four wall levels, D3Q19 FP64):
    BOUNDARY masks, ranks and fractions bit for bit; at rest the momentum exchange is zero to
    rounding; in the flow, two rounds of two coarse steps with populations bit for bit and the
    force to a relative 1e-9."""
    print(f"links agreeing with masks: U {U}", flush=True)
    tri = capi.icosphere((0.3, 0.5, 0.5), 0.08, 4, 0.05, 20240)
    p = capi.default_params(velocity_set=capi.D3Q19, precision=capi.FP64, collision=capi.BGK, Nx=32,
                            MAX_LEVELS_WALL=4, MAX_LEVELS_INTERIOR=1, v0=0.04 * 0.16 / 200.0,
                            bc_preset=capi.BC_FPSC3D, bc_velocity=U, ic_u=U, P_REFINE=1 << 30,
                            G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=16, G_BIN_LEVELS=10, G_BIN_SPEC=7,
                            S_BC_TYPE=0, S_FORCE_TYPE=0, S_FORCE_ORDER=1)
    g, o = LbmMesh(p, 60000), OracleMesh(p, 60000)
    for m in (g, o):
        m.set_option(capi.OPT_LINK_MASKS, 1)
        m.set_geometry(tri)
    assert g.M_Advance_RefineNearWall() == o.M_Advance_RefineNearWall() == capi.AGAL_OK
    g.sync()
    assert_same(g, o)
    assert_links_same(g, o)
    for k in range(2):
        for m in (g, o):
            for _ in range(2):
                m.S_Advance(0)
        g.sync()
        assert_same(g, o)
        fg, fo = g.S_ComputeForces(), o.S_ComputeForces()
        np.testing.assert_allclose(fg, fo, rtol=1e-9, atol=1e-15)
        if U == 0.0:
            assert np.abs(fg).max() < 1e-15
    if U > 0:
        assert np.all(np.isfinite(fg)) and np.abs(fg).max() > 0
    g.close()
    o.close()
