"""GPU parity of the geometry path: voxelization with binned rays, near-wall marks and rebuilds,
boundary cells, the sphere flow's steps (inflow / outflow hook, bounce-back on the body) and the
momentum-exchange force, against the CPU restatement (which tests every triangle per ray).

Integer state and populations must be identical; the force's per-block sums are too, and their
FP64 total is summed in another order on the GPU (strided lanes and a tree) than on the CPU (block
order), so it is compared to a relative 1e-9."""
from __future__ import annotations

import os

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from oracle_lib import OracleMesh
from parity import body, compare_states, sphere_params

pytestmark = pytest.mark.gpu


def pair(p, cap, tri):
    g, o = LbmMesh(p, cap), OracleMesh(p, cap)
    for m in (g, o):
        m.set_geometry(tri)
    return g, o


def assert_same(g, o):
    bad = compare_states(g.state(), o.state())
    assert not bad, "\n".join(bad)


def test_voxelize_root_grid_binned_rays_match_all_triangles():
    """32^3 root cells against a 5,120-triangle body: cell and block masks, marks on level 0."""
    tri = body(center=(0.45, 0.5, 0.52), radius=0.3, subdivisions=4, noise=0.05, seed=9)
    p = sphere_params(vs=capi.D3Q19, N=32, levels=1, collision=capi.BGK)
    g, o = pair(p, 1024, tri)
    for m in (g, o):
        m.M_Geometry_Voxelize_S1(0)
        m.M_Geometry_Voxelize_S2(0)
    g.sync()
    assert_same(g, o)
    st = g.state()
    assert (st["cell_mask"] == capi.CELL_SOLID).sum() > 3500  # 4/3 pi (0.3 * 32)^3 = 3706 cells
    assert (st["cell_mask"] == capi.CELL_BOUNDARY).sum() > 1000


@pytest.mark.parametrize("vs,prec,coll,N", [(capi.D3Q19, capi.FP32, capi.BGK, 32), (capi.D3Q19, capi.FP64, capi.BGK, 16),
                                            (capi.D3Q27, capi.FP32, capi.MRT, 16), (capi.D3Q27, capi.FP16, capi.MRT, 16)])
def test_sphere_flow_near_wall_hierarchy_steps_and_forces(vs, prec, coll, N):
    """The body of the sphere-flow workload (radius 0.08 L at x = 0.3 L, +-5% noise) on a small
    root grid with three wall levels; three rounds of two coarse steps."""
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.05, seed=1)
    # Plain bounce-back on the body; face-cell links for the first-order momentum exchange.
    p = sphere_params(vs=vs, N=N, levels=3, prec=prec, collision=coll, D=0.16, G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=4,
                      G_BIN_LEVELS=3, G_BIN_SPEC=1, S_BC_TYPE=0, S_FORCE_TYPE=0, S_FORCE_ORDER=1)
    g, o = pair(p, 20000, tri)
    sg = g.M_Advance_RefineNearWall()
    so = o.M_Advance_RefineNearWall()
    assert sg == so == capi.AGAL_OK
    assert_same(g, o)
    assert g.sizes.n_ids[2] > 1000
    for k in range(3):
        for m in (g, o):
            for _ in range(2):
                m.S_Advance(0)
        g.sync()
        assert_same(g, o)
        fg, fo = g.S_ComputeForces(), o.S_ComputeForces()
        np.testing.assert_allclose(fg, fo, rtol=1e-9, atol=1e-14)
    assert np.abs(fg).max() > 0


def test_stl_geometry_hierarchy():
    """The reference's monkey input, scaled into the domain: an open surface, so the parity count
    decides what is solid; both sides must decide it identically."""
    tri = capi.load_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "geometry", "monkey.stl"))
    lo, hi = tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)
    tri = ((tri - (lo + hi) / 2) / (hi - lo).max() * 0.5 + 0.5).astype(np.float32)
    p = sphere_params(vs=capi.D3Q19, N=16, levels=2, collision=capi.BGK)
    g, o = pair(p, 4096, tri)
    assert g.M_Advance_RefineNearWall() == o.M_Advance_RefineNearWall()
    assert_same(g, o)
