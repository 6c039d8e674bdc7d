"""Parity at the benchmark workloads' own sizes.

* D3Q19 lid-driven cavity, 64^3 root, 3 levels by the |grad u| criterion, Re 1000, rebuild every
  100 coarse steps: the HIP library and the CPU restatement run the same driver through two
  solution-based rebuilds; the full state is compared bit for bit and the block pool checked
  against brute force.
* D2Q9 cavity 128^2, Re 100, FP32 and FP64, its full 20 000 steps.
* The sphere flow at its own size (D3Q27 MRT, root 64^3, four wall levels, the 5 120-triangle noisy
  icosphere, links that agree with the cell masks): hierarchy, masks, links, two coarse steps and
  the force.
* Size-independent property at 256^3 (16.7 M cells): on a periodic grid a state that repeats every
  8 cells evolves as the 8^3 periodic state does, so every cell of the large GPU run must equal the
  matching cell of the small CPU run, bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from amr_checks import check_state
from oracle_lib import OracleMesh, set_threads
from parity import compare_states, ldc_params

pytestmark = pytest.mark.gpu


def ldc64_params(**kw):
    """The benchmark's cavity: |grad u| criterion (2^-2 on the root level, halved per level)."""
    return ldc_params(vs=capi.D3Q19, N=64, Re=1000.0, levels=3, P_REFINE=100, S_CRITERION=2, **kw)


def test_ldc64_three_levels_two_rebuilds():
    set_threads(16)
    p = ldc64_params()
    cap = 120000
    g, o = LbmMesh(p, cap), OracleMesh(p, cap)
    for m in (g, o):
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
    for i0, n in ((0, 101), (101, 101)):
        rg = g.advance_loop(i0, n)
        ro = o.advance_loop(i0, n)
        assert rg[0] == ro[0]
        sg, so = g.state(), o.state()
        bad = compare_states(sg, so)
        assert not bad, (i0 + n, bad)
    assert sg["n_ids"][2] > 0
    bad = check_state(sg, 3, (16, 16, 16), n_max=cap)
    assert not bad, bad[:5]


@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_d2q9_ldc128(prec):
    set_threads(16)
    p = ldc_params(vs=capi.D2Q9, N=128, Re=100.0, levels=1, prec=prec)
    g, o = LbmMesh(p, 1100), OracleMesh(p, 1100)
    for m in (g, o):
        m.S_SetIC(0)
        m.advance_loop(0, 20000)
    bad = compare_states(g.state(), o.state())
    assert not bad, bad


def test_sphere_flow_full_size():
    set_threads(16)
    print("sphere flow 64^3, 4 wall levels, D3Q27 MRT FP32", flush=True)
    tri = capi.icosphere((0.3, 0.5, 0.5), 0.08, 4, 0.05, 20240)
    U, D = 0.04, 0.16
    p = capi.default_params(velocity_set=capi.D3Q27, precision=capi.FP32, collision=capi.MRT, Nx=64,
                            MAX_LEVELS_WALL=4, MAX_LEVELS_INTERIOR=1, v0=U * D / 200.0, bc_preset=capi.BC_FPSC3D,
                            bc_velocity=U, ic_u=U, P_REFINE=1 << 30, G_BIN_MAKE_LINKS=1, G_BIN_DENSITY=16,
                            G_BIN_LEVELS=10, G_BIN_SPEC=7, S_BC_TYPE=0, S_FORCE_TYPE=0, S_FORCE_ORDER=1, s6=1.2)
    g, o = LbmMesh(p, 60000), OracleMesh(p, 60000)
    for m in (g, o):
        m.set_option(capi.OPT_LINK_MASKS, 1)
        m.set_geometry(tri)
    assert g.M_Advance_RefineNearWall() == o.M_Advance_RefineNearWall() == capi.AGAL_OK
    g.sync()
    assert list(g.sizes.n_ids[:4]) == [4096, 3168, 6728, 21216]
    bad = compare_states(g.state(), o.state())
    assert not bad, bad
    lg, lo = g.links(), o.links()
    np.testing.assert_array_equal(lg["X_b"].view(np.uint32), lo["X_b"].view(np.uint32))
    for k in range(2):
        for m in (g, o):
            m.S_Advance(0)
        g.sync()
        bad = compare_states(g.state(), o.state())
        assert not bad, (k, bad)
        fg, fo = g.S_ComputeForces(), o.S_ComputeForces()
        np.testing.assert_allclose(fg, fo, rtol=1e-9, atol=1e-15)
    assert np.abs(fg).max() > 0


@pytest.mark.parametrize("prec", [capi.FP32, capi.FP64])
def test_periodic_tiling_256(prec):
    small, big = 8, 256
    kw = dict(vs=capi.D3Q19, Re=50.0, levels=1, prec=prec, bc_preset=capi.BC_NONE,
              PERIODIC_X=1, PERIODIC_Y=1, PERIODIC_Z=1)
    ps = ldc_params(N=small, **kw)
    pb = ldc_params(N=big, **kw)
    # Same lattice viscosity on both grids: v0 and dx scale by small/big, a power of two, so
    # omega = dx / (3 v0 + dx/2) is bit-identical in both runs.
    pb.v0 = ps.v0 * small / big
    o = OracleMesh(ps, (small // 4) ** 3 + 1)
    g = LbmMesh(pb, (big // 4) ** 3 + 1)
    o.S_SetIC(0)
    g.S_SetIC(0)
    f0 = o.populations()
    rng = np.random.default_rng(9)
    f0 = f0 * (1 + 0.05 * rng.uniform(-1, 1, f0.shape)).astype(f0.dtype)
    o.set_populations(f0)

    def cell_xyz(n, nb_axis):
        b = np.arange(nb_axis ** 3)
        bx, by, bz = b % nb_axis, (b // nb_axis) % nb_axis, b // nb_axis ** 2
        t = np.arange(64)
        I, J, K = t % 4, (t // 4) % 4, t // 16
        x = (4 * bx[:, None] + I[None, :]).ravel()
        y = (4 * by[:, None] + J[None, :]).ravel()
        z = (4 * bz[:, None] + K[None, :]).ravel()
        return x, y, z

    xs, ys, zs = cell_xyz(small, small // 4)
    lin_small = np.empty(small ** 3, dtype=np.int64)
    lin_small[xs + small * (ys + small * zs)] = np.arange(small ** 3)
    xb, yb, zb = cell_xyz(big, big // 4)
    src = lin_small[(xb % small) + small * ((yb % small) + small * (zb % small))]
    g.set_populations(f0[:, src])
    del f0
    steps = 6
    for _ in range(steps):
        o.S_Advance(0)
        g.S_Advance(0)
    fo = o.populations()
    fg = g.populations()
    assert np.array_equal(fg, fo[:, src])
