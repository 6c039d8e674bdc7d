"""Parity at the benchmark workloads' own sizes.

* D3Q19 lid-driven cavity, 64^3 root, 3 levels by the |grad u| criterion, Re 1000, rebuild every
  100 coarse steps: the HIP library and the CPU restatement run the same driver through two
  solution-based rebuilds; the full state is compared bit for bit and the block pool checked
  against brute force.
* D2Q9 cavity 128^2, Re 100, FP32 and FP64.
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
        m.advance_loop(0, 3000)
    bad = compare_states(g.state(), o.state())
    assert not bad, bad


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
