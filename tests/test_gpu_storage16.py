"""GPU parity and accuracy of the 16-bit population pools (f - w_q in binary16 / bfloat16, FP32
math).

Parity: the HIP library and the CPU restatement store the same rounded values at the same points
(initial condition, after the fused collide + hook, interpolation, averaging, uploads), so the
full state must agree bit for bit, as for FP32.  Accuracy: on one mesh sequence (the 16-bit run
takes the FP64 run's refinement marks), the velocity over the leaf cells stays within 1e-3
(relative L2) of FP64 for binary16."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from agal_amd.twins import advance_twins, velocity_error
from oracle_lib import OracleMesh
from parity import compare_states, ldc_params

pytestmark = pytest.mark.gpu

P16 = [capi.FP16, capi.BF16]


def assert_same(g, o):
    bad = compare_states(g.state(), o.state())
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("prec", P16)
@pytest.mark.parametrize("vs,N", [(capi.D2Q9, 32), (capi.D3Q19, 16), (capi.D3Q27, 8)])
def test_uniform_ldc_16bit(vs, N, prec):
    p = ldc_params(vs=vs, N=N, Re=50.0, levels=1, prec=prec)
    g, o = LbmMesh(p, 512), OracleMesh(p, 512)
    for m in (g, o):
        m.S_SetIC(0)
        for _ in range(60):
            m.S_Advance(0)
    assert_same(g, o)


@pytest.mark.parametrize("prec", P16)
@pytest.mark.parametrize("vs", [capi.D2Q9, capi.D3Q19])
def test_solution_adaptive_16bit(vs, prec):
    N = 32 if vs == capi.D2Q9 else 16
    p = ldc_params(vs=vs, N=N, Re=60.0, levels=3, P_REFINE=15, prec=prec)
    g, o = LbmMesh(p, 8192), OracleMesh(p, 8192)
    for m in (g, o):
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
    for k in range(4):
        rg, ro = g.advance_loop(15 * k, 15), o.advance_loop(15 * k, 15)
        assert rg[0] == ro[0]
        assert_same(g, o)
    assert g.sizes.n_ids[1] > 0


def test_stream_guard_16bit():
    p = ldc_params(vs=capi.D3Q19, N=8, Re=10.0, levels=1, prec=capi.FP16)
    g, o = LbmMesh(p, 64), OracleMesh(p, 64)
    for m in (g, o):
        m.S_SetIC(0)
    f = o.populations()
    rng = np.random.default_rng(5)
    f = (f * (1.0 + 1.5 * rng.standard_normal(f.shape))).astype(np.float32)
    for m in (g, o):
        m.set_populations(f)
        for _ in range(3):
            m.S_Advance(0)
    g.sync()
    assert_same(g, o)


def test_fp16_velocity_error_vs_fp64():
    """D3Q19 cavity, 32^3 root, 3 levels, Re 400, rebuild every 50 coarse steps, 400 steps."""
    kw = dict(vs=capi.D3Q19, N=32, Re=400.0, levels=3, P_REFINE=50)
    lead = LbmMesh(ldc_params(prec=capi.FP64, **kw), 60000)
    f16 = LbmMesh(ldc_params(prec=capi.FP16, **kw), 60000)
    bf16 = LbmMesh(ldc_params(prec=capi.BF16, **kw), 60000)
    f32 = LbmMesh(ldc_params(prec=capi.FP32, **kw), 60000)
    for m in (lead, f16, bf16, f32):
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
    advance_twins(lead, [f16, bf16, f32], 0, 400)
    assert lead.sizes.n_ids[2] > 0
    e16 = velocity_error(lead, f16)
    e32 = velocity_error(lead, f32)
    ebf = velocity_error(lead, bf16)
    print(f"velocity error vs FP64: FP32 {e32['rel_l2_velocity']:.3e}, FP16 {e16['rel_l2_velocity']:.3e}, "
          f"BF16 {ebf['rel_l2_velocity']:.3e} over {e16['leaf_cells']} leaf cells")
    assert e32["rel_l2_velocity"] < e16["rel_l2_velocity"] < 1e-3
    assert ebf["rel_l2_velocity"] < 5e-2
