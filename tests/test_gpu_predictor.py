"""GPU parity and behaviour of the learned refinement predictor.

* Features and criterion labels of a developed adaptive cavity: labels identical to the CPU
  restatement; features to 1e-6 (square roots of the statistics are the only place the two may
  round differently).
* Forward pass: logits bit-identical to the restatement (same FP16 weights, same summation order).
* Training: gradients of successive Adam steps against the restatement's, and the loss going down.
* The learned advance loop on the cavity: the predictor refines ahead of the plain criterion, which
  misses fewer of the blocks it flags than the plain method does on its own.
* The study's two controls without the network (criterion marks at the inference cadence, and
  criterion marks dilated by one block): valid meshes and fewer missed blocks than the plain method.
"""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh
from amr_checks import check_state_fast
from oracle_lib import MlpOracle, OracleMesh
from parity import body, compare_states, ldc_params, sphere_params

pytestmark = pytest.mark.gpu


def developed_cavity(N=16, steps=60):
    p = ldc_params(vs=capi.D3Q19, N=N, Re=40.0, levels=3, P_REFINE=15, S_CRITERION=2, N_REFINE_START=-2.0)
    g, o = LbmMesh(p, 8192), OracleMesh(p, 8192)
    for m in (g, o):
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        m.advance_loop(0, steps)
        for L in range(m.sizes.max_levels):
            m.S_RefreshVariables(L)
    g.sync()
    assert not compare_states(g.state(), o.state())
    return g, o


def test_features_and_labels_match_the_restatement():
    g, o = developed_cavity()
    g.predictor_init(7, 1024)
    fg, fo = g.predictor_features(), o.predictor_features()
    assert np.abs(fg - fo).max() <= 1e-6
    lg, lo = g.criterion_labels(), o.criterion_labels()
    np.testing.assert_array_equal(lg, lo)
    assert (lg == 1).sum() > 0 and (lg == 0).sum() > 0
    # Labels leave the refinement marks untouched.
    assert np.array_equal(g.block_ints(capi.BLK_REF), o.block_ints(capi.BLK_REF))


def test_features_on_the_sphere_hierarchy_use_the_geometry_distance():
    """Features of the sphere flow's wall hierarchy (a body set): the signed distance feature (5) is
    the distance to the nearest triangle centroid, negative inside the body -- identical inputs,
    features to 1e-6 of the restatement's, both signs present."""
    tri = body(center=(0.3, 0.5, 0.5), radius=0.08, subdivisions=3, noise=0.05, seed=1)
    p = sphere_params(vs=capi.D3Q19, N=16, levels=3, collision=capi.BGK, D=0.16)
    g, o = LbmMesh(p, 20000), OracleMesh(p, 20000)
    for m in (g, o):
        m.set_geometry(tri)
        m.M_Advance_RefineNearWall()
        for _ in range(3):
            m.S_Advance(0)
        for L in range(m.sizes.max_levels):
            m.S_RefreshVariables(L)
    g.sync()
    assert not compare_states(g.state(), o.state())
    g.predictor_init(7, 1024)
    fg, fo = g.predictor_features(), o.predictor_features()
    assert np.abs(fg - fo).max() <= 1e-6
    act = g.block_ints(capi.BLK_REF) != capi.REF_INACTIVE
    d = fg[act, 5]
    assert (d < 0).any() and (d > 0).any()


def _rand_params(rng):
    w = np.zeros(capi.MLP_NPARAM, dtype=np.float32)
    w[:768] = rng.uniform(-1, 1, 768) * np.sqrt(6 / 12)
    w[832:4928] = rng.uniform(-1, 1, 4096) * np.sqrt(6 / 64)
    w[4992:5056] = rng.uniform(-1, 1, 64) * np.sqrt(6 / 64)
    w[768:832] = rng.uniform(-0.05, 0.05, 64)
    return w


def test_forward_logits_bit_identical():
    rng = np.random.default_rng(11)
    p = ldc_params(vs=capi.D3Q19, N=8)
    g = LbmMesh(p, 64)
    g.predictor_init(1, 4096)
    w = _rand_params(rng)
    g.predictor_set_params(w)
    assert np.array_equal(g.predictor_params(1), w.astype(np.float16).astype(np.float32))
    x = rng.standard_normal((3000, 12)).astype(np.float32)
    z, prob = g.predictor_forward(x)
    np.testing.assert_array_equal(z, MlpOracle(w).forward(x))
    np.testing.assert_allclose(prob, 1 / (1 + np.exp(-z.astype(np.float64))), atol=1e-6)


def test_training_steps_match_the_restatement():
    rng = np.random.default_rng(12)
    p = ldc_params(vs=capi.D3Q19, N=8)
    g = LbmMesh(p, 64)
    g.predictor_init(1, 2048)
    w = _rand_params(rng)
    g.predictor_set_params(w)
    o = MlpOracle(w)
    losses = []
    for s in range(40):
        x = rng.standard_normal((2048, 12)).astype(np.float32)
        y = ((x[:, 2] - 0.5 * x[:, 7]) > 0.3).astype(np.float32)
        lg = g.predictor_train_step(x, y, lr=2e-3, pos_weight=2.0)
        lo = o.train_step(x, y, lr=2e-3, pos_weight=2.0)
        losses.append(lg)
        if s < 3:
            gg = g.predictor_params(2)
            scale = np.abs(o.g).max()
            np.testing.assert_allclose(gg, o.g, rtol=0, atol=1e-4 * scale)
            assert abs(lg - lo) <= 1e-5 * max(1.0, abs(lo))
    assert losses[-1] < 0.5 * losses[0]


def test_learned_loop_refines_ahead_of_the_criterion():
    """The same cavity run twice with identical training: once with the predictions applied, once
    with the plain criterion alone.  After training, the predictor leaves fewer flagged blocks
    unrefined at the rebuilds."""
    p = ldc_params(vs=capi.D3Q19, N=32, Re=200.0, levels=3, P_REFINE=20, S_CRITERION=2, N_REFINE_START=-2.0)
    res = {}
    for use in (0, 1):
        m = LbmMesh(p, 60000)
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        opts = capi.learn_opts(train_steps=400, infer_every=5, batch=2048, adam_steps=16, lr=1e-3, threshold=0.5,
                               pos_weight=4.0, seed=3, use_predictions=use)
        res[use] = m.advance_loop_learned(0, 800, opts)
        m.close()
    plain, learned = res[0], res[1]
    print("plain", plain)
    print("learned", learned)
    assert plain["flagged"] > 0 and learned["flagged"] > 0
    assert plain["adam_steps_done"] == learned["adam_steps_done"] > 0
    assert learned["predicted_refinements"] > 0
    rate = lambda s: s["missed"] / s["flagged"]
    assert rate(learned) < rate(plain)


def test_learned_loop_controls_without_the_network():
    """The two controls of the learned study on the same cavity: the plain criterion's refinement
    marks applied at the inference cadence (mode 2), and its marks grown by one block at the
    rebuilds (mode 3).  Both keep every invariant of the block pool and miss fewer flagged blocks
    than the plain method; the dilated run holds more refined blocks."""
    p = ldc_params(vs=capi.D3Q19, N=32, Re=200.0, levels=3, P_REFINE=20, S_CRITERION=2, N_REFINE_START=-2.0)
    res, sizes = {}, {}
    for use in (0, 2, 3):
        m = LbmMesh(p, 60000)
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        opts = capi.learn_opts(train_steps=0, infer_every=5, batch=2048, adam_steps=0, lr=1e-3, threshold=0.5,
                               pos_weight=4.0, seed=3, use_predictions=use)
        res[use] = m.advance_loop_learned(0, 400, opts)
        st = m.state(with_f=False, with_cells=False)
        bad = check_state_fast(st, (8, 8, 8), n_max=60000)
        assert not bad, f"mode {use}: " + "\n".join(bad[:10])
        sizes[use] = [int(m.sizes.n_ids[L]) for L in range(3)]
        m.close()
    print(res, sizes)
    rate = lambda s: s["missed"] / s["flagged"]
    assert res[0]["flagged"] > 0 and res[2]["flagged"] > 0 and res[3]["flagged"] > 0
    assert res[2]["predicted_refinements"] > 0 and res[2]["early_rebuilds"] > 0
    assert res[0]["predicted_refinements"] == res[3]["predicted_refinements"] == 0
    assert rate(res[2]) < rate(res[0])
    assert rate(res[3]) <= rate(res[0])
    assert sum(sizes[3][1:]) > sum(sizes[0][1:])
