"""Pins the restatement's 16-bit population pools (FP16 / BF16 storage of f - w_q, FP32 math).

The conversions are checked against numpy's binary16 and torch's bfloat16 (both round to nearest
even); the pool rule "a stored value reads back as w + round16(f - w)" is checked to be a
projection (storing a read-back value changes nothing, for every 16-bit pattern and every lattice
weight); the stepping is checked against the FP32 restatement it must stay close to."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
import oracle_lib as O
from oracle_lib import OracleMesh
from parity import ldc_params

torch = pytest.importorskip("torch")


def sample_floats(rng, n=200000):
    mags = 2.0 ** rng.uniform(-30, 17, n)
    x = (mags * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    # exact ties of binary16 and bfloat16 and their neighbours
    h = rng.integers(0, 0x7BFF, 20000).astype(np.uint16)
    lo = h.view(np.float16).astype(np.float32)
    hi = (h + 1).view(np.float16).astype(np.float32)
    mid = ((lo.astype(np.float64) + hi.astype(np.float64)) / 2).astype(np.float32)
    b = rng.integers(0, 0x7F7F, 20000).astype(np.uint32) << 16
    bmid = (b | 0x8000).view(np.float32)
    special = np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e30, -1e30, 2.0 ** -24, 2.0 ** -25,
                        1.5 * 2.0 ** -25, 2.0 ** -14, 2.0 ** -14 * (1 - 2.0 ** -11)], dtype=np.float32)
    return np.concatenate([x, mid, np.nextafter(mid, np.float32(np.inf)), np.nextafter(mid, np.float32(-np.inf)),
                           bmid, special, -special])


def test_binary16_conversion_matches_numpy():
    x = sample_floats(np.random.default_rng(1))
    ours = O.to16(1, x)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref), x[ours != ref][:10]
    allbits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    dec = O.from16(1, allbits)
    ref = allbits.view(np.float16).astype(np.float32)
    fin = np.isfinite(ref)
    assert np.array_equal(dec[fin], ref[fin]) and np.all(np.isnan(dec[~fin]) == np.isnan(ref[~fin]))


def test_bfloat16_conversion_matches_torch():
    x = sample_floats(np.random.default_rng(2))
    ours = O.to16(2, x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref), x[ours != ref][:10]
    allbits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    dec = O.from16(2, allbits)
    ref = torch.from_numpy(allbits.view(np.int16)).view(torch.bfloat16).float().numpy()
    fin = np.isfinite(ref)
    assert np.array_equal(dec[fin], ref[fin])


@pytest.mark.parametrize("mode", [1, 2])
def test_pool_round_trip_is_a_projection(mode):
    """w + round16(f - w) evaluated on a value that already reads back from the pool returns it:
    re-storing decoded populations (inert cells, relabelled slots) never changes what they read
    back as.  (The stored bits may change where that is invisible: -0 becomes +0, and two
    deviations closer than an FP32 ulp of w read back as the same value.)"""
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    d = O.from16(mode, bits)
    keep = np.isfinite(d) & (np.abs(d) < 8)
    bits, d = bits[keep], d[keep]
    for Q in (9, 19, 27):
        _, w, _, _ = O.velset(Q)
        for wq in np.unique(w.astype(np.float32)):
            v = d + wq                      # float32 read-back
            back = O.to16(mode, v - wq)     # stored again
            assert np.array_equal(O.from16(mode, back) + wq, v), (Q, wq)
            changed = back != bits
            dd = O.from16(mode, back[changed]) - d[changed]
            assert np.all(np.abs(dd) <= 2 * np.spacing(wq)), (Q, wq)


@pytest.mark.parametrize("prec,tol", [(capi.FP16, 2e-3), (capi.BF16, 2e-2)])
def test_16bit_pool_tracks_fp32(prec, tol):
    """The LDC flow with a 16-bit pool stays within the storage rounding of the FP32 run."""
    p32 = ldc_params(vs=capi.D3Q19, N=16, Re=100.0, levels=2, P_REFINE=20)
    p16 = ldc_params(vs=capi.D3Q19, N=16, Re=100.0, levels=2, P_REFINE=20, prec=prec)
    a, b = OracleMesh(p32, 2048), OracleMesh(p16, 2048)
    for m in (a, b):
        m.S_SetIC(0)
        m.mark_blocks([21, 22, 25, 26, 37, 38, 41, 42], capi.REF_MARK_REFINE)
        m.M_RefineAndCoarsenBlocks()
        for _ in range(60):
            m.S_Advance(0)
        for L in range(2):
            m.S_RefreshVariables(L)
    sa, sb = a.state(), b.state()
    assert sa["n_ids"] == sb["n_ids"]
    ua, ub = a.macros()[1:4], b.macros()[1:4]
    err = np.sqrt(np.sum((ua - ub) ** 2) / np.sum(ua ** 2))
    assert 0 < err < tol, err
    # every stored population of the 16-bit pool is representable: w + round16(f - w)
    f = b.populations()
    _, w, _, _ = O.velset(19)
    n = int(sb["id_max"]) * 64
    for q in range(19):
        wq = np.float32(w[q])
        dq = f[q, :n] - wq
        assert np.array_equal(O.from16(1 if prec == capi.FP16 else 2, O.to16(1 if prec == capi.FP16 else 2, dq)) + wq,
                              f[q, :n])


def test_guard_sign_without_the_add():
    """The fused step tests a stored 16-bit deviation s against the guard dec(s, w) > 0 as
    float(s) > -w, with no add.  Every binary16 and bfloat16 pattern against every lattice weight
    (D2Q9, D3Q19, D3Q27) in float32 arithmetic: the two tests agree, NaNs and infinities
    included."""
    bits = np.arange(1 << 16, dtype=np.uint32)
    halves = bits.astype(np.uint16).view(np.float16).astype(np.float32)
    bfloats = (bits << 16).view(np.float32)
    weights = np.array([4 / 9, 1 / 9, 1 / 36, 1 / 3, 1 / 18, 8 / 27, 2 / 27, 1 / 54, 1 / 216], dtype=np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        for vals in (halves, bfloats):
            for w in weights:
                with_add = (vals + w) > np.float32(0)
                without = vals > -w
                assert np.array_equal(with_add, without), float(w)
