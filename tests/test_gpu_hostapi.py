"""The header-library host API on the HIP library: the same driver run (Parser -> Mesh ->
Solver_LBM -> M_AdvanceLoop with rebuilds, probes, .vthb output, restart) on the GPU and on the CPU
restatement gives bit-identical populations; a run resumed from the restart file equals the
continuous run on the GPU; the pool survives M_RetrieveFromGPU -> M_LoadToGPU unchanged."""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
import hostapi_util as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return H.build(str(tmp_path_factory.mktemp("hostapi_gpu")))


@pytest.mark.parametrize("prec,dt", [(capi.FP32, np.float32), (capi.FP16, np.float32), (capi.FP64, np.float64)])
def test_driver_gpu_equals_restatement(exe, tmp_path, prec, dt):
    g, o = tmp_path / "g", tmp_path / "o"
    g.mkdir()
    o.mkdir()
    print(f"host API driver, precision {prec}", flush=True)
    rg = H.run(exe, "run", "gpu", H.ORACLE, g, "A", 60, 0, prec, timeout=300)
    ro = H.run(exe, "run", "oracle", H.ORACLE, o, "A", 60, 0, prec, timeout=300)
    assert rg["roundtrip"] and rg["n_ids"] == ro["n_ids"] and rg["n_gaps"] == ro["n_gaps"]
    assert rg["n_ids"][1] > 0
    fg = np.fromfile(g / "A_f.bin", dtype=dt)
    fo = np.fromfile(o / "A_f.bin", dtype=dt)
    assert np.array_equal(fg, fo)
    assert (g / "restart.bin").read_bytes() == (o / "restart.bin").read_bytes()


def test_resume_from_restart_on_gpu(exe, tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    H.run(exe, "run", "gpu", H.ORACLE, a, "A", 60, 0, capi.FP32, timeout=300)
    H.run(exe, "run", "gpu", H.ORACLE, b, "B", 30, 0, capi.FP32, timeout=300)
    rc = H.run(exe, "run", "gpu", H.ORACLE, b, "C", 30, 1, capi.FP32, timeout=300)
    assert rc["last_iter"] == 59 and rc["roundtrip"]
    assert np.array_equal(np.fromfile(a / "A_f.bin", dtype=np.float32), np.fromfile(b / "C_f.bin", dtype=np.float32))


def test_output_properties_on_gpu_fields(exe, tmp_path):
    r = H.run(exe, "props", "gpu", H.ORACLE, tmp_path, timeout=300)
    assert r["err_vorticity"] < 1e-12 and r["err_q"] < 1e-12
