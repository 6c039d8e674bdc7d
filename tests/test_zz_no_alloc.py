"""The allocation-free population download (agal_download_populations_noalloc) returns exactly what
agal_download_populations returns, on two calls in a row (the second with fewer blocks, so the
shared staging area still holds the first call's larger contents), and rejects a request larger
than its fixed staging area with a clear error."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from agal_amd import capi
from agal_amd.mesh import LbmMesh

pytestmark = pytest.mark.gpu


def _download(m: LbmMesh, fn, n_blocks: int, dtype) -> tuple[int, np.ndarray]:
    cells = n_blocks * 64
    out = np.full(m.params.velocity_set * cells, np.nan, dtype=dtype)
    st = fn(m.h, np_ptr(out), n_blocks, cells)
    return st, out


def np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("prec", [capi.FP32, capi.FP16, capi.FP64])
def test_noalloc_download_equals_download(prec):
    p = capi.default_params(velocity_set=capi.D3Q19, precision=prec, Nx=32, MAX_LEVELS_WALL=1, MAX_LEVELS_INTERIOR=2,
                            v0=0.05 * 32 / 100.0 / 32, bc_preset=capi.BC_LDC3D, bc_velocity=0.05, P_REFINE=10,
                            N_REFINE_START=-3, N_REFINE_INC=1, N_REFINE_MAX=1)
    print(f"no-alloc download: D3Q19 LDC 32^3, 2 levels, precision {prec}", flush=True)
    lib = capi.load_library()
    fns = capi.bind(lib, "agal_", ["download_populations", "download_populations_noalloc"])
    m = LbmMesh(p, 16384)
    try:
        m.S_SetIC(0)
        m.M_FreezeRefinedCells(0)
        m.advance_loop(0, 30)
        m.sync()
        dt = np.float64 if prec == capi.FP64 else np.float32
        sz = m.refresh_sizes()
        n_all = int(sz.id_max[sz.max_levels])
        assert n_all > 512
        for n_blocks in (n_all, n_all // 3):  # second call smaller than the first
            st_a, a = _download(m, fns["download_populations"], n_blocks, dt)
            st_b, b = _download(m, fns["download_populations_noalloc"], n_blocks, dt)
            assert st_a == capi.AGAL_OK and st_b == capi.AGAL_OK
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), n_blocks
        # A request beyond the fixed staging area is refused with a message, before any transfer.
        per_block = p.velocity_set * 64 * (8 if prec == capi.FP64 else 4)
        too_many = capi.AGAL_NOALLOC_SCRATCH_BYTES // per_block + 1
        assert too_many <= 16384
        big = np.zeros(8, dtype=dt)
        st = fns["download_populations_noalloc"](m.h, np_ptr(big), too_many, too_many * 64)
        assert st == capi.AGAL_ERR_ARG
        assert b"staging bytes" in lib.agal_last_error(m.h)
    finally:
        m.close()
