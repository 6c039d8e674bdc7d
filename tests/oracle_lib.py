"""Binding of the CPU restatement (oracle/liboracle.so) -- the parity checker.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

from agal_amd import capi
from agal_amd.mesh import LbmMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_NAMES = [k for k in capi.PROTOTYPES
          if not k.startswith("timing") and not k.startswith("predictor") and
          k not in ("set_option", "upload_cell_masks", "upload_block_ints", "criterion_labels", "advance_loop_learned",
                        "download_populations_noalloc")]
_lib = None


def oracle_library():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise FileNotFoundError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(ORACLE_PATH)
        for extra, (res, args) in {
            "set_threads": (C.c_int, [C.c_int]),
            "set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_longlong]),
            "plant": (C.c_int, [C.c_void_p, C.c_int]),
            "collide": (C.c_int, [C.c_void_p, C.c_int]),
            "impose_bc": (C.c_int, [C.c_void_p, C.c_int]),
            "stream": (C.c_int, [C.c_void_p, C.c_int]),
            "ray_count": (C.c_longlong, [C.POINTER(C.c_float), C.c_longlong, C.c_double, C.c_double, C.c_double]),
            "features": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_longlong]),
            "criterion_labels": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_longlong]),
            "mlp_forward": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
            "mlp_train_step": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                         C.c_double, C.c_double, C.POINTER(C.c_float), C.POINTER(C.c_double)]),
            "plog2": (C.c_float, [C.c_float]),
            "upload_cell_masks": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_longlong]),
            "get_level_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
            "velset": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "child_neighbor_rule": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "cell_mask_rule": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int]),
            "in_half_set": (C.c_int, [C.c_int, C.c_int]),
            "upload_macros": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong, C.c_longlong]),
            "to16": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_longlong]),
            "from16": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_longlong]),
        }.items():
            f = getattr(_lib, "oracle_" + extra)
            f.restype, f.argtypes = res, args
    return _lib


def set_threads(n: int) -> int:
    return oracle_library().oracle_set_threads(int(n))


class OracleMesh(LbmMesh):
    def __init__(self, params, capacity):
        lib = oracle_library()
        fns = capi.bind(lib, "oracle_", _NAMES)
        super().__init__(params, capacity, fns=fns, lib=lib, prefix="oracle_")

    def set_option(self, option, value):
        """Only options that change results exist here (AGAL_OPT_LINK_MASKS)."""
        self._call_raw("set_option", int(option), int(value))

    def plant(self, code: int):
        """Plant a known mistake in the restatement (see Oracle::plant); 0 removes it."""
        self._call_raw("plant", int(code))

    def collide(self, L):
        self._call_raw("collide", L)

    def impose_bc(self, L):
        self._call_raw("impose_bc", L)

    def stream(self, L):
        self._call_raw("stream", L)

    def predictor_features(self, n_blocks=None):
        import numpy as np
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty((nb, 12), dtype=np.float32)
        oracle_library().oracle_features(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), nb)
        return out

    def criterion_labels(self, n_blocks=None):
        import numpy as np
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty(nb, dtype=np.int32)
        oracle_library().oracle_criterion_labels(self.h, out.ctypes.data_as(C.POINTER(C.c_int)), nb)
        return out

    def level_params(self):
        n = self.sizes.max_levels
        dxf = (C.c_double * n)()
        tau = (C.c_double * n)()
        oracle_library().oracle_get_level_params(self.h, dxf, tau, n)
        return list(dxf), list(tau)

    def set_cell_masks(self, m):
        import numpy as np
        m = np.ascontiguousarray(m, dtype=np.int32)
        nb = m.size // self.sizes.cells_per_block
        oracle_library().oracle_upload_cell_masks(self.h, m.ctypes.data_as(C.POINTER(C.c_int)), nb)

    def set_macros(self, m):
        import numpy as np
        m = np.ascontiguousarray(m, dtype=self.dtype)
        nc = m.shape[1]
        self._call_raw("upload_macros", m.ctypes.data_as(C.c_void_p), nc // self.sizes.cells_per_block, nc)

    def import_state(self, st: dict):
        """Adopt a block pool and populations downloaded from another context (LbmMesh.state())."""
        import numpy as np
        lib = oracle_library()
        nb = int(st["id_max"])
        P = C.POINTER(C.c_int)

        def ip(a):
            return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(P)

        keep = []

        def arr(a, dt=np.int32):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float) if dt == np.float32 else P)

        s = self.sizes
        id_max = [0] * (s.max_levels + 1)
        id_max[s.max_levels] = nb
        for L in range(s.max_levels):
            ids = st["id_sets"][L]
            id_max[L] = int(ids.max()) + 1 if ids.size else 0
        id_max[0] = s.n_coarsecblocks
        f = lib.oracle_import_pool_arrays
        f.restype = C.c_int
        st_ = f(self.h, C.c_longlong(nb), arr(st["nbr"]), arr(st["nbr_child"]), arr(st["ref"]), arr(st["level"]),
                arr(st["onb"]), arr(st["blk_mask"]), arr(st["coords"], np.float32), arr(st["cell_mask"]),
                arr(np.array(st["n_ids"])), arr(np.array(id_max)), arr(np.concatenate(st["id_sets"])),
                C.c_int(len(st["gap_set"])), arr(st["gap_set"]))
        if st_ != capi.AGAL_OK:
            raise capi.AgalError(st_, "oracle_import_pool_arrays")
        self.refresh_sizes()
        self.set_populations(st["f"])

    def _call_raw(self, name, *args):
        st = getattr(oracle_library(), "oracle_" + name)(self.h, *args)
        if st != capi.AGAL_OK:
            raise capi.AgalError(st, "oracle_" + name)
        return st


def velset(Q: int):
    """The restated velocity set as numpy arrays: c (3 x Qmax), w, pb, conn_map."""
    import numpy as np
    c = (C.c_int * 81)()
    w = (C.c_double * 27)()
    pb = (C.c_int * 27)()
    cm = (C.c_int * 27)()
    qmax = oracle_library().oracle_velset(Q, c, w, pb, cm)
    cc = np.array(list(c), dtype=np.int64).reshape(3, 27)[:, :qmax]
    return cc, np.array(list(w))[:qmax], np.array(list(pb))[:qmax], np.array(list(cm))[:qmax]


def child_neighbor_rule(Q: int, c: int, p: int):
    a, b = C.c_int(), C.c_int()
    oracle_library().oracle_child_neighbor_rule(Q, c, p, C.byref(a), C.byref(b))
    return a.value, b.value


def cell_mask_rule(Q: int, skip, t: int) -> int:
    arr = (C.c_int * 27)(*([int(x) for x in skip] + [0] * (27 - len(skip))))
    return oracle_library().oracle_cell_mask_rule(Q, arr, t)


def in_half_set(dim: int, p: int) -> bool:
    return bool(oracle_library().oracle_in_half_set(dim, p))


def to16(mode: int, x):
    """The restatement's float32 -> 16-bit conversion (1 binary16, 2 bfloat16) as raw bits."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    st = oracle_library().oracle_to16(mode, x.ctypes.data, out.ctypes.data, x.size)
    assert st == capi.AGAL_OK
    return out


def from16(mode: int, bits):
    import numpy as np
    bits = np.ascontiguousarray(bits, dtype=np.uint16)
    out = np.empty(bits.shape, dtype=np.float32)
    st = oracle_library().oracle_from16(mode, bits.ctypes.data, out.ctypes.data, bits.size)
    assert st == capi.AGAL_OK
    return out


class MlpOracle:
    """The restated network: FP32 master parameters, Adam state, one-step training."""

    def __init__(self, w):
        import numpy as np
        self.w = np.array(w, dtype=np.float32)
        self.m = np.zeros_like(self.w)
        self.v = np.zeros_like(self.w)
        self.g = np.zeros_like(self.w)
        self.t = C.c_int(0)

    def forward(self, x):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
        z = np.empty(x.shape[0], dtype=np.float32)
        f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        oracle_library().oracle_mlp_forward(f(self.w), f(x), x.shape[0], f(z))
        return z

    def train_step(self, x, y, lr=1e-3, pos_weight=4.0):
        import numpy as np
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
        y = np.ascontiguousarray(y, dtype=np.float32)
        f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        loss = C.c_double()
        oracle_library().oracle_mlp_train_step(f(self.w), f(self.m), f(self.v), C.byref(self.t), f(x), f(y),
                                               x.shape[0], lr, pos_weight, f(self.g), C.byref(loss))
        return loss.value


def plog2(v: float) -> float:
    return oracle_library().oracle_plog2(v)
