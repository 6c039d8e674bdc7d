"""Host-side mirror of the reference's Mesh / Solver_LBM setup API over the C ABI.

`LbmMesh` owns one context created through a bound C ABI (the HIP library, `agal_*`).  Method
names follow the reference (inc/mesh.h, inc/solver_lbm.h): M_RefineAndCoarsenBlocks,
S_SetIC, S_Advance, ...; each raises `AgalError` on a non-zero status, except the pool-exhausted
refinement cancel, which the reference also survives (src/mesh/mesh_amr.cu:2197-2206) and which
is reported through the return value.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import AgalError, Params, Sizes, np_ptr


class LbmMesh:
    """One mesh + LBM solver on one device, driven through a bound C ABI."""

    def __init__(self, params: Params, capacity: int, *, fns: dict | None = None, lib=None,
                 prefix: str = "agal_", device: int = 0):
        if fns is None:
            lib = capi.load_library() if lib is None else lib
            fns = capi.bind(lib, prefix, capi.PROTOTYPES.keys())
        self._f = fns
        self._lib = lib
        self.params = params
        self.dtype = np.float64 if params.precision == capi.FP64 else np.float32
        h = C.c_void_p()
        st = fns["create"](C.byref(params), int(capacity), int(device), C.byref(h))
        if st != capi.AGAL_OK:
            raise AgalError(st, prefix + "create")
        self.h = h
        self.prefix = prefix
        self.refresh_sizes()

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self._f["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args, ok=(capi.AGAL_OK,)):
        st = self._f[name](self.h, *args)
        if st not in ok:
            detail = ""
            if self._lib is not None and hasattr(self._lib, "agal_last_error") and self.prefix == "agal_":
                detail = (self._lib.agal_last_error(self.h) or b"").decode()
            raise AgalError(st, self.prefix + name, detail)
        return st

    def refresh_sizes(self) -> Sizes:
        s = Sizes()
        self._call("get_sizes", C.byref(s))
        self.sizes = s
        return s

    def sync(self):
        self._call("sync")

    def set_option(self, option, value):
        self._call("set_option", int(option), int(value))

    # -- solver interface (inc/solver.h:62-71) -------------------------------------------------
    def S_SetIC(self, L):
        self._call("set_ic", int(L))

    def S_Advance(self, L0=None):
        self._call("advance", int(self.params.N_LEVEL_START if L0 is None else L0))

    def collide_stream(self, L):
        self._call("collide_stream", int(L))

    def S_Interpolate(self, L, var):
        self._call("interpolate", int(L), int(var))

    def S_Average(self, L, var):
        self._call("average", int(L), int(var))

    def S_RefreshVariables(self, L):
        self._call("refresh", int(L))

    def S_ComputeRefCriteria(self):
        self._call("ref_criteria")

    # -- mesh interface ------------------------------------------------------------------------
    def M_RefineAndCoarsenBlocks(self):
        counts = (C.c_int * 4)()
        st = self._call("refine_coarsen", counts, ok=(capi.AGAL_OK, capi.AGAL_ERR_POOL_EXHAUSTED))
        self.refresh_sizes()
        return st, list(counts)

    def M_Advance_RefineWithSolution(self):
        counts = (C.c_int * 4)()
        st = self._call("refine_with_solution", counts, ok=(capi.AGAL_OK, capi.AGAL_ERR_POOL_EXHAUSTED))
        self.refresh_sizes()
        return st, list(counts)

    def mark_uniform(self, L):
        self._call("mark_uniform", int(L))

    def mark_blocks(self, ids, mark):
        a = np.ascontiguousarray(np.asarray(ids, dtype=np.int32))
        self._call("mark_blocks", np_ptr(a, C.c_int), int(a.size), int(mark))

    def M_Advance_RefineNearWall(self):
        """Returns the status (AGAL_OK or AGAL_ERR_POOL_EXHAUSTED)."""
        st = self._call("refine_near_wall", ok=(capi.AGAL_OK, capi.AGAL_ERR_POOL_EXHAUSTED))
        self.refresh_sizes()
        return st

    def set_geometry(self, tri: np.ndarray):
        """Closed triangle surface, (n, 3, 3) or (n, 9) float32 in domain coordinates."""
        t = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
        self._call("set_geometry", np_ptr(t, C.c_float), int(t.shape[0]))

    def M_Geometry_Voxelize_S1(self, L):
        self._call("voxelize", int(L))

    def M_Geometry_Voxelize_S2(self, L):
        self._call("check_masks", int(L))

    def S_ComputeForces(self) -> np.ndarray:
        """S_ReportForces: the force (x, y, z) accumulated since the last report -- S_FORCE_TYPE 0: the
        momentum exchange of the last coarse step (needs G_BIN_MAKE_LINKS 1); 1: the control-volume
        balance over the S_FORCEVOLUME_* box in the last coarse step.  The accumulators are reset."""
        out = (C.c_double * 3)()
        self._call("forces", out)
        return np.array(list(out))

    def links(self) -> dict:
        """Face-cell links: onb_solid (block -> rank), X_b (27, 64 * n_solidb) fractions, n_solidb,
        corrections of the validation."""
        nb = self.n_blocks()
        info = (C.c_longlong * 3)()
        onb = np.empty(nb, dtype=np.int32)
        self._call("download_links", np_ptr(onb), nb, None, 0, info)
        ns = int(info[0])
        xb = np.empty(27 * 64 * ns, dtype=np.float32)
        if ns > 0 and info[2]:
            self._call("download_links", None, 0, np_ptr(xb), xb.size, info)
        else:
            xb[:] = -1.0
        return {"onb_solid": onb, "X_b": xb.reshape(27, 64 * ns), "n_solidb": ns, "corrections": int(info[1]),
                "present": bool(info[2])}

    def link_bins(self) -> dict:
        """The face bins of the link search: start (n_bins + 1), faces, bins per unit length, kept faces."""
        info = (C.c_longlong * 4)()
        self._call("download_link_bins", None, 0, None, 0, info)
        nbins, nent = int(info[0]), int(info[1])
        start = np.zeros(max(nbins + 1, 0), dtype=np.int32)
        faces = np.zeros(nent, dtype=np.int32)
        if nbins > 0:
            self._call("download_link_bins", np_ptr(start), start.size, None, 0, info)
        if nent > 0:
            self._call("download_link_bins", None, 0, np_ptr(faces), nent, info)
        return {"start": start, "faces": faces, "G": int(info[2]), "kept": int(info[3])}

    # -- learned refinement predictor (no reference counterpart) -----------------------------------
    def predictor_init(self, seed=1, batch_capacity=4096):
        self._call("predictor_init", int(seed), int(batch_capacity))

    def predictor_params(self, which=0) -> np.ndarray:
        out = np.empty(capi.MLP_NPARAM, dtype=np.float32)
        self._call("predictor_get", int(which), np_ptr(out, C.c_float))
        return out

    def predictor_set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == capi.MLP_NPARAM
        self._call("predictor_set", np_ptr(w, C.c_float))

    def predictor_forward(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, capi.MLP_NFEAT)
        n = x.shape[0]
        z = np.empty(n, dtype=np.float32)
        p = np.empty(n, dtype=np.float32)
        self._call("predictor_forward", np_ptr(x, C.c_float), n, np_ptr(z, C.c_float), np_ptr(p, C.c_float))
        return z, p

    def predictor_train_step(self, x, y, lr=1e-3, pos_weight=4.0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, capi.MLP_NFEAT)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_double()
        self._call("predictor_train_step", np_ptr(x, C.c_float), np_ptr(y, C.c_float), x.shape[0], float(lr),
                   float(pos_weight), C.byref(loss))
        return loss.value

    def predictor_features(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty((nb, capi.MLP_NFEAT), dtype=np.float32)
        self._call("predictor_features", np_ptr(out, C.c_float), nb)
        return out

    def criterion_labels(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty(nb, dtype=np.int32)
        self._call("criterion_labels", np_ptr(out, C.c_int), nb)
        return out

    def advance_loop_learned(self, i0, n, opts=None) -> dict:
        o = capi.learn_opts() if opts is None else opts
        s = capi.LearnStats()
        self._call("advance_loop_learned", int(i0), int(n), C.byref(o), C.byref(s))
        self.refresh_sizes()
        return s.as_dict()

    def M_FreezeRefinedCells(self, var=0):
        self._call("freeze_refined", int(var))

    def advance_loop(self, i0, n):
        ts, tr = C.c_double(), C.c_double()
        st = self._call("advance_loop", int(i0), int(n), C.byref(ts), C.byref(tr),
                        ok=(capi.AGAL_OK, capi.AGAL_ERR_POOL_EXHAUSTED))
        self.refresh_sizes()
        return st, ts.value, tr.value

    # -- data exchange -------------------------------------------------------------------------
    def n_blocks(self) -> int:
        """Blocks up to the largest ID ever used (id_max over all levels)."""
        s = self.refresh_sizes()
        return int(s.id_max[s.max_levels])

    def populations(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        nc = nb * self.sizes.cells_per_block
        out = np.empty((self.sizes.n_q, nc), dtype=self.dtype)
        self._call("download_populations", np_ptr(out), nb, nc)
        return out

    def set_populations(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=self.dtype)
        nc = f.shape[1]
        nb = nc // self.sizes.cells_per_block
        self._call("upload_populations", np_ptr(f), nb, nc)

    def macros(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        nc = nb * self.sizes.cells_per_block
        out = np.empty((4, nc), dtype=self.dtype)
        self._call("download_macros", np_ptr(out), nb, nc)
        return out

    def cell_masks(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty(nb * self.sizes.cells_per_block, dtype=np.int32)
        self._call("download_cell_masks", np_ptr(out, C.c_int), nb)
        return out

    def block_ints(self, which, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        rows = self.sizes.n_q_max if which in (capi.BLK_NBR, capi.BLK_NBR_CHILD) else 1
        out = np.empty((rows, nb), dtype=np.int32)
        self._call("download_block_ints", int(which), np_ptr(out, C.c_int), nb, nb)
        return out if rows > 1 else out[0]

    def set_block_ints(self, which, a: np.ndarray):
        """Overwrite a block table for blocks [0, a.shape[-1]) (M_LoadToGPU for one array)."""
        a = np.ascontiguousarray(np.atleast_2d(a), dtype=np.int32)
        nb = a.shape[1]
        self._call("upload_block_ints", int(which), np_ptr(a, C.c_int), nb, nb)
        self.refresh_sizes()

    def coords(self, n_blocks=None) -> np.ndarray:
        nb = self.n_blocks() if n_blocks is None else int(n_blocks)
        out = np.empty((self.sizes.n_dim, nb), dtype=np.float32)
        self._call("download_coords", np_ptr(out, C.c_float), nb, nb)
        return out

    def id_set(self, L) -> np.ndarray:
        n = int(self.refresh_sizes().n_ids[L])
        out = np.empty(n, dtype=np.int32)
        self._call("download_id_set", int(L), np_ptr(out, C.c_int), n)
        return out

    def gap_set(self) -> np.ndarray:
        n = int(self.refresh_sizes().n_gaps)
        out = np.empty(n, dtype=np.int32)
        self._call("download_gap_set", np_ptr(out, C.c_int), n)
        return out

    def import_pool(self, st: dict):
        """M_LoadToGPU of a whole pool (the dict state() returns) through agal_import_pool."""
        keep = []

        def arr(a, dt=np.int32):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float if dt == np.float32 else C.c_int))

        p = capi.Pool()
        p.n_blocks = int(st["id_max"])
        p.nbr, p.nbr_child = arr(st["nbr"]), arr(st["nbr_child"])
        p.ref, p.level, p.onb, p.blk_mask = arr(st["ref"]), arr(st["level"]), arr(st["onb"]), arr(st["blk_mask"])
        p.coords = arr(st["coords"], np.float32)
        p.cell_mask = arr(st["cell_mask"])
        p.n_ids = arr(np.array(st["n_ids"]))
        sets = [np.asarray(x, dtype=np.int32) for x in st["id_sets"]]
        p.id_sets = arr(np.concatenate(sets) if sets else np.zeros(0))
        p.n_gaps = len(st["gap_set"])
        p.gap_set = arr(st["gap_set"])
        if st.get("f") is not None:
            f = np.ascontiguousarray(st["f"], dtype=self.dtype)
            keep.append(f)
            p.populations = f.ctypes.data_as(C.c_void_p)
        self._call("import_pool", C.byref(p))
        self.refresh_sizes()

    def state(self, with_f: bool = True, with_cells: bool = True) -> dict:
        """Everything the parity tests compare, restricted to active blocks where it matters
        (without the populations when with_f is False, without cell masks and the per-block masks
        when with_cells is False)."""
        s = self.refresh_sizes()
        ML = s.max_levels
        st = {
            "n_ids": [int(s.n_ids[L]) for L in range(ML)],
            "id_max": int(s.id_max[ML]),
            "n_gaps": int(s.n_gaps),
            "id_sets": [self.id_set(L) for L in range(ML)],
            "gap_set": self.gap_set(),
        }
        nb = st["id_max"]
        st["ref"] = self.block_ints(capi.BLK_REF, nb)
        st["level"] = self.block_ints(capi.BLK_LEVEL, nb)
        st["nbr"] = self.block_ints(capi.BLK_NBR, nb)
        st["nbr_child"] = self.block_ints(capi.BLK_NBR_CHILD, nb)
        st["coords"] = self.coords(nb)
        if with_cells:
            st["onb"] = self.block_ints(capi.BLK_ONB, nb)
            st["blk_mask"] = self.block_ints(capi.BLK_MASK, nb)
            st["cell_mask"] = self.cell_masks(nb)
        if with_f:
            st["f"] = self.populations(nb)
        return st
