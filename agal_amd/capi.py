"""ctypes view of the C ABI declared in include/agal_capi.h.

The structures mirror `agal_params` / `agal_sizes` field for field.  `load_library()` loads the
HIP shared library built in-tree (agal_amd/lib/libagal.so) and raises if it is missing: there is
no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libagal.so")

# ---- constants (include/agal_capi.h) -------------------------------------------------------
AGAL_OK = 0
AGAL_ERR_ARG = 1
AGAL_ERR_HIP = 2
AGAL_ERR_ALLOC = 3
AGAL_ERR_CAPACITY = 4
AGAL_ERR_POOL_EXHAUSTED = 5
AGAL_ERR_STATE = 6
AGAL_ERR_OVERFLOW = 7
AGAL_ERR_NOT_BUILT = 8
# Size of the fixed staging area of agal_download_populations_noalloc (AGAL_NOALLOC_SCRATCH_BYTES).
AGAL_NOALLOC_SCRATCH_BYTES = 64 << 20

D2Q9, D3Q19, D3Q27 = 9, 19, 27
FP32, FP64, FP16, BF16 = 0, 1, 2, 3
BGK, MRT = 0, 2
BC_NONE, BC_LDC2D, BC_LDC3D, BC_FPSC2D, BC_FPSC3D = 0, 1, 2, 3, 4
INTERP_LINEAR, INTERP_CUBIC = 0, 2          # interpolation order (agal_params.interp_order)
INTERP_INTERFACE, INTERP_ADDED = 0, 1       # interpolation mode (agal_interpolate var)
AVERAGE_INTERFACE, AVERAGE_BLOCK, AVERAGE_GRID = 0, 1, 2
REF_UNREFINED, REF_REFINED, REF_REFINED_WCHILD, REF_REFINED_PERM = 0, 1, 2, 3
REF_MARK_REFINE, REF_MARK_COARSEN, REF_NEW, REF_REMOVE, REF_INACTIVE = 4, 5, 6, 7, 8
REF_UNREFINED_VIO, REF_REFINED_WCHILD_PERM = 9, 10
N_SKIPID = -999
CELL_INTERIOR, CELL_INTERFACE, CELL_GHOST, CELL_SOLID, CELL_BOUNDARY = 0, 1, 2, -1, -6
BLOCKMASK_REGULAR, BLOCKMASK_INTERFACE, BLOCKMASK_SOLIDB, BLOCKMASK_SOLIDA, BLOCKMASK_SOLID = 0, 1, -1, -2, -3
BLK_NBR, BLK_NBR_CHILD, BLK_REF, BLK_LEVEL, BLK_ONB, BLK_MASK = 0, 1, 2, 3, 4, 5
OPT_GRAPHS, OPT_MRT_MFMA, OPT_EXC_CAP, OPT_GRAPH_STEPS, OPT_LINK_MASKS = 1, 2, 3, 4, 5


class Params(C.Structure):
    _fields_ = [
        ("velocity_set", C.c_int),
        ("precision", C.c_int),
        ("collision", C.c_int),
        ("interp_order", C.c_int),
        ("Nx", C.c_int),
        ("L_c", C.c_double),
        ("L_fy", C.c_double),
        ("L_fz", C.c_double),
        ("PERIODIC_X", C.c_int),
        ("PERIODIC_Y", C.c_int),
        ("PERIODIC_Z", C.c_int),
        ("MAX_LEVELS_WALL", C.c_int),
        ("MAX_LEVELS_INTERIOR", C.c_int),
        ("N_LEVEL_START", C.c_int),
        ("v0", C.c_double),
        ("S_CRITERION", C.c_int),
        ("N_REFINE_START", C.c_double),
        ("N_REFINE_INC", C.c_double),
        ("N_REFINE_MAX", C.c_double),
        ("P_REFINE", C.c_int),
        ("bc_preset", C.c_int),
        ("bc_velocity", C.c_double),
        ("ic_rho", C.c_double),
        ("ic_u", C.c_double),
        ("ic_v", C.c_double),
        ("ic_w", C.c_double),
        ("s1", C.c_double),
        ("s2", C.c_double),
        ("s3", C.c_double),
        ("s4", C.c_double),
        ("s5", C.c_double),
        ("s6", C.c_double),
        ("G_NEAR_WALL_DISTANCE", C.c_double),
        ("G_BIN_DENSITY", C.c_int),
        ("G_BIN_LEVELS", C.c_int),
        ("G_BIN_SPEC", C.c_int),
        ("G_BIN_MAKE_LINKS", C.c_int),
        ("S_BC_TYPE", C.c_int),
        ("S_FORCE_TYPE", C.c_int),
        ("S_FORCE_ORDER", C.c_int),
        ("S_FORCEVOLUME_Xm", C.c_double),
        ("S_FORCEVOLUME_XM", C.c_double),
        ("S_FORCEVOLUME_Ym", C.c_double),
        ("S_FORCEVOLUME_YM", C.c_double),
        ("S_FORCEVOLUME_Zm", C.c_double),
        ("S_FORCEVOLUME_ZM", C.c_double),
    ]


class LearnOpts(C.Structure):
    _fields_ = [
        ("train_steps", C.c_int),
        ("infer_every", C.c_int),
        ("batch", C.c_int),
        ("adam_steps", C.c_int),
        ("lr", C.c_double),
        ("threshold", C.c_double),
        ("pos_weight", C.c_double),
        ("seed", C.c_ulonglong),
        ("use_predictions", C.c_int),
    ]


class LearnStats(C.Structure):
    _fields_ = [
        ("flagged", C.c_longlong),
        ("missed", C.c_longlong),
        ("flagged_train", C.c_longlong),
        ("missed_train", C.c_longlong),
        ("predicted_refinements", C.c_longlong),
        ("adam_steps_done", C.c_int),
        ("rebuilds", C.c_int),
        ("early_rebuilds", C.c_int),
        ("last_loss", C.c_double),
        ("step_ms", C.c_double),
        ("rebuild_ms", C.c_double),
        ("predictor_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_NFEAT, MLP_NPARAM = 12, 5057


class Pool(C.Structure):
    """agal_pool: a whole block pool for agal_import_pool (M_LoadToGPU)."""
    _fields_ = [
        ("n_blocks", C.c_longlong),
        ("nbr", C.POINTER(C.c_int)), ("nbr_child", C.POINTER(C.c_int)),
        ("ref", C.POINTER(C.c_int)), ("level", C.POINTER(C.c_int)), ("onb", C.POINTER(C.c_int)),
        ("blk_mask", C.POINTER(C.c_int)),
        ("coords", C.POINTER(C.c_float)),
        ("cell_mask", C.POINTER(C.c_int)),
        ("n_ids", C.POINTER(C.c_int)),
        ("id_sets", C.POINTER(C.c_int)),
        ("n_gaps", C.c_int),
        ("gap_set", C.POINTER(C.c_int)),
        ("populations", C.c_void_p),
    ]


class Sizes(C.Structure):
    _fields_ = [
        ("n_dim", C.c_int),
        ("n_q", C.c_int),
        ("n_q_max", C.c_int),
        ("cells_per_block", C.c_int),
        ("n_children", C.c_int),
        ("max_levels", C.c_int),
        ("n_maxcblocks", C.c_longlong),
        ("n_maxcells", C.c_longlong),
        ("n_coarsecblocks", C.c_int),
        ("n_ids", C.c_int * 8),
        ("id_max", C.c_int * 9),
        ("n_gaps", C.c_int),
    ]


def default_params(**kw) -> Params:
    """The reference parser's defaults (inc/parser.h:62-144) plus the presets named here."""
    p = Params()
    p.velocity_set = D3Q19
    p.precision = FP32
    p.collision = BGK
    p.interp_order = 0
    p.Nx = 64
    p.L_c, p.L_fy, p.L_fz = 1.0, 1.0, 1.0
    p.PERIODIC_X = p.PERIODIC_Y = p.PERIODIC_Z = 0
    p.MAX_LEVELS_WALL = 5
    p.MAX_LEVELS_INTERIOR = 4
    p.N_LEVEL_START = 0
    p.v0 = 1.5625e-5
    p.S_CRITERION = 0
    p.N_REFINE_START, p.N_REFINE_INC, p.N_REFINE_MAX = -3.0, 1.0, 1.0
    p.P_REFINE = 32
    p.bc_preset = BC_LDC3D
    p.bc_velocity = 0.05
    p.ic_rho, p.ic_u, p.ic_v, p.ic_w = 1.0, 0.0, 0.0, 0.0
    p.s1, p.s2, p.s3, p.s4, p.s5, p.s6 = 1.19, 1.4, 1.2, 1.2, 1.2, 1.98
    p.G_NEAR_WALL_DISTANCE = 0.05
    p.G_BIN_DENSITY, p.G_BIN_LEVELS, p.G_BIN_SPEC, p.G_BIN_MAKE_LINKS = 1, 1, 1, 0
    p.S_BC_TYPE, p.S_FORCE_TYPE, p.S_FORCE_ORDER = 0, 0, 0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class AgalError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        super().__init__(f"{where} failed with status {status}" + (f": {detail}" if detail else ""))


_P = C.c_void_p
_I = C.c_int
_LL = C.c_longlong

# name -> (restype, argtypes) without the prefix; shared by the HIP library and the test oracle.
PROTOTYPES = {
    "create": (_I, [C.POINTER(Params), _LL, _I, C.POINTER(_P)]),
    "destroy": (_I, [_P]),
    "get_sizes": (_I, [_P, C.POINTER(Sizes)]),
    "sync": (_I, [_P]),
    "set_ic": (_I, [_P, _I]),
    "advance": (_I, [_P, _I]),
    "collide_stream": (_I, [_P, _I]),
    "interpolate": (_I, [_P, _I, _I]),
    "average": (_I, [_P, _I, _I]),
    "refresh": (_I, [_P, _I]),
    "ref_criteria": (_I, [_P]),
    "refine_coarsen": (_I, [_P, C.POINTER(_I)]),
    "refine_with_solution": (_I, [_P, C.POINTER(_I)]),
    "mark_uniform": (_I, [_P, _I]),
    "mark_blocks": (_I, [_P, C.POINTER(_I), _I, _I]),
    "freeze_refined": (_I, [_P, _I]),
    "advance_loop": (_I, [_P, _I, _I, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "download_populations": (_I, [_P, _P, _LL, _LL]),
    "download_links": (_I, [_P, _P, _LL, _P, _LL, _P]),
    "download_link_bins": (_I, [_P, _P, _LL, _P, _LL, _P]),
    "download_populations_noalloc": (_I, [_P, _P, _LL, _LL]),
    "upload_populations": (_I, [_P, _P, _LL, _LL]),
    "download_macros": (_I, [_P, _P, _LL, _LL]),
    "download_cell_masks": (_I, [_P, C.POINTER(_I), _LL]),
    "upload_cell_masks": (_I, [_P, C.POINTER(_I), _LL]),
    "upload_block_ints": (_I, [_P, _I, C.POINTER(_I), _LL, _LL]),
    "download_block_ints": (_I, [_P, _I, C.POINTER(_I), _LL, _LL]),
    "download_coords": (_I, [_P, C.POINTER(C.c_float), _LL, _LL]),
    "download_id_set": (_I, [_P, _I, C.POINTER(_I), _I]),
    "download_gap_set": (_I, [_P, C.POINTER(_I), _I]),
    "import_pool": (_I, [_P, C.POINTER(Pool)]),
    "set_option": (_I, [_P, _I, _LL]),
    "timing_reset": (_I, [_P, _I]),
    "timing_get": (_I, [_P, C.POINTER(C.c_double), C.POINTER(_LL), C.POINTER(C.c_double)]),
    "refine_near_wall": (_I, [_P]),
    "set_geometry": (_I, [_P, C.POINTER(C.c_float), _LL]),
    "voxelize": (_I, [_P, _I]),
    "check_masks": (_I, [_P, _I]),
    "forces": (_I, [_P, C.POINTER(C.c_double)]),
    "predictor_init": (_I, [_P, C.c_ulonglong, _I]),
    "predictor_set": (_I, [_P, C.POINTER(C.c_float)]),
    "predictor_get": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "predictor_forward": (_I, [_P, C.POINTER(C.c_float), _I, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "predictor_train_step": (_I, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), _I, C.c_double, C.c_double,
                                  C.POINTER(C.c_double)]),
    "predictor_features": (_I, [_P, C.POINTER(C.c_float), _LL]),
    "criterion_labels": (_I, [_P, C.POINTER(_I), _LL]),
    "advance_loop_learned": (_I, [_P, _I, _I, C.POINTER(LearnOpts), C.POINTER(LearnStats)]),
}

# Host helpers without a context.
HOST_PROTOTYPES = {
    "icosphere": (_LL, [C.c_double, C.c_double, C.c_double, C.c_double, _I, C.c_double, C.c_ulonglong,
                        C.POINTER(C.c_float), _LL]),
    "load_stl": (_LL, [C.c_char_p, C.POINTER(C.c_float), _LL]),
    "learn_opts_default": (None, [C.POINTER(LearnOpts)]),
}

# Every symbol include/agal_capi.h declares (checked by tests/test_capi_symbols.py).
EXPORTED = ["agal_params_default", "agal_status_string", "agal_last_error"] + [
    "agal_" + k for k in list(PROTOTYPES) + list(HOST_PROTOTYPES)
]


def bind(lib: C.CDLL, prefix: str, names) -> dict:
    fns = {}
    for name in names:
        restype, argtypes = PROTOTYPES[name]
        f = getattr(lib, prefix + name)
        f.restype = restype
        f.argtypes = argtypes
        fns[name] = f
    return fns


_LIB = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises FileNotFoundError when it has not been built."""
    global _LIB
    if _LIB is not None and path == LIB_PATH:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    lib.agal_status_string.restype = C.c_char_p
    lib.agal_status_string.argtypes = [_I]
    lib.agal_last_error.restype = C.c_char_p
    lib.agal_last_error.argtypes = [_P]
    lib.agal_params_default.restype = None
    lib.agal_params_default.argtypes = [C.POINTER(Params)]
    for name, (res, args) in HOST_PROTOTYPES.items():
        f = getattr(lib, "agal_" + name)
        f.restype, f.argtypes = res, args
    if path == LIB_PATH:
        _LIB = lib
    return lib


def np_ptr(a: np.ndarray, ctype=None):
    if ctype is None:
        return a.ctypes.data_as(C.c_void_p)
    return a.ctypes.data_as(C.POINTER(ctype))


def icosphere(center, radius: float, subdivisions: int, noise: float = 0.0, seed: int = 0) -> np.ndarray:
    """Triangles (n, 3, 3) float32 of the seeded noisy icosphere built by agal_icosphere."""
    lib = load_library()
    n = lib.agal_icosphere(*map(float, center), float(radius), int(subdivisions), float(noise), int(seed), None, 0)
    if n < 0:
        raise AgalError(AGAL_ERR_ARG, "agal_icosphere")
    out = np.empty((n, 3, 3), dtype=np.float32)
    m = lib.agal_icosphere(*map(float, center), float(radius), int(subdivisions), float(noise), int(seed),
                           np_ptr(out, C.c_float), n)
    if m != n:
        raise AgalError(AGAL_ERR_ARG, "agal_icosphere")
    return out


def load_stl(path: str) -> np.ndarray:
    """Triangles (n, 3, 3) float32 of an ASCII STL file (agal_load_stl)."""
    lib = load_library()
    n = lib.agal_load_stl(path.encode(), None, 0)
    if n < 0:
        raise AgalError(AGAL_ERR_ARG, f"agal_load_stl({path})", f"code {n}")
    out = np.empty((n, 3, 3), dtype=np.float32)
    if lib.agal_load_stl(path.encode(), np_ptr(out, C.c_float), n) != n:
        raise AgalError(AGAL_ERR_ARG, f"agal_load_stl({path})")
    return out


def learn_opts(**kw) -> LearnOpts:
    """agal_learn_opts_default() with the given fields replaced."""
    o = LearnOpts()
    load_library().agal_learn_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o
