"""ctypes binding of libpot3d_hip.so (declarations in include/pot3d_hip.h).

The library is the only compute path: if it is missing, importing the
solver raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpot3d_hip.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# name -> (restype, argtypes), as declared in include/pot3d_hip.h
SIGNATURES = {
    "load_external_lu_solver": (None, [_vp, _vp, _vp, C.c_int, C.c_int]),
    "external_lu_solver": (None, [_vp]),
    "unload_external_lu_solver": (None, []),
    "p3d_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int] + [_dp] * 10
                   + [C.c_double, C.c_int]),
    "p3d_destroy": (C.c_int, [_vp]),
    "p3d_last_error": (C.c_char_p, []),
    "p3d_set_option": (C.c_int, [_vp, C.c_int, C.c_double]),
    "p3d_set_br0": (C.c_int, [_vp, _dp]),
    "p3d_set_br1": (C.c_int, [_vp, _dp]),
    "p3d_rhs": (C.c_int, [_vp, _dp]),
    "p3d_solve": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp]),
    "p3d_solve_dev": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, _ip, _dp]),
    "p3d_fill_phi": (C.c_int, [_vp, _dp, _dp]),
    "p3d_energy": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "p3d_readbr": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                             C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp]),
    "p3d_apply_A": (C.c_int, [_vp, _dp, _dp]),
    "p3d_apply_prec": (C.c_int, [_vp, _dp]),
    "p3d_ilu_dinv": (C.c_int, [_vp, _fp]),
    "p3d_get_stats": (C.c_int, [_vp, _vp]),
    "p3d_history": (C.c_int, [_vp, _dp, C.c_int]),
    "p3d_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "p3d_nn_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, _dp, C.c_uint]),
    "p3d_nn_destroy": (C.c_int, [_vp]),
    "p3d_nn_num_params": (C.c_size_t, [_vp]),
    "p3d_nn_set_scales": (C.c_int, [_vp, C.c_double, C.c_double]),
    "p3d_nn_get_params": (C.c_int, [_vp, _fp]),
    "p3d_nn_set_params": (C.c_int, [_vp, _fp]),
    "p3d_nn_add_sample": (C.c_int, [_vp, _dp, _dp]),
    "p3d_nn_loss_grad": (C.c_int, [_vp, _dp, _dp, _dp, _fp]),
    "p3d_nn_train": (C.c_int, [_vp, C.c_int, C.c_double, C.c_uint, _fp]),
    "p3d_nn_set_augment": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "p3d_nn_predict": (C.c_int, [_vp, _dp, _dp, _vp]),
    "p3d_nn_prolong": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                 _ip, _dp, _ip, _dp, _ip, _dp, _dp]),
    "p3d_nn_prolong_dev": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                     _ip, _dp, _ip, _dp, _ip, _dp, _vp]),
    "p3d_nn_prolong_dev_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_int,
                                             C.c_int, _ip, _dp, _ip, _dp, _ip, _dp, _vp]),
    "p3d_nn_mfma_check": (C.c_int, [C.POINTER(C.c_ushort), C.POINTER(C.c_ushort), _fp]),
}

# option keys and error codes (include/pot3d_hip.h)
P3D_OK, P3D_EINVAL, P3D_ENOMEM, P3D_EHIP, P3D_ENOCONV, P3D_ENONFINITE, P3D_ESTATE = range(7)
OPT_STOP, OPT_MIXED, OPT_INNER_TOL, OPT_CHUNK, OPT_GRAPH, OPT_SWEEP, OPT_VERBOSE, \
    OPT_GUESS_CHECK, OPT_TILE_K, OPT_TILE_WAVES, OPT_ILU_BLOCKS_R, OPT_ILU_BLOCKS_T, \
    OPT_ILU_BLOCKS_P, OPT_TILE_STAGE, OPT_STENCIL, OPT_APPLY_FP32, OPT_FUSE_RESIDUAL, \
    OPT_BLOCK_SWEEP = range(1, 19)


class Stats(C.Structure):
    _fields_ = [("ncg", C.c_int), ("n_outer", C.c_int), ("status", C.c_int), ("guess", C.c_int),
                ("epsn", C.c_double), ("relres", C.c_double), ("true_relres", C.c_double),
                ("t_solve", C.c_double), ("t_setup", C.c_double)]


_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m pot3d_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return load().p3d_last_error().decode()


class Pot3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"p3d error {code}: {msg}")
        self.code = code


def check(rc: int):
    if rc != P3D_OK:
        raise Pot3dError(rc, last_error())
    return rc
