"""Builds and loads the scalar C oracle (oracle.c) -- test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "oracle.c")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")

_lib = None


def build():
    """Compiles oracle.c with gcc (-ffp-contract=off keeps every op rounded)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
           SRC, "-o", tmp, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    P, I64, I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.or_level_table.restype = ctypes.c_int
    lib.or_level_table.argtypes = [I32, I32, ctypes.c_float, I32, P, P, P, P, P]
    lib.or_hashgrid_corners.restype = None
    lib.or_hashgrid_corners.argtypes = [P, I64, I32, P, P, P, P, P, P, P]
    lib.or_hashgrid_fwd.restype = None
    lib.or_hashgrid_fwd.argtypes = [P, I64, P, I32, I32, P, P, P, P, P, P]
    lib.or_ray_aabb.restype = None
    lib.or_ray_aabb.argtypes = [P, P, I64, P, P, P]
    lib.or_ray_marching_count.restype = ctypes.c_int64
    lib.or_ray_marching_count.argtypes = [P, P, P, P, I64, P, I32, P, ctypes.c_float, I32, P]
    lib.or_ray_marching_fill.restype = None
    lib.or_ray_marching_fill.argtypes = [P, P, P, P, I64, P, I32, P, ctypes.c_float, I32, P, P, P]
    F32 = ctypes.c_float
    lib.or_march_bg_count.restype = ctypes.c_int64
    lib.or_march_bg_count.argtypes = [P, P, P, P, I64, P, I32, P, F32, F32, I32, P]
    lib.or_march_bg_fill.restype = None
    lib.or_march_bg_fill.argtypes = [P, P, P, P, I64, P, I32, P, F32, F32, I32, P, P, P]
    lib.or_visible_prefix.restype = None
    lib.or_visible_prefix.argtypes = [P, P, P, P, I64, F32, P]
    _lib = lib
    return lib


def p(a):
    """Pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
