"""CPU restatement of fresnel's path-tracer arithmetic: the parity checker.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU baseline; never by the fresnel_amd package.

`lib()` loads liboracle.so (built from fresnel_oracle.cpp by `build()` /
the Makefile).  `backend` provides the reference `_cpu` class set on top of it.
"""

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'liboracle.so')
_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(['make', '-s', '-B', '-C', _HERE, 'liboracle.so'], check=True)


def lib():
    """The loaded oracle library (ctypes.CDLL) with argument types declared."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    F = ctypes.c_float
    U = ctypes.c_uint
    I = ctypes.c_int
    fp = ctypes.POINTER(ctypes.c_float)
    up = ctypes.POINTER(ctypes.c_uint32)
    bp = ctypes.POINTER(ctypes.c_uint8)
    sz = ctypes.c_size_t
    L.oracle_philox.argtypes = [up, sz, up]
    L.oracle_u01.argtypes = [up, sz, fp]
    L.oracle_scene_new.restype = P
    L.oracle_scene_free.argtypes = [P]
    L.oracle_scene_set_camera.argtypes = [P, fp, fp, fp, F, F, F, F, I]
    L.oracle_scene_set_background.argtypes = [P, fp, F]
    L.oracle_scene_set_lights.argtypes = [P, U, fp, fp, fp]
    L.oracle_scene_set_brute.argtypes = [P, I]
    L.oracle_add_sphere.argtypes = [P, U, I, fp, fp, F, U, fp, fp, fp]
    L.oracle_add_cylinder.argtypes = [P, U, I, fp, fp, F, U, fp, fp, fp]
    L.oracle_add_polyhedron.argtypes = [P, U, I, fp, fp, F, U, fp, fp, fp, U, fp, fp, fp, F, F]
    L.oracle_add_mesh.argtypes = [P, U, I, fp, fp, F, U, fp, fp, U, fp, fp]
    L.oracle_add_polygon.argtypes = [P, U, I, fp, fp, F, U, fp, fp, fp, U, fp, F]
    L.oracle_scene_commit.argtypes = [P]
    L.oracle_scene_num_prims.argtypes = [P]
    L.oracle_scene_num_prims.restype = U
    L.oracle_path_pass.argtypes = [P, U, U, U, U, U, fp, bp, I, fp, I]
    L.oracle_preview.argtypes = [P, U, U, U, U, fp, bp, I, fp, I]
    L.oracle_camera_rays.argtypes = [P, U, U, U, U, I, fp]
    L.oracle_trace.argtypes = [P, fp, sz, F, up, up, fp, fp, I]
    L.oracle_count_ties.argtypes = [P, fp, sz, F, up]
    L.oracle_sphere_t64.argtypes = [fp, fp, fp, F, ctypes.c_double]
    L.oracle_sphere_t64.restype = ctypes.c_double
    L.oracle_cr_sincosf.argtypes = [fp, sz, fp, fp]
    L.oracle_sphere_t64_batch.argtypes = [fp, fp, fp, sz, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def ptr(a, ctype=ctypes.c_float):
    """ctypes pointer to a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(ctypes.POINTER(ctype))
