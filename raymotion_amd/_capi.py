"""ctypes view of librm_hip.so (include/rm_hip.h, include/rm_scene.h).

The record dtypes below are the byte layouts the reference uploads to its GPU
(src/engine/metal_types.h:3-57). Loading fails loudly when the library has not
been built: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
# RAYMOTION_HIP_LIB selects another build of the same library (the
# host-sanitizer variant librm_hip_san.so).
LIB_PATH = Path(os.environ.get("RAYMOTION_HIP_LIB") or PKG_DIR / "librm_hip.so")

RM_OK = 0
RM_ERR_NO_DEVICE = 1
RM_ERR_INVALID = 2
RM_ERR_ALLOC = 3
RM_ERR_DEVICE = 4
RM_ERR_NO_SCENE = 5

V3 = ("<f4", (3,))
TRIANGLE_DTYPE = np.dtype([
    ("v0", *V3), ("v1", *V3), ("v2", *V3), ("n", *V3), ("nn", *V3), ("n0", *V3), ("n1", *V3), ("n2", *V3),
    ("ar", "<f4"), ("ag", "<f4"), ("ab", "<f4"), ("er", "<f4"), ("eg", "<f4"), ("eb", "<f4"),
    ("ior", "<f4"), ("rough", "<f4"), ("metallic", "<f4"), ("alpha", "<f4"), ("shader", "<i4"),
    ("tu0", "<f4"), ("tv0", "<f4"), ("tu1", "<f4"), ("tv1", "<f4"), ("tu2", "<f4"), ("tv2", "<f4"),
    ("texture_offset", "<i4"), ("texture_width", "<i4"), ("texture_height", "<i4"),
])
NODE_DTYPE = np.dtype([("mn", *V3), ("mx", *V3), ("right", "<i4"), ("offset", "<i4"), ("count", "<i4")])
GBUFFER_DTYPE = np.dtype([
    ("pos", *V3), ("normal", *V3), ("albedo", *V3), ("wo", *V3), ("emission", *V3),
    ("rough", "<f4"), ("metallic", "<f4"), ("ior", "<f4"), ("kind", "<i4"),
])
RESERVOIR_DTYPE = np.dtype([
    ("light_pos", *V3), ("light_normal", *V3), ("emission", *V3), ("alpha", "<f4"), ("p_hat", "<f4"),
    ("sum_ws", "<f4"), ("sum_wu", "<f4"), ("m", "<u4"), ("valid", "<u4"),
])
assert TRIANGLE_DTYPE.itemsize == 176 and NODE_DTYPE.itemsize == 36
assert GBUFFER_DTYPE.itemsize == 76 and RESERVOIR_DTYPE.itemsize == 60


class Camera(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("x", "y", "z", "focal", "m0", "m1", "m2", "m3", "m4", "m5", "m6",
                                             "m7", "m8")]


class Frame(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_int32), ("bounces", ctypes.c_int32),
        ("resolution", ctypes.c_float), ("clamp_value", ctypes.c_float),
        ("adapt_min", ctypes.c_int32), ("adapt_step", ctypes.c_int32),
        ("adapt_rel", ctypes.c_float), ("adapt_abs", ctypes.c_float),
        ("nomis", ctypes.c_int32),
        ("restir_candidates", ctypes.c_int32), ("restir_spatial", ctypes.c_int32), ("restir_mcap", ctypes.c_int32),
        ("restir_radius", ctypes.c_float),
        ("denoise", ctypes.c_int32), ("denoise_iterations", ctypes.c_int32), ("temporal", ctypes.c_int32),
        ("camera", Camera),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("tris", ctypes.c_void_p), ("ntri", ctypes.c_size_t),
        ("nodes", ctypes.c_void_p), ("nnode", ctypes.c_size_t),
        ("bvh_tri", ctypes.c_void_p),
        ("light_tri", ctypes.c_void_p), ("light_cdf", ctypes.c_void_p), ("nlight", ctypes.c_size_t),
        ("light_total", ctypes.c_float),
        ("texels", ctypes.c_void_p), ("ntexel", ctypes.c_size_t),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("path", ctypes.c_int32), ("lanes", ctypes.c_int32), ("has_history", ctypes.c_int32),
        ("refit", ctypes.c_int32),
        ("rays", ctypes.c_uint64 * 6), ("node_visits", ctypes.c_uint64 * 6), ("tri_tests", ctypes.c_uint64 * 6),
        ("ms", ctypes.c_float * 8), ("bvh_nodes", ctypes.c_int32), ("stack_depth", ctypes.c_int32),
        ("scene_ms", ctypes.c_float * 2), ("wave_iters", ctypes.c_int32), ("wave_tail", ctypes.c_uint32),
    ]


# Exported symbols and their prototypes: (name, restype, argtypes).
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_SIGNATURES = [
    ("rm_create", _I, [ctypes.POINTER(_P), _I]),
    ("rm_destroy", None, [_P]),
    ("rm_last_error", ctypes.c_char_p, [_P]),
    ("rm_set_scene", _I, [_P, ctypes.POINTER(SceneDesc)]),
    ("rm_render", _I, [_P, ctypes.POINTER(Frame), ctypes.c_uint64, _P, _P]),
    ("rm_trace", _I, [_P, _P, _P, _P, _SZ, _I, _P, _P, _P, _P]),
    ("rm_trace_device", _I, [_P, _P, _P, _P, _SZ, _I, _P, _P]),
    ("rm_read_tree", _I, [_P, _P, _SZ, _P, _SZ, _P]),
    ("rm_read_restir", _I, [_P, _P, _P, _P, _P]),
    ("rm_restir_merge", _I, [_P, ctypes.POINTER(Frame), ctypes.c_uint64, _P, _P, _P, _P, _P, _P]),
    ("rm_denoise_image", _I, [_P, ctypes.POINTER(Frame), _P, _P, _P, _P, _P, _P]),
    ("rm_render_adjoint", _I, [_P, ctypes.POINTER(Frame), ctypes.c_uint64, _P, _P, _P, _I, _P]),
    ("rm_set_group_materials", _I, [_P, _P, _SZ, _I, _P, _P, _P, _SZ, ctypes.c_float]),
    ("rm_kpn_param_count", _SZ, []),
    ("rm_kpn_set_weights", _I, [_P, _P, _SZ]),
    ("rm_kpn_get_weights", _I, [_P, _P, _SZ]),
    ("rm_render_features", _I, [_P, ctypes.POINTER(Frame), ctypes.c_uint64, _P, _P]),
    ("rm_kpn_dataset_clear", _I, [_P]),
    ("rm_kpn_dataset_add", _I, [_P, _P, _P, _P, _I, _I]),
    ("rm_kpn_train", _I, [_P, _I, _I, ctypes.c_float, ctypes.c_uint64, _P]),
    ("rm_kpn_denoise", _I, [_P, _P, _P, _I, _I, _P]),
    ("rm_kpn_gradient", _I, [_P, _P, _P, _P, _I, _I, _I, _P, _P]),
    ("rm_set_profiling", _I, [_P, _I]),
    ("rm_set_schedule", _I, [_P, _I]),
    ("rm_get_stats", _I, [_P, ctypes.POINTER(Stats)]),
    ("rm_scene_create", _P, []),
    ("rm_scene_destroy", None, [_P]),
    ("rm_scene_clear", None, [_P]),
    ("rm_scene_add_triangles", _I, [_P, _SZ, _P, _P, _P, _P, _P]),
    ("rm_scene_set_positions", _I, [_P, _SZ, _P, _P]),
    ("rm_scene_build_bvh", _I, [_P, _I]),
    ("rm_scene_refit_bvh", _I, [_P]),
    ("rm_scene_counts", None, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    ("rm_scene_nodes", _I, [_P, _P, _P, _P]),
    ("rm_scene_precompute", _I, [_P, _P, _P, _P, _P, _P]),
    ("rm_scene_export", _I, [_P, _P, _P, _P, _P, _P, _P]),
    ("rm_scene_upload", _I, [_P, _P]),
    ("rm_scene_upload_gpu_tree", _I, [_P, _P, _I]),
    ("rm_set_scene_gpu_tree", _I, [_P, _P, _I]),
    ("rm_default_frame", None, [ctypes.POINTER(Frame), _I, _I, _I, _I]),
    ("rm_camera_set", None, [_P, _P, ctypes.c_double, ctypes.POINTER(Camera)]),
    ("rm_camera_orbit", None, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.POINTER(Camera)]),
]
EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C raymotion_amd/csrc)")
        handle = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
        for name, restype, argtypes in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = handle
    return _lib


def ptr(a: np.ndarray | None) -> int | None:
    """Raw data pointer of a C-contiguous array (None stays NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
