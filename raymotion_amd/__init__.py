"""raymotion_amd: MI355X (gfx950) rendering backend for raymotion scenes.

The GPU path lives in librm_hip.so (hand-written HIP kernels behind the C-ABI
of include/rm_hip.h); this package is the Python host side: scene building
through the C++ engine, the backend context, and the benchmark scenes.
"""
from .scene import Backend, BackendError, Exported, Scene, camera_struct, default_frame, material_rows
from ._capi import Camera, Frame, Stats, lib

__all__ = ["Backend", "BackendError", "Exported", "Scene", "camera_struct", "default_frame", "material_rows",
           "Camera", "Frame", "Stats", "lib"]
