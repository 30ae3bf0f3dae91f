"""ctypes binding of libvkhr_hip.so (the C ABI declared in include/vkhr_hip.h).

This module only loads and describes the library: there is no fallback.  If
the shared object is missing or does not export a declared symbol, importing
raises, so nothing can silently run without the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent
LIB_PATH = PKG_DIR / "libvkhr_hip.so"
HEADER = ROOT / "include" / "vkhr_hip.h"

VKHR_OK = 0
STATUS = {0: "VKHR_OK", -1: "VKHR_ERR_INVALID", -2: "VKHR_ERR_HIP", -3: "VKHR_ERR_OVERFLOW",
          -4: "VKHR_ERR_NOMEM", -5: "VKHR_ERR_NO_DEVICE"}


class VkhrError(RuntimeError):
    def __init__(self, fn: str, code: int, message: str):
        super().__init__(f"{fn} -> {STATUS.get(code, code)}: {message}")
        self.code = code


class AABB(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("radius", C.c_float), ("size", C.c_float * 3), ("volume", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("projection", C.c_float * 16), ("position", C.c_float * 3),
                ("look_at_distance", C.c_float), ("znear", C.c_float), ("zfar", C.c_float),
                ("resolution", C.c_float * 2)]


class Light(C.Structure):
    _fields_ = [("vector", C.c_float * 4), ("intensity", C.c_float * 4), ("matrix", C.c_float * 16),
                ("origin", C.c_float * 3), ("znear", C.c_float), ("zfar", C.c_float), ("pad_", C.c_float * 3)]


class StrandParams(C.Structure):
    _fields_ = [("volume_bounds", AABB), ("volume_resolution", C.c_float * 3), ("strand_radius", C.c_float),
                ("hair_color", C.c_float * 3), ("hair_opacity", C.c_float), ("hair_shininess", C.c_float),
                ("strand_ratio", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("shading_model", C.c_int32), ("adsm_kernel_size", C.c_int32), ("adsm_sampling_type", C.c_int32),
                ("adsm_stride_size", C.c_int32), ("adsm_on", C.c_int32), ("ctsm_kernel_size", C.c_int32),
                ("ctsm_sampling_type", C.c_int32), ("ctsm_bias", C.c_float), ("ctsm_on", C.c_int32),
                ("shadow_technique", C.c_int32), ("isosurface", C.c_float), ("raycast_steps", C.c_float),
                ("occlusion_radius", C.c_float), ("ao_exponent", C.c_float), ("ao_clamp", C.c_float),
                ("lod_magnified_distance", C.c_float), ("renderer", C.c_int32),
                ("lod_minified_distance", C.c_float), ("benchmarking", C.c_int32)]


class Volume(C.Structure):
    _fields_ = [("density", C.c_void_p), ("cells", C.c_void_p), ("tangents", C.c_void_p),
                ("density_f32", C.c_void_p), ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("reserved_", C.c_uint32), ("occupancy", C.c_void_p)]


class Ppll(C.Structure):
    _fields_ = [("heads", C.c_void_p), ("nodes", C.c_void_p), ("counter", C.c_void_p), ("node_count", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("reserved_", C.c_uint32)]


class FragCapture(C.Structure):
    _fields_ = [("segment", C.c_void_p), ("pixel", C.c_void_p), ("color", C.c_void_p)]


class DiffParams(C.Structure):
    _fields_ = [("absorption", C.c_float), ("adsm_thickness", C.c_float), ("shininess", C.c_float),
                ("reserved_", C.c_float)]


class BatchStyle(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("n_vertices", C.c_uint32), ("n_indices", C.c_uint32),
                ("indices", C.c_void_p), ("tangents", C.c_void_p), ("bounds", AABB), ("volume", C.c_uint32),
                ("reserved_", C.c_uint32), ("counts", C.c_void_p), ("tangent_acc", C.c_void_p),
                ("density", C.c_void_p), ("tangent_i8x4", C.c_void_p), ("cells", C.c_void_p), ("occupancy", C.c_void_p),
                ("rgba", C.c_void_p),
                ("camera", Camera), ("light", Light), ("strand", StrandParams), ("params", RenderParams)]


class AdjointStyle(C.Structure):
    _fields_ = [("volume", Volume), ("d_density", C.c_void_p), ("d_params", C.c_void_p)]


class AdjointRender(C.Structure):
    _fields_ = [("style", C.c_uint32), ("reserved_", C.c_uint32), ("camera", Camera), ("light", Light),
                ("strand", StrandParams), ("params", RenderParams), ("diff", DiffParams), ("target", C.c_void_p),
                ("rgba", C.c_void_p), ("record", C.c_void_p), ("loss", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("groups", C.c_uint32)]


class ViewingPlane(C.Structure):
    _fields_ = [("x", C.c_float * 3), ("y", C.c_float * 3), ("z", C.c_float * 3), ("point", C.c_float * 3)]


class TraceParams(C.Structure):
    _fields_ = [("method", C.c_uint32), ("shadows_on", C.c_int32), ("ao_radius", C.c_float),
                ("exponent", C.c_float), ("diffuse", C.c_float * 3), ("seed", C.c_uint32), ("frame", C.c_uint32),
                ("ignore_self", C.c_int32)]


P = C.c_void_p
U32 = C.c_uint32
I32 = C.c_int

# name -> (argtypes); every function returns int status except noted.
SIGNATURES = {
    "vkhr_abi_version": [],
    "vkhr_init": [],
    "vkhr_release_scratch": [P],
    "vkhr_aabb_compute": [P, U32, C.POINTER(AABB), P],
    "vkhr_generate_indices": [P, U32, P, P],
    "vkhr_sort_segments": [P, U32, P, U32, C.POINTER(AABB), P, P],
    "vkhr_sort_strands": [P, U32, P, U32, C.POINTER(AABB), P, P],
    "vkhr_generate_tangents": [P, P, U32, P, P],
    "vkhr_voxelize_vertices": [P, U32, P, C.POINTER(AABB), U32, U32, U32, P, P, P],
    "vkhr_voxelize_segments": [P, U32, P, U32, P, C.POINTER(AABB), U32, U32, U32, P, P, P],
    "vkhr_volume_finalize": [P, P, U32, U32, U32, I32, P, P, P, P, P, P],
    "vkhr_volume_finalize_noalloc": [P, P, U32, U32, U32, I32, P, P, P, P, P, P],
    "vkhr_volume_build_segments": [P, U32, P, U32, P, C.POINTER(AABB), U32, U32, U32, I32, P, P, P, P, P, P, P, P],
    "vkhr_raymarch": [C.POINTER(Volume), C.POINTER(Camera), C.POINTER(Light), C.POINTER(StrandParams),
                      C.POINTER(RenderParams), U32, U32, P, P, P, P, P],
    "vkhr_raymarch_ppll": [C.POINTER(Volume), C.POINTER(Camera), C.POINTER(Light), C.POINTER(StrandParams),
                           C.POINTER(RenderParams), U32, U32, P, C.POINTER(Ppll), P],
    "vkhr_ppll_clear": [C.POINTER(Ppll), P],
    "vkhr_shadow_depth": [P, U32, P, U32, P, C.c_float, U32, U32, P, P],
    "vkhr_raster_strands": [P, P, P, U32, P, U32, C.POINTER(Camera), C.POINTER(Light), C.POINTER(StrandParams),
                            C.POINTER(RenderParams), C.POINTER(Volume), P, U32, U32, P, C.POINTER(Ppll),
                            C.POINTER(FragCapture), P],
    "vkhr_ppll_resolve": [C.POINTER(Ppll), P, P, P],
    "vkhr_raymarch_record": [C.POINTER(Volume), C.POINTER(Camera), C.POINTER(Light), C.POINTER(StrandParams),
                             C.POINTER(RenderParams), C.POINTER(DiffParams), U32, U32, P, P, P, P, P],
    "vkhr_raymarch_adjoint": [C.POINTER(Volume), C.POINTER(Camera), C.POINTER(Light), C.POINTER(StrandParams),
                              C.POINTER(RenderParams), C.POINTER(DiffParams), U32, U32, P, P, P, P, P],
    "vkhr_batch_create": [C.POINTER(BatchStyle), U32, U32, U32, C.POINTER(Tuning), C.POINTER(C.c_void_p)],
    "vkhr_batch_launch": [P, P],
    "vkhr_batch_destroy": [P],
    "vkhr_adjoint_batch_create": [C.POINTER(AdjointStyle), U32, C.POINTER(AdjointRender), U32, U32, U32,
                                  C.POINTER(C.c_void_p)],
    "vkhr_adjoint_batch_launch": [P, P],
    "vkhr_adjoint_batch_destroy": [P],
    "vkhr_volume_downsample": [P, U32, U32, U32, I32, P, P],
    "vkhr_bvh_build": [P, P, U32, P, U32, P, C.POINTER(C.c_void_p)],
    "vkhr_bvh_destroy": [P],
    "vkhr_bvh_info": [P, P, P],
    "vkhr_bvh_nodes": [P, P, P, P, P],
    "vkhr_trace_rays": [P, P, U32, I32, P, P, P, P],
    "vkhr_raytrace": [P, P, P, C.POINTER(ViewingPlane), C.POINTER(Light), C.POINTER(TraceParams), U32, U32, P, P],
    "vkhr_volume_filter": [P, U32, U32, U32, U32, P, P],
}
# entry points that do not return a status
SIZE_FUNCTIONS = {"vkhr_march_record_bytes": [U32, U32], "vkhr_occupancy_words": [U32, U32, U32],
                  "vkhr_scratch_bytes": [P]}


def declared_symbols() -> list[str]:
    """Every function name declared in include/vkhr_hip.h."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vkhr_[a-z0-9_]+)\s*\(", text)))


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C vkhr_amd)")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loading torch first makes the library bind to that one instance, so
    # device pointers and streams are shared; two runtimes in one process
    # would not see each other's allocations.
    import torch  # noqa: F401
    lib = C.CDLL(str(LIB_PATH))
    lib.vkhr_last_error.restype = C.c_char_p
    lib.vkhr_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    for name, args in SIZE_FUNCTIONS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_size_t
    return lib


lib = _load()


def _check_args(name: str, fn, args) -> None:
    """Catches arguments passed in the wrong position before they reach the
    device: a small integer where a pointer goes, or a pointer-sized value
    where a 32-bit integer goes."""
    for i, (v, t) in enumerate(zip(args, fn.argtypes or ())):
        if not isinstance(v, int) or isinstance(v, bool):
            continue
        if t is C.c_void_p and 0 < v < 65536:
            raise ValueError(f"{name}: argument {i} is {v}, not a pointer")
        if t in (C.c_uint32, C.c_int) and not (-(1 << 31) <= v < (1 << 32)):
            raise ValueError(f"{name}: argument {i} ({v:#x}) does not fit 32 bits")


def call(name: str, *args) -> None:
    """Calls a status-returning entry point and raises VkhrError on failure."""
    fn = getattr(lib, name)
    _check_args(name, fn, args)
    rc = fn(*args)
    if rc != VKHR_OK:
        raise VkhrError(name, rc, lib.vkhr_last_error().decode(errors="replace"))


def ptr(t) -> int | None:
    """Device/host address of a torch tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_of(torch_stream=None):
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def make_aabb(d: dict) -> AABB:
    a = AABB()
    for i in range(3):
        a.origin[i] = float(d["origin"][i])
        a.size[i] = float(d["size"][i])
    a.radius = float(d["radius"])
    a.volume = float(d["volume"])
    return a


def aabb_dict(a: AABB) -> dict:
    import numpy as np
    return dict(origin=np.array(a.origin[:], np.float32), size=np.array(a.size[:], np.float32),
                radius=np.float32(a.radius), volume=np.float32(a.volume))
