"""ORACLE (test infrastructure): numpy-facing wrapper of liboracle.so's
strict-IEEE restatement of src/vkhr/scene_graph/hair_style.cc:171-357."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
    return C.CDLL(str(LIB_PATH))


lib = _load()
P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def aabb(xyz: np.ndarray) -> dict:
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros(8, np.float32)
    lib.oracle_aabb(_p(xyz), C.c_uint64(len(xyz)), _p(out))
    return dict(origin=out[0:3].copy(), radius=np.float32(out[3]), size=out[4:7].copy(), volume=np.float32(out[7]))


def aabb_array(b: dict) -> np.ndarray:
    return np.array([*b["origin"], b["radius"], *b["size"], b["volume"]], np.float32)


def generate_indices(offsets: np.ndarray) -> np.ndarray:
    offsets = np.ascontiguousarray(offsets, np.uint32)
    n = len(offsets) - 1
    n_idx = 2 * (int(offsets[-1]) - n)
    out = np.zeros(max(n_idx, 1), np.uint32)
    lib.oracle_generate_indices(_p(offsets), C.c_uint32(n), _p(out))
    return out[:n_idx]


def generate_tangents(xyz: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float32)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    out = np.zeros_like(xyz)
    lib.oracle_generate_tangents(_p(xyz), _p(offsets), C.c_uint32(len(offsets) - 1), _p(out))
    return out


def voxelize(xyz, bounds: dict, res, indices=None, tangents=None, want_sums=False, want_sums_all=True):
    """Returns dict(counts=u32, dens_u8=u8 (reference saturating view),
    sums_ref (reference's first-255 tangent sums), sums_all) for vertex mode
    (indices None) or segment mode.  want_sums_all=False leaves out the sums
    over every contribution (the reference keeps only sums_ref)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    nx, ny, nz = res
    n = nx * ny * nz
    counts = np.zeros(n, np.uint32)
    dens = np.zeros(n, np.uint8)
    sums_ref = np.zeros(3 * n, np.float32) if want_sums else None
    sums_all = np.zeros(3 * n, np.float32) if want_sums and want_sums_all else None
    tan = None if tangents is None else np.ascontiguousarray(tangents, np.float32)
    b = aabb_array(bounds)
    if indices is None:
        lib.oracle_voxelize_vertices(_p(xyz), C.c_uint64(len(xyz)), _p(tan), _p(b), C.c_uint32(nx),
                                     C.c_uint32(ny), C.c_uint32(nz), _p(counts), _p(dens), _p(sums_ref),
                                     _p(sums_all))
    else:
        idx = np.ascontiguousarray(indices, np.uint32)
        lib.oracle_voxelize_segments(_p(xyz), _p(idx), C.c_uint64(len(idx)), _p(tan), _p(b), C.c_uint32(nx),
                                     C.c_uint32(ny), C.c_uint32(nz), _p(counts), _p(dens), _p(sums_ref),
                                     _p(sums_all))
    return dict(counts=counts, dens_u8=dens, sums_ref=sums_ref, sums_all=sums_all)


def quantize_tangents(sums: np.ndarray, denominators: np.ndarray) -> np.ndarray:
    sums = np.ascontiguousarray(sums, np.float32)
    den = np.ascontiguousarray(denominators, np.uint32)
    out = np.zeros(4 * len(den), np.int8)
    lib.oracle_quantize_tangents(_p(sums), _p(den), C.c_uint64(len(den)), _p(out))
    return out


def normalize(dens: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(dens, np.uint8).copy()
    lib.oracle_normalize(_p(d), C.c_uint64(len(d)))
    return d


def voxelize_counts_omp(xyz, bounds: dict, res, indices) -> np.ndarray:
    """Segment-mode counts only, the DDA split over the OpenMP team (a
    multi-core timing variant; the same counts as voxelize)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    nx, ny, nz = res
    counts = np.zeros(nx * ny * nz, np.uint32)
    lib.oracle_voxelize_segments_counts_omp(_p(xyz), _p(idx), C.c_uint64(len(idx)), _p(aabb_array(bounds)),
                                            C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), _p(counts))
    return counts


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP team (pixel loops, count-only DDA)."""
    lib.oracle_set_threads(C.c_int(n))


def max_threads() -> int:
    lib.oracle_max_threads.restype = C.c_int
    return int(lib.oracle_max_threads())
