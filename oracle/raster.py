"""ORACLE (test infrastructure): numpy wrapper of the float64 strand
rasterizer / PPLL resolve transcription in raster_oracle.cc."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .voxel import lib

P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def shadow_depth(xyz, indices, light_vp_cm, line_width, size):
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    m = np.ascontiguousarray(light_vp_cm, np.float32)
    out = np.zeros(size[0] * size[1], np.float32)
    lib.oracle_shadow_depth(_p(xyz), C.c_uint32(len(xyz)), _p(idx), C.c_uint64(len(idx)), _p(m),
                            C.c_float(line_width), C.c_uint32(size[0]), C.c_uint32(size[1]), _p(out))
    return out.reshape(size[1], size[0])


def raster(xyz, tangents, thickness, indices, cam, light, strand, render, density=None, density_f32=None,
           res=(1, 1, 1), shadow_map=None, depth_buffer=None, pairs=None, capacity=1 << 16):
    """Returns dict(segment, pixel, rgba (n,4), depth, kept) of every fragment
    of the listed index pairs (all when None).  capacity: the fragment count
    to allocate for first (a second pass runs when it is exceeded)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    tan = np.ascontiguousarray(tangents, np.float32)
    th = np.ascontiguousarray(thickness, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    pl = None if pairs is None else np.ascontiguousarray(pairs, np.uint64)
    dens = None if density is None else np.ascontiguousarray(density, np.uint8)
    f32 = None if density_f32 is None else np.ascontiguousarray(density_f32, np.float32)
    sm = None if shadow_map is None else np.ascontiguousarray(shadow_map, np.float32)
    sh = (0, 0) if shadow_map is None else (shadow_map.shape[1], shadow_map.shape[0])
    db = None if depth_buffer is None else np.ascontiguousarray(depth_buffer, np.float32)
    lib.oracle_raster.restype = C.c_uint64
    cap = max(int(capacity), 1)
    while True:
        seg = np.zeros(cap, np.uint32)
        pix = np.zeros(cap, np.uint32)
        rgba = np.zeros((cap, 4), np.float64)
        depth = np.zeros(cap, np.float64)
        kept = np.zeros(cap, np.uint8)
        n = lib.oracle_raster(_p(xyz), _p(tan), _p(th), C.c_uint32(len(xyz)), _p(idx), C.c_uint64(len(idx)), _p(pl),
                              C.c_uint64(0 if pl is None else len(pl)), C.byref(cam), C.byref(light),
                              C.byref(strand), C.byref(render), _p(dens), _p(f32), C.c_uint32(res[0]),
                              C.c_uint32(res[1]), C.c_uint32(res[2]), _p(sm), C.c_uint32(sh[0]), C.c_uint32(sh[1]),
                              _p(db), C.c_uint64(cap), _p(seg), _p(pix), _p(rgba), _p(depth), _p(kept))
        if n <= cap:
            return dict(segment=seg[:n], pixel=pix[:n], rgba=rgba[:n], depth=depth[:n], kept=kept[:n].astype(bool))
        cap = int(n)


def resolve(heads, nodes, width, height, background=None):
    h = np.ascontiguousarray(heads, np.uint32)
    n = np.ascontiguousarray(nodes, np.uint32)
    bg = None if background is None else np.ascontiguousarray(background, np.float32)
    out = np.zeros((height * width, 4), np.float64)
    lib.oracle_resolve(_p(h), _p(n), C.c_uint32(len(n) // 3), C.c_uint32(width), C.c_uint32(height), _p(bg),
                       _p(out))
    return out.reshape(height, width, 4)


def pack_unorm4x8(rgba):
    """packUnorm4x8 with ties to even, as the kernels do."""
    q = np.rint(np.clip(np.asarray(rgba, np.float32), 0, 1) * np.float32(255)).astype(np.uint32)
    return q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (q[..., 3] << 24)
