"""ORACLE (test infrastructure only): numpy wrapper of the brute-force
strand ray tracer restatement in trace_oracle.cc (the reference's Embree
tracer of flat linear curves, src/vkhr/ray_tracer.cc:76-169)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .voxel import lib

P = C.c_void_p
MISS = 0xFFFFFFFF


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


lib.oracle_trace_draw.restype = C.c_uint32
lib.oracle_trace_draw.argtypes = [C.c_uint32] * 4
lib.oracle_flat_curve.restype = C.c_int
lib.oracle_flat_curve.argtypes = [P, P, P, C.c_float, P, C.c_float, C.c_float, C.c_float, P, P]


def draw(seed, pixel, frame, k) -> int:
    return int(lib.oracle_trace_draw(seed, pixel, frame, k))


def flat_curve(o, d, p0, r0, p1, r1, tnear=0.0, tfar=np.inf):
    """(t, u) of the ray-facing ribbon hit, or None."""
    o, d, p0, p1 = (np.ascontiguousarray(x, np.float32) for x in (o, d, p0, p1))
    t, u = np.zeros(1, np.float32), np.zeros(1, np.float32)
    ok = lib.oracle_flat_curve(_p(o), _p(d), _p(p0), C.c_float(r0), _p(p1), C.c_float(r1), C.c_float(tnear),
                               C.c_float(tfar), _p(t), _p(u))
    return (float(t[0]), float(u[0])) if ok else None


def closest(xyz, radius, indices, rays):
    """rays: (n, 8) float32 (origin, tnear, direction, tfar) -> t, u, segment."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    rad = np.ascontiguousarray(radius, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    rays = np.ascontiguousarray(rays, np.float32)
    n = len(rays)
    t, u, s = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.uint32)
    lib.oracle_trace_closest(_p(xyz), _p(rad), _p(idx), C.c_uint64(len(idx)), _p(rays), C.c_uint64(n), _p(t), _p(u),
                             _p(s))
    return t, u, s


def any_hit(xyz, radius, indices, rays):
    xyz = np.ascontiguousarray(xyz, np.float32)
    rad = np.ascontiguousarray(radius, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(len(rays), np.uint8)
    lib.oracle_trace_any(_p(xyz), _p(rad), _p(idx), C.c_uint64(len(idx)), _p(rays), C.c_uint64(len(rays)), _p(out))
    return out.astype(bool)


def raytrace_frame(xyz, radius, tangents, indices, plane, light_origin, intensity, diffuse, exponent=50.0,
                   method=0, shadows_on=True, ao_radius=2.5, seed=0, frame=0, width=64, height=64, accum=None,
                   ignore_self=False):
    """Adds one sample per pixel (Raytracer::draw) to accum (H, W, 3) float64."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    rad = np.ascontiguousarray(radius, np.float32)
    tan = np.ascontiguousarray(tangents, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    pl = np.ascontiguousarray(plane, np.float32)
    light = np.ascontiguousarray(np.concatenate([light_origin, intensity]), np.float32)
    dif = np.ascontiguousarray(diffuse, np.float32)
    if accum is None:
        accum = np.zeros((height, width, 3), np.float64)
    lib.oracle_raytrace_frame(_p(xyz), _p(rad), _p(tan), _p(idx), C.c_uint64(len(idx)), _p(pl), _p(light), _p(dif),
                              C.c_float(exponent), C.c_int(method), C.c_int(1 if shadows_on else 0),
                              C.c_float(ao_radius), C.c_uint32(seed), C.c_uint32(frame), C.c_uint32(width),
                              C.c_uint32(height), C.c_int(1 if ignore_self else 0), _p(accum))
    return accum
