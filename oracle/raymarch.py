"""ORACLE (test infrastructure): numpy wrapper of the float64 raymarch
transcription in raymarch_oracle.cc."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .voxel import lib

P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def raymarch(density, tangents, res, cam, light, strand, render, width, height, pixels=None,
             density_f32=None, depth_buffer=None, forced=None):
    """Evaluates the listed pixels (all when None).  Returns dict(pixels,
    rgba (n,4) float64, depth (n,), info (n,4) int32, diag (n,4) float64)."""
    nx, ny, nz = res
    if pixels is None:
        pixels = np.arange(width * height, dtype=np.int64)
    pixels = np.ascontiguousarray(pixels, np.int64)
    n = len(pixels)
    rgba = np.zeros((n, 4), np.float64)
    depth = np.zeros(n, np.float64)
    info = np.zeros((n, 4), np.int32)
    diag = np.zeros((n, 4), np.float64)
    dens = None if density is None else np.ascontiguousarray(density, np.uint8)
    tan = None if tangents is None else np.ascontiguousarray(tangents, np.int8)
    f32 = None if density_f32 is None else np.ascontiguousarray(density_f32, np.float32)
    db = None if depth_buffer is None else np.ascontiguousarray(depth_buffer, np.float32)
    fo = None if forced is None else np.ascontiguousarray(forced, np.int32)
    lib.oracle_raymarch(_p(dens), _p(tan), _p(f32), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz),
                        C.byref(cam), C.byref(light), C.byref(strand), C.byref(render), C.c_uint32(width),
                        C.c_uint32(height), _p(db), _p(pixels), C.c_uint64(n), _p(fo), _p(rgba), _p(depth),
                        _p(info), _p(diag))
    return dict(pixels=pixels, rgba=rgba, depth=depth, info=info, diag=diag)


def sample_density(density, res, origin, size, p, density_f32=None):
    nx, ny, nz = res
    lib.oracle_sample_density.restype = C.c_double
    o = np.ascontiguousarray(origin, np.float32)
    s = np.ascontiguousarray(size, np.float32)
    pp = np.ascontiguousarray(p, np.float64)
    d = None if density is None else np.ascontiguousarray(density, np.uint8)
    f = None if density_f32 is None else np.ascontiguousarray(density_f32, np.float32)
    return lib.oracle_sample_density(_p(d), _p(f), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), _p(o), _p(s),
                                     _p(pp))


def kajiya_kay(diffuse, specular, p, tangent, light, eye):
    out = np.zeros(3, np.float64)
    arr = [np.ascontiguousarray(x, np.float64) for x in (diffuse, specular, tangent, light, eye)]
    lib.oracle_kajiya_kay(_p(arr[0]), _p(arr[1]), C.c_double(p), _p(arr[2]), _p(arr[3]), _p(arr[4]), _p(out))
    return out


def volume_adsm(density, res, origin, size, p, light_origin, steps, alpha):
    nx, ny, nz = res
    lib.oracle_volume_adsm.restype = C.c_double
    strands = C.c_double(0)
    o, s = np.ascontiguousarray(origin, np.float32), np.ascontiguousarray(size, np.float32)
    pp, lo = np.ascontiguousarray(p, np.float64), np.ascontiguousarray(light_origin, np.float64)
    d = np.ascontiguousarray(density, np.uint8)
    v = lib.oracle_volume_adsm(_p(d), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), _p(o), _p(s), _p(pp), _p(lo),
                               C.c_float(steps), C.c_double(alpha), C.byref(strands))
    return v, strands.value


def lao(density, res, origin, size, p, radius=2.5, exponent=10.0, clamp=0.16):
    nx, ny, nz = res
    lib.oracle_lao.restype = C.c_double
    o, s = np.ascontiguousarray(origin, np.float32), np.ascontiguousarray(size, np.float32)
    d = np.ascontiguousarray(density, np.uint8)
    pp = np.ascontiguousarray(p, np.float64)
    return lib.oracle_lao(_p(d), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), _p(o), _p(s), _p(pp),
                          C.c_double(radius), C.c_double(exponent), C.c_double(clamp))


def raymarch_grad(tangents, density_f32, res, cam, light, strand, render, diff, width, height, d_rgba=None):
    """Float64 differentiable forward + analytic gradient.  diff = (alpha,
    adsm_thickness, shininess).  Returns (rgba (H,W,4), d_density (N,), d_params (3,))."""
    nx, ny, nz = res
    tan = np.ascontiguousarray(tangents, np.int8)
    f32 = np.ascontiguousarray(density_f32, np.float32)
    dp = np.ascontiguousarray(diff, np.float32)
    g = np.zeros((height * width, 4), np.float32) if d_rgba is None else np.ascontiguousarray(d_rgba, np.float32)
    rgba = np.zeros((height * width, 4), np.float64)
    dd = np.zeros(nx * ny * nz, np.float64)
    dpar = np.zeros(3, np.float64)
    lib.oracle_raymarch_grad(_p(tan), _p(f32), C.c_uint32(nx), C.c_uint32(ny), C.c_uint32(nz), C.byref(cam),
                             C.byref(light), C.byref(strand), C.byref(render), _p(dp), C.c_uint32(width),
                             C.c_uint32(height), _p(g), _p(rgba), _p(dd), _p(dpar))
    return rgba.reshape(height, width, 4), dd, dpar
