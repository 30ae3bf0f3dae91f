"""The strand ray tracer over torch device buffers (the C ABI's vkhr_bvh_* /
vkhr_trace_rays / vkhr_raytrace): the reference's Embree renderer,
vkhr::Raytracer (src/vkhr/ray_tracer.cc:28-169), as a ground truth for the
raymarcher's shadows and ambient occlusion."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import capi
from .capi import call, ptr

METHODS = dict(shaded=0, combined_shadows=1, direct_shadows=2, ambient_occlusion=3)


def viewing_plane(eye, center, up, fov_y: float, width: int, height: int) -> capi.ViewingPlane:
    """Camera::recalculate_viewing_plane (camera.cc:251-266) in float32:
    x = left = up x forward, y = x x forward, z = the direction through the
    corner of pixel (0, 0); pixel (i, j) looks along (i+jx) x + (j+jy) y + z."""
    f32 = np.float32
    eye, center, up = (np.asarray(v, np.float64).astype(f32) for v in (eye, center, up))
    fwd = center - eye
    fwd = (fwd / np.sqrt(np.dot(fwd, fwd))).astype(f32)
    left = np.cross(up, fwd).astype(f32)
    y = np.cross(left, fwd).astype(f32)
    scale = f32(1.0) / f32(math.tan(0.5 * fov_y))
    z = (f32(-0.5) * f32(width) * left + f32(-0.5) * f32(height) * y + f32(0.5) * f32(height) * fwd * scale)
    p = capi.ViewingPlane()
    p.x[:], p.y[:], p.z[:], p.point[:] = (list(map(float, v)) for v in (left, y, z.astype(f32), eye))
    return p


def plane_array(p: capi.ViewingPlane) -> np.ndarray:
    return np.array(list(p.x) + list(p.y) + list(p.z) + list(p.point), np.float32)


class Bvh:
    """A BVH of one style's segments (flat linear curves of per-vertex radius)."""

    def __init__(self, xyz: torch.Tensor, radius: torch.Tensor, indices: torch.Tensor, stream=None):
        h = C.c_void_p()
        call("vkhr_bvh_build", ptr(xyz), ptr(radius), xyz.shape[0], ptr(indices), indices.numel(),
             capi.stream_of(stream), C.byref(h))
        self.handle = h
        self.n_indices = indices.numel()

    def info(self):
        n = C.c_uint32()
        b = (C.c_float * 6)()
        call("vkhr_bvh_info", self.handle, C.byref(n), b)
        return int(n.value), np.array(b[:], np.float32)

    def nodes(self):
        """(nodes (n-1, 16) float32, parent (2n-1,) uint32, leaf segment (n,) uint32) on the host."""
        n, _ = self.info()
        inner = max(n - 1, 0)
        nodes = np.zeros((inner, 16), np.float32)
        parent = np.zeros(inner + n, np.uint32)
        leaf = np.zeros(n, np.uint32)
        call("vkhr_bvh_nodes", self.handle, nodes.ctypes.data_as(C.c_void_p), parent.ctypes.data_as(C.c_void_p),
             leaf.ctypes.data_as(C.c_void_p), capi.stream_of())
        torch.cuda.synchronize()
        return nodes, parent, leaf

    def trace_rays(self, rays: torch.Tensor, any_hit: bool = False):
        """rays (n, 8) float32 on the device -> (t, u, segment) device tensors."""
        n = rays.shape[0]
        dev = rays.device
        t = torch.empty(n, dtype=torch.float32, device=dev)
        u = torch.empty(n, dtype=torch.float32, device=dev)
        seg = torch.empty(n, dtype=torch.int32, device=dev)
        call("vkhr_trace_rays", self.handle, ptr(rays), n, 1 if any_hit else 0, ptr(t), ptr(u), ptr(seg),
             capi.stream_of())
        return t, u, seg

    def raytrace(self, tangents, indices, plane: capi.ViewingPlane, light: capi.Light, width: int, height: int,
                 accum: torch.Tensor, method: int = 0, shadows_on: bool = True, ao_radius: float = 2.5,
                 exponent: float = 50.0, diffuse=(0.32, 0.228, 0.128), seed: int = 0, frame: int = 0,
                 ignore_self: bool = False):
        """Adds one sample per pixel (Raytracer::draw) to accum (H, W, 3) float64."""
        p = capi.TraceParams()
        p.method, p.shadows_on, p.ao_radius, p.exponent = method, 1 if shadows_on else 0, ao_radius, exponent
        p.diffuse[:] = list(diffuse)
        p.seed, p.frame, p.ignore_self = seed, frame, 1 if ignore_self else 0
        call("vkhr_raytrace", self.handle, ptr(tangents), ptr(indices), C.byref(plane), C.byref(light), C.byref(p),
             width, height, ptr(accum), capi.stream_of())

    def __del__(self):
        if getattr(self, "handle", None) and capi.lib is not None:
            capi.lib.vkhr_bvh_destroy(self.handle)
            self.handle = None
