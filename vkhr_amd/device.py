"""Torch-tensor conveniences over the C ABI (device memory and streams only:
every computation is a libvkhr_hip entry point).  Used by tests and bench."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import call, ptr, stream_of


def to_dev(a: np.ndarray, device="cuda") -> torch.Tensor:
    """Copies a numpy array to the device; unsigned 32/64-bit data travel as
    the same-width signed dtype (the kernels read the bits)."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).to(device)


def aabb(xyz: torch.Tensor) -> dict:
    out = capi.AABB()
    call("vkhr_aabb_compute", ptr(xyz), xyz.shape[0], C.byref(out), stream_of())
    return capi.aabb_dict(out)


def generate_indices(offsets: torch.Tensor) -> torch.Tensor:
    n = offsets.shape[0] - 1
    n_idx = 2 * (int(offsets[-1].item()) - n)
    out = torch.empty(max(n_idx, 1), dtype=torch.int32, device=offsets.device)
    call("vkhr_generate_indices", ptr(offsets), n, ptr(out), stream_of())
    return out[:n_idx]


def sort_segments(xyz: torch.Tensor, indices: torch.Tensor, bounds: dict) -> torch.Tensor:
    """The index pairs in the Morton order of the segment midpoints (a
    permutation of the pairs, made once at load)."""
    out = torch.empty_like(indices)
    b = capi.make_aabb(bounds)
    call("vkhr_sort_segments", ptr(xyz), xyz.shape[0], ptr(indices), indices.shape[0], C.byref(b), ptr(out),
         stream_of())
    return out


def sort_strands(xyz: torch.Tensor, indices: torch.Tensor, bounds: dict) -> torch.Tensor:
    """The index pairs strand by strand, the strands in the Morton order of
    their root vertex (a permutation of the pairs, made once at load)."""
    out = torch.empty_like(indices)
    b = capi.make_aabb(bounds)
    call("vkhr_sort_strands", ptr(xyz), xyz.shape[0], ptr(indices), indices.shape[0], C.byref(b), ptr(out),
         stream_of())
    return out


def generate_tangents(xyz: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
    out = torch.zeros_like(xyz)
    call("vkhr_generate_tangents", ptr(xyz), ptr(offsets), offsets.shape[0] - 1, ptr(out), stream_of())
    return out


def voxelize(xyz: torch.Tensor, bounds: dict, res, indices: torch.Tensor | None = None,
             tangents: torch.Tensor | None = None, want_sums: bool = False, counts=None, sums=None,
             want_counts: bool = True):
    """Vertex mode when indices is None, else segment mode.  Returns (counts, sums);
    with want_sums and not want_counts the counts stay packed in sums (counts is None)."""
    nx, ny, nz = res
    n = nx * ny * nz
    dev = xyz.device
    if not want_sums:
        want_counts = True
    if counts is None and want_counts:
        counts = torch.empty(n, dtype=torch.int32, device=dev)
    if want_sums and sums is None:
        sums = torch.empty(2 * n, dtype=torch.int64, device=dev)   # fixed-point tangent accumulator
    b = capi.make_aabb(bounds)
    if indices is None:
        call("vkhr_voxelize_vertices", ptr(xyz), xyz.shape[0], ptr(tangents), C.byref(b), nx, ny, nz,
             ptr(counts), ptr(sums) if want_sums else None, stream_of())
    else:
        call("vkhr_voxelize_segments", ptr(xyz), xyz.shape[0], ptr(indices), indices.shape[0], ptr(tangents),
             C.byref(b), nx, ny, nz, ptr(counts), ptr(sums) if want_sums else None, stream_of())
    return counts, (sums if want_sums else None)


def occupancy_words(res) -> int:
    return int(capi.lib.vkhr_occupancy_words(*res))


def finalize(counts: torch.Tensor | None, res, sums: torch.Tensor | None = None, normalize: bool = True,
             want_f32: bool = False, want_cells: bool = True, want_occupancy: bool = True, noalloc: bool = False):
    """counts may be None when sums holds them packed.  The brick occupancy
    comes with the cells unless want_occupancy is False.  noalloc: through
    vkhr_volume_finalize_noalloc (static scratch, no library allocation)."""
    nx, ny, nz = res
    n = nx * ny * nz
    dev = (counts if counts is not None else sums).device
    dens = torch.empty(n, dtype=torch.uint8, device=dev)
    tan = torch.empty(4 * n, dtype=torch.int8, device=dev) if sums is not None else None
    f32 = torch.empty(n, dtype=torch.float32, device=dev) if want_f32 else None
    cells = torch.empty((nx + 1) * (ny + 1) * (nz + 1), dtype=torch.int64, device=dev) if want_cells else None
    occ = (torch.empty(occupancy_words(res), dtype=torch.int32, device=dev)
           if want_cells and want_occupancy else None)
    call("vkhr_volume_finalize_noalloc" if noalloc else "vkhr_volume_finalize", ptr(counts), ptr(sums), nx, ny, nz,
         int(normalize), ptr(dens), ptr(tan),
         ptr(f32), ptr(cells), ptr(occ), stream_of())
    return dict(density=dens, tangents=tan, density_f32=f32, cells=cells, occupancy=occ)


def build_volume(xyz: torch.Tensor, bounds, res, indices: torch.Tensor, tangents: torch.Tensor | None = None,
                 normalize: bool = True, want_counts: bool = False, want_tangents: bool = True,
                 want_f32: bool = False, want_cells: bool = True, want_occupancy: bool = True,
                 workspace: torch.Tensor | None = None, out: dict | None = None):
    """vkhr_volume_build_segments: the voxelization and the finalize in one
    call (LDS-binned).  Returns the dict of finalize plus counts (or None);
    out: a previous result whose buffers are reused."""
    nx, ny, nz = res
    n = nx * ny * nz
    dev = xyz.device
    if out is None:
        out = dict(
            counts=torch.empty(n, dtype=torch.int32, device=dev) if want_counts else None,
            density=torch.empty(n, dtype=torch.uint8, device=dev),
            tangents=torch.empty(4 * n, dtype=torch.int8, device=dev) if want_tangents else None,
            density_f32=torch.empty(n, dtype=torch.float32, device=dev) if want_f32 else None,
            cells=torch.empty((nx + 1) * (ny + 1) * (nz + 1), dtype=torch.int64, device=dev) if want_cells else None,
            occupancy=(torch.empty(occupancy_words(res), dtype=torch.int32, device=dev)
                       if want_cells and want_occupancy else None))
    if workspace is None:
        workspace = torch.empty(2 * n, dtype=torch.int64, device=dev)
    b = capi.make_aabb(bounds)
    call("vkhr_volume_build_segments", ptr(xyz), xyz.shape[0], ptr(indices), indices.shape[0],
         ptr(tangents) if out["tangents"] is not None else None, C.byref(b), nx, ny, nz, int(normalize),
         ptr(out["counts"]), ptr(out["density"]), ptr(out["tangents"]), ptr(out["density_f32"]), ptr(out["cells"]),
         ptr(out["occupancy"]), ptr(workspace), stream_of())
    return out


def volume_struct(vol: dict, res) -> capi.Volume:
    v = capi.Volume()
    v.density = ptr(vol.get("density"))
    v.cells = ptr(vol.get("cells"))
    v.tangents = ptr(vol.get("tangents"))
    v.density_f32 = ptr(vol.get("density_f32"))
    v.occupancy = ptr(vol.get("occupancy"))
    v.nx, v.ny, v.nz = res
    return v


def raymarch(vol: dict, res, cam, light, strand, render, width: int, height: int, depth_buffer=None,
             want_depth=True, want_info=False, out=None):
    """Returns dict(rgba (H,W,4) float32, depth (H,W) or None, info (H,W,4) int32 or None)."""
    dev = next(t for t in vol.values() if t is not None).device
    rgba = out if out is not None else torch.empty((height, width, 4), dtype=torch.float32, device=dev)
    depth = torch.empty((height, width), dtype=torch.float32, device=dev) if want_depth else None
    info = torch.empty((height, width, 4), dtype=torch.int32, device=dev) if want_info else None
    v = volume_struct(vol, res)
    call("vkhr_raymarch", C.byref(v), C.byref(cam), C.byref(light), C.byref(strand), C.byref(render), width,
         height, ptr(depth_buffer), ptr(rgba), ptr(depth), ptr(info), stream_of())
    return dict(rgba=rgba, depth=depth, info=info)


def style_volume(style, res, device, segments=True, want_f32=False):
    """Voxelize + finalize a synthetic style on the device (the reference's
    load path).  Returns (bounds, volume dict)."""
    xyz = to_dev(style.vertices, device)
    off = to_dev(style.strand_offsets, device)
    tan = generate_tangents(xyz, off)
    b = aabb(xyz)
    idx = generate_indices(off) if segments else None
    counts, sums = voxelize(xyz, b, res, indices=idx, tangents=tan, want_sums=True)
    vol = finalize(counts, res, sums=sums, normalize=True, want_f32=want_f32, want_cells=True)
    return b, vol


class PPLL:
    """Device storage of a per-pixel linked list (heads, 12-byte nodes, counter)."""

    def __init__(self, width: int, height: int, node_count: int, device="cuda"):
        self.width, self.height, self.node_count = width, height, node_count
        self.heads = torch.empty(width * height, dtype=torch.int32, device=device)
        self.nodes = torch.empty(3 * node_count, dtype=torch.int32, device=device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=device)
        self.c = capi.Ppll()
        self.c.heads, self.c.nodes, self.c.counter = ptr(self.heads), ptr(self.nodes), ptr(self.counter)
        self.c.node_count, self.c.width, self.c.height = node_count, width, height

    def clear(self):
        call("vkhr_ppll_clear", C.byref(self.c), stream_of())

    def resolve(self, background=None, out=None):
        if out is None:
            out = torch.empty((self.height, self.width, 4), dtype=torch.float32, device=self.heads.device)
        call("vkhr_ppll_resolve", C.byref(self.c), ptr(background), ptr(out), stream_of())
        return out

    def used(self) -> int:
        return int(self.counter.item() & 0xFFFFFFFF)


def shadow_depth(xyz: torch.Tensor, indices: torch.Tensor, light_vp_cm, line_width: float, size=(1024, 1024),
                 out=None):
    m = (C.c_float * 16)(*[float(v) for v in light_vp_cm])
    if out is None:
        out = torch.empty((size[1], size[0]), dtype=torch.float32, device=xyz.device)
    call("vkhr_shadow_depth", ptr(xyz), xyz.shape[0], ptr(indices), indices.shape[0], m, float(line_width), size[0],
         size[1], ptr(out), stream_of())
    return out


def raster_strands(xyz, tangents, thickness, indices, cam, light, strand, render, vol: dict | None, res,
                   shadow_map, ppll: PPLL, depth_buffer=None, capture: dict | None = None):
    v = volume_struct(vol, res) if vol is not None else None
    cap = None
    if capture is not None:
        cap = capi.FragCapture()
        cap.segment, cap.pixel, cap.color = ptr(capture.get("segment")), ptr(capture.get("pixel")), \
            ptr(capture.get("color"))
    sw, sh = (shadow_map.shape[1], shadow_map.shape[0]) if shadow_map is not None else (0, 0)
    call("vkhr_raster_strands", ptr(xyz), ptr(tangents), ptr(thickness), xyz.shape[0], ptr(indices), indices.shape[0],
         C.byref(cam), C.byref(light), C.byref(strand), C.byref(render), C.byref(v) if v is not None else None,
         ptr(shadow_map), sw, sh, ptr(depth_buffer), C.byref(ppll.c), C.byref(cap) if cap is not None else None,
         stream_of())


def raymarch_ppll(vol: dict, res, cam, light, strand, render, ppll: PPLL, depth_buffer=None):
    v = volume_struct(vol, res)
    call("vkhr_raymarch_ppll", C.byref(v), C.byref(cam), C.byref(light), C.byref(strand), C.byref(render),
         ppll.width, ppll.height, ptr(depth_buffer), C.byref(ppll.c), stream_of())


def diff_params(absorption: float, adsm_thickness: float = 11.0, shininess: float = 80.0) -> capi.DiffParams:
    d = capi.DiffParams()
    d.absorption, d.adsm_thickness, d.shininess = absorption, adsm_thickness, shininess
    return d


def raymarch_record(vol: dict, res, cam, light, strand, render, diff, width, height, record=None, out=None):
    """Differentiable forward: returns (rgba (H,W,4), record bytes tensor)."""
    dev = next(t for t in vol.values() if t is not None).device
    if record is None:
        record = torch.empty(int(capi.lib.vkhr_march_record_bytes(width, height)), dtype=torch.uint8, device=dev)
    rgba = out if out is not None else torch.empty((height, width, 4), dtype=torch.float32, device=dev)
    v = volume_struct(vol, res)
    call("vkhr_raymarch_record", C.byref(v), C.byref(cam), C.byref(light), C.byref(strand), C.byref(render),
         C.byref(diff) if diff is not None else None, width, height, None, ptr(rgba), None, ptr(record), stream_of())
    return rgba, record


def raymarch_adjoint(vol: dict, res, cam, light, strand, render, diff, width, height, record, d_rgba,
                     d_density=None, d_params=None):
    dev = d_rgba.device
    if d_density is None:
        d_density = torch.zeros(res[0] * res[1] * res[2], dtype=torch.float32, device=dev)
    if d_params is None:
        d_params = torch.zeros(3, dtype=torch.float32, device=dev)
    v = volume_struct(vol, res)
    call("vkhr_raymarch_adjoint", C.byref(v), C.byref(cam), C.byref(light), C.byref(strand), C.byref(render),
         C.byref(diff) if diff is not None else None, width, height, ptr(record), ptr(d_rgba.contiguous()),
         ptr(d_density), ptr(d_params), stream_of())
    return d_density, d_params
