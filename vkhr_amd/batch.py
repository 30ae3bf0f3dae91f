"""Batched launch of many hair styles (the C ABI's vkhr_batch_*), plus the
per-style sequence it replaces, over torch device buffers."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import torch

from . import capi, params, synth
from . import device as D
from .capi import call, ptr

TUNING_PATH = Path(__file__).resolve().parent / "tuning.json"


def load_tuning(key: str = "batch") -> capi.Tuning | None:
    """The autotuner's choice for a workload key (vkhr_amd/tuning.json), or None."""
    if not TUNING_PATH.exists():
        return None
    table = json.loads(TUNING_PATH.read_text())
    if key not in table:
        return None
    t = capi.Tuning()
    t.tile_w, t.tile_h, t.blocks_per_cu = (table[key][k] for k in ("tile_w", "tile_h", "blocks_per_cu"))
    t.groups = int(table[key].get("groups", 1))
    return t


class StyleBuffers:
    """Device inputs and outputs of one style of a batch."""

    def __init__(self, style: synth.Style, n: int, width: int, height: int, device, azimuth: float = 0.0,
                 steps: float = 512.0, with_counts: bool = True, morton: bool = True):
        """with_counts=False writes no separate uint32 count volume.
        morton: the index pairs strand by strand with the strands in the
        Morton order of their roots (vkhr_sort_strands, once at load), the
        order the binned volume build runs fastest on; the volumes do not
        depend on the order."""
        self.n = n
        self.xyz = D.to_dev(style.vertices, device)
        off = D.to_dev(style.strand_offsets, device)
        self.tan = D.generate_tangents(self.xyz, off)
        self.bounds = D.aabb(self.xyz)
        self.idx = D.generate_indices(off)
        if morton and self.idx.numel() >= 2:
            self.idx = D.sort_strands(self.xyz, self.idx, self.bounds)
        nv = n ** 3
        self.counts = torch.empty(nv, dtype=torch.int32, device=device) if with_counts else None
        self.acc = torch.empty(2 * nv, dtype=torch.int64, device=device)
        self.density = torch.empty(nv, dtype=torch.uint8, device=device)
        self.tangents = torch.empty(4 * nv, dtype=torch.int8, device=device)
        self.cells = torch.empty((n + 1) ** 3, dtype=torch.int64, device=device)
        self.occupancy = torch.empty(D.occupancy_words((n, n, n)), dtype=torch.int32, device=device)
        self.rgba = torch.empty((height, width, 4), dtype=torch.float32, device=device)
        sc = synth.make_scene(self.bounds, width, height, azimuth_deg=azimuth)
        self.cam, self.light = params.camera(sc), params.light(sc)
        self.sp = params.strand(self.bounds, (n, n, n))
        self.rp = params.render(raycast_steps=steps)
        self.width, self.height = width, height

    def c_style(self) -> capi.BatchStyle:
        b = capi.BatchStyle()
        b.xyz, b.n_vertices = ptr(self.xyz), self.xyz.shape[0]
        b.indices, b.n_indices, b.tangents = ptr(self.idx), self.idx.shape[0], ptr(self.tan)
        b.bounds = capi.make_aabb(self.bounds)
        b.volume = self.n
        b.counts, b.tangent_acc, b.density = ptr(self.counts), ptr(self.acc), ptr(self.density)
        b.tangent_i8x4, b.cells, b.rgba = ptr(self.tangents), ptr(self.cells), ptr(self.rgba)
        b.occupancy = ptr(self.occupancy)
        b.camera, b.light, b.strand, b.params = self.cam, self.light, self.sp, self.rp
        return b

    def run_plain(self, out: dict | None = None, one_call: bool = False):
        """The per-style sequence: voxelize_segments -> finalize -> raymarch
        (one_call: the one-call volume build -> raymarch), into separate
        output buffers (out) or this style's own."""
        o = out if out is not None else dict(counts=self.counts, acc=self.acc, density=self.density,
                                             tangents=self.tangents, cells=self.cells, occupancy=self.occupancy,
                                             rgba=self.rgba)
        res = (self.n,) * 3
        if one_call:
            vol = dict(counts=o["counts"], density=o["density"], tangents=o["tangents"], density_f32=None,
                       cells=o["cells"], occupancy=o.get("occupancy"))
            D.build_volume(self.xyz, self.bounds, res, self.idx, self.tan, normalize=True, workspace=o["acc"],
                           out=vol)
        else:
            sums = o["acc"] is not None   # else counts only (no tangent volume)
            D.voxelize(self.xyz, self.bounds, res, indices=self.idx, tangents=self.tan if sums else None,
                       want_sums=sums, counts=o["counts"], sums=o["acc"], want_counts=o["counts"] is not None)
            call("vkhr_volume_finalize", ptr(o["counts"]), ptr(o["acc"]), self.n, self.n, self.n, 1,
                 ptr(o["density"]), ptr(o["tangents"]), None, ptr(o["cells"]), ptr(o.get("occupancy")),
                 capi.stream_of())
        vol = dict(density=o["density"], cells=o["cells"], tangents=o["tangents"], occupancy=o.get("occupancy"))
        D.raymarch(vol, res, self.cam, self.light, self.sp, self.rp, self.width, self.height, want_depth=False,
                   out=o["rgba"])


class Batch:
    def __init__(self, buffers: list[StyleBuffers], width: int, height: int, tuning: capi.Tuning | None = None):
        self.buffers = buffers
        arr = (capi.BatchStyle * len(buffers))(*[b.c_style() for b in buffers])
        h = C.c_void_p()
        call("vkhr_batch_create", arr, len(buffers), width, height, C.byref(tuning) if tuning else None, C.byref(h))
        self.handle = h
        self.stream = torch.cuda.Stream()

    def launch(self, wait=True):
        """Replays the batch graph on its own stream (graph capture needs a
        created stream); with wait, the current stream is ordered around it."""
        if wait:
            self.stream.wait_stream(torch.cuda.current_stream())
        call("vkhr_batch_launch", self.handle, C.c_void_p(self.stream.cuda_stream))
        if wait:
            torch.cuda.current_stream().wait_stream(self.stream)

    def __del__(self):
        if getattr(self, "handle", None):
            capi.lib.vkhr_batch_destroy(self.handle)
            self.handle = None
