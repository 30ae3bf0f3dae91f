"""Batched adjoint (the C ABI's vkhr_adjoint_batch_*): one training step of
the differentiable raymarch over several styles x views, and the per-render
sequence it replaces, over torch device buffers."""
from __future__ import annotations

import ctypes as C

import torch

from . import capi, params, synth
from . import device as D
from .capi import call, ptr


class AdjointProblem:
    """Styles of the batch generator voxelized (segment DDA, normalized) into
    n^3 float densities, each seen from `views` cameras on a ring; the
    targets are renders of the densities with +10% Gaussian noise (clamped
    at 0), as in the single-style differentiable configuration."""

    def __init__(self, device, n_styles=8, views=8, n=128, width=512, height=512, steps=256.0, seed=0,
                 diff=(0.3, 11.0, 80.0)):
        self.width, self.height, self.res = width, height, (n, n, n)
        styles, _ = synth.batch_styles(n_styles)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.styles, self.renders = [], []
        self.diff = D.diff_params(*diff)
        for si, st in enumerate(styles):
            b, vol = D.style_volume(st, self.res, device, segments=True, want_f32=True)
            dens = vol["density_f32"]
            noise = torch.randn(dens.shape, generator=g).to(device)
            noisy = torch.clamp(dens * (1.0 + 0.1 * noise), min=0.0)
            self.styles.append(dict(bounds=b, density=dens, noisy=noisy, tangents=vol["tangents"],
                                    d_density=torch.zeros_like(dens),
                                    d_params=torch.zeros(3, dtype=torch.float32, device=device)))
            for v in range(views):
                sc = synth.make_scene(b, width, height, azimuth_deg=360.0 * v / views)
                cam, light = params.camera(sc), params.light(sc)
                sp, rp = params.strand(b, self.res), params.render(raycast_steps=steps)
                target, _ = D.raymarch_record(dict(density_f32=noisy, tangents=vol["tangents"]), self.res, cam,
                                              light, sp, rp, self.diff, width, height)
                self.renders.append(dict(style=si, cam=cam, light=light, sp=sp, rp=rp, target=target,
                                         rgba=torch.empty_like(target),
                                         record=torch.empty(int(capi.lib.vkhr_march_record_bytes(width, height)),
                                                            dtype=torch.uint8, device=device),
                                         loss=torch.zeros(1, dtype=torch.float32, device=device)))

    def vol(self, s):
        st = self.styles[s]
        return dict(density_f32=st["density"], tangents=st["tangents"])

    def run_plain(self):
        """The per-render sequence: record forward, L2 loss gradient, adjoint
        accumulated into the style's gradient (zeroed first)."""
        for st in self.styles:
            st["d_density"].zero_()
            st["d_params"].zero_()
        for r in self.renders:
            st = self.styles[r["style"]]
            rgba, rec = D.raymarch_record(self.vol(r["style"]), self.res, r["cam"], r["light"], r["sp"], r["rp"],
                                          self.diff, self.width, self.height, record=r["record"], out=r["rgba"])
            d_rgba = 2.0 * (rgba - r["target"])
            D.raymarch_adjoint(self.vol(r["style"]), self.res, r["cam"], r["light"], r["sp"], r["rp"], self.diff,
                               self.width, self.height, rec, d_rgba, st["d_density"], st["d_params"])


class AdjointBatch:
    """The whole step as one graph: forward, loss gradient, adjoint and the
    per-style gradient fold (gradients overwritten at each launch)."""

    def __init__(self, prob: AdjointProblem):
        self.prob = prob
        styles = (capi.AdjointStyle * len(prob.styles))()
        for i, st in enumerate(prob.styles):
            styles[i].volume = D.volume_struct(prob.vol(i), prob.res)
            styles[i].d_density, styles[i].d_params = ptr(st["d_density"]), ptr(st["d_params"])
        renders = (capi.AdjointRender * len(prob.renders))()
        for i, r in enumerate(prob.renders):
            c = renders[i]
            c.style = r["style"]
            c.camera, c.light, c.strand, c.params, c.diff = r["cam"], r["light"], r["sp"], r["rp"], prob.diff
            c.target, c.rgba, c.record, c.loss = ptr(r["target"]), ptr(r["rgba"]), ptr(r["record"]), ptr(r["loss"])
        h = C.c_void_p()
        call("vkhr_adjoint_batch_create", styles, len(prob.styles), renders, len(prob.renders), prob.width,
             prob.height, C.byref(h))
        self.handle = h
        self.stream = torch.cuda.Stream()

    def launch(self, wait=True):
        if wait:
            self.stream.wait_stream(torch.cuda.current_stream())
        call("vkhr_adjoint_batch_launch", self.handle, C.c_void_p(self.stream.cuda_stream))
        if wait:
            torch.cuda.current_stream().wait_stream(self.stream)

    def __del__(self):
        if getattr(self, "handle", None):
            capi.lib.vkhr_adjoint_batch_destroy(self.handle)
            self.handle = None
