"""Library-level behaviour on the GPU: volume sizes that would wrap 32-bit
size arithmetic are refused on the host before any kernel is enqueued, and
the per-stream scratch areas are released on request."""
import ctypes as C

import numpy as np
import pytest
import torch

from vkhr_amd import capi, params, synth
from vkhr_amd import device as D
from vkhr_amd.capi import VkhrError, call, ptr, stream_of

pytestmark = pytest.mark.gpu

HUGE = 0xFFFFFFFF


def _expect(code, name, *args):
    with pytest.raises(VkhrError) as e:
        call(name, *args)
    assert e.value.code == code, str(e.value)
    return str(e.value)


def test_oversized_volume_sides_are_refused_before_any_launch(device):
    buf = torch.zeros(1 << 16, dtype=torch.int64, device=device)
    p = ptr(buf)
    s = stream_of()
    # finalize: a side of 2^32 - 1 (its + 1 wrapped to 0 in 32 bits)
    msg = _expect(-3, "vkhr_volume_finalize", p, p, HUGE, 8, 8, 1, p, p, p, p, p, s)
    assert "side above" in msg
    _expect(-3, "vkhr_volume_finalize", p, p, 8, HUGE, 8, 1, p, p, None, None, None, s)
    # sides whose 64-bit product wraps to 0: 2^22 * 2^22 * 2^21 = 2^65
    b = capi.make_aabb(dict(origin=[0, 0, 0], size=[1, 1, 1], radius=1.0, volume=1.0))
    _expect(-3, "vkhr_voxelize_segments", p, 2, p, 2, p, C.byref(b), 1 << 22, 1 << 22, 1 << 21, p, None, s)
    _expect(-3, "vkhr_voxelize_vertices", p, 1, p, C.byref(b), HUGE, 1, 1, p, None, s)
    # the renderers' volumes
    bounds = dict(origin=np.zeros(3, np.float32), size=np.ones(3, np.float32), radius=np.float32(1.7),
                  volume=np.float32(1.0))
    sc = synth.make_scene(bounds, 16, 16)
    cam, light = params.camera(sc), params.light(sc)
    sp, rp = params.strand(bounds, (8, 8, 8)), params.render(raycast_steps=16.0)
    v = capi.Volume()
    v.cells = v.density = v.tangents = v.occupancy = p
    v.nx, v.ny, v.nz = HUGE, 8, 8
    _expect(-3, "vkhr_raymarch", C.byref(v), C.byref(cam), C.byref(light), C.byref(sp), C.byref(rp), 16, 16,
            None, p, None, None, s)
    v.nx, v.ny, v.nz = 8, 1 << 13, 1 << 13   # (ny+1)(nz+1) >= 2^24: 24-bit cell index multiplies
    _expect(-3, "vkhr_raymarch", C.byref(v), C.byref(cam), C.byref(light), C.byref(sp), C.byref(rp), 16, 16,
            None, p, None, None, s)
    fv = capi.Volume()
    fv.density_f32 = fv.tangents = p
    fv.nx, fv.ny, fv.nz = HUGE, HUGE, 2
    d = D.diff_params(0.1)
    _expect(-3, "vkhr_raymarch_record", C.byref(fv), C.byref(cam), C.byref(light), C.byref(sp), C.byref(rp),
            C.byref(d), 16, 16, None, p, None, p, s)
    _expect(-3, "vkhr_raymarch_adjoint", C.byref(fv), C.byref(cam), C.byref(light), C.byref(sp), C.byref(rp),
            C.byref(d), 16, 16, p, p, p, p, s)
    _expect(-3, "vkhr_volume_downsample", p, HUGE, 2, 2, 0, p, s)
    _expect(-3, "vkhr_volume_filter", p, 2, HUGE, 2, 3, ptr(buf[1:]), s)
    torch.cuda.synchronize()
    # nothing was written
    assert int(buf.abs().sum().item()) == 0


def _finalized(counts, res, stream):
    with torch.cuda.stream(stream):
        return D.finalize(counts, res, normalize=True, want_cells=True)


def test_scratch_areas_are_released_per_stream(device):
    rng = np.random.default_rng(5)
    res = (40, 24, 32)
    counts = D.to_dev(rng.integers(0, 400, size=res[0] * res[1] * res[2], dtype=np.uint32), device)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    first = [_finalized(counts, res, s) for s in streams]
    torch.cuda.synchronize()
    for s in streams:
        assert capi.lib.vkhr_scratch_bytes(stream_of(s)) >= 1 << 20
    call("vkhr_release_scratch", stream_of(streams[0]))
    assert capi.lib.vkhr_scratch_bytes(stream_of(streams[0])) == 0
    assert capi.lib.vkhr_scratch_bytes(stream_of(streams[1])) >= 1 << 20
    call("vkhr_release_scratch", stream_of(streams[1]))
    assert capi.lib.vkhr_scratch_bytes(stream_of(streams[1])) == 0
    call("vkhr_release_scratch", stream_of(streams[1]))   # releasing twice is a no-op
    # a released stream gets a new area on its next call; the results are unchanged
    again = [_finalized(counts, res, s) for s in streams]
    torch.cuda.synchronize()
    for a, b in zip(first, again):
        for k in ("density", "cells", "occupancy"):
            assert torch.equal(a[k], b[k]), k
    for s in streams:
        call("vkhr_release_scratch", stream_of(s))


def test_scratch_growth_frees_the_outgrown_area(device):
    """The adjoint needs a larger area than the finalize on the same stream:
    the area grows (the old buffer is freed after the stream drains) and the
    earlier call's results are not disturbed."""
    rng = np.random.default_rng(6)
    res = (16, 16, 16)
    counts = D.to_dev(rng.integers(0, 300, size=16 ** 3, dtype=np.uint32), device)
    s = torch.cuda.Stream()
    vol = _finalized(counts, res, s)
    small = capi.lib.vkhr_scratch_bytes(stream_of(s))
    with torch.cuda.stream(s):
        # the adjoint's cell accumulator of a 64^3 volume (65^3 x 32 B = 8.8 MB)
        # is larger than the first area (1 MiB)
        res64 = (64, 64, 64)
        f32 = torch.rand(64 ** 3, device=device)
        tang = torch.zeros(4 * 64 ** 3, dtype=torch.int8, device=device)
        bounds = dict(origin=np.zeros(3, np.float32), size=np.ones(3, np.float32), radius=np.float32(1.7),
                      volume=np.float32(1.0))
        sc = synth.make_scene(bounds, 32, 32)
        cam, light = params.camera(sc), params.light(sc)
        sp, rp = params.strand(bounds, res64), params.render(raycast_steps=32.0, shading_model=1)
        fv = dict(density_f32=f32, tangents=tang)
        rgba, rec = D.raymarch_record(fv, res64, cam, light, sp, rp, D.diff_params(0.2), 32, 32)
        D.raymarch_adjoint(fv, res64, cam, light, sp, rp, D.diff_params(0.2), 32, 32, rec, torch.ones_like(rgba))
    torch.cuda.synchronize()
    assert capi.lib.vkhr_scratch_bytes(stream_of(s)) >= 65 ** 3 * 32 > small
    again = _finalized(counts, res, s)
    torch.cuda.synchronize()
    assert torch.equal(vol["cells"], again["cells"])
    call("vkhr_release_scratch", stream_of(s))
    assert capi.lib.vkhr_scratch_bytes(stream_of(s)) == 0
