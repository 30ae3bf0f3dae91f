"""GPU raymarch parity against the float64 transcription of the shaders:
max abs error <= 1e-4 per channel on the clamped fragment colour (tolerance
in tests/parity.py), NDC depth within 1e-6, identical fragment masks and
isosurface decisions except at documented near-ties."""
import math

import numpy as np
import pytest
import torch

from tests.parity import compare_raymarch
from tests.test_oracle_raymarch import axis_scene
from vkhr_amd import capi, params, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu


def _np_vol(vol):
    return dict(density=vol["density"].cpu().numpy(), tangents=vol["tangents"].cpu().numpy())


def _render(vol, res, cam, light, sp, rp, w, h, depth_buffer=None):
    out = D.raymarch(vol, res, cam, light, sp, rp, w, h, depth_buffer=depth_buffer, want_depth=True,
                     want_info=True)
    return {k: (v.cpu().numpy() if v is not None else None) for k, v in out.items()}


def _upload_volume(dens, tan, res, device):
    """A host volume (u8 density + i8x4 tangents) in the device layouts."""
    nx, ny, nz = res
    counts = D.to_dev(dens.astype(np.uint32), device)
    vol = D.finalize(counts, res, normalize=False, want_cells=True)
    vol["tangents"] = D.to_dev(tan, device)
    return vol


def test_axis_scene_closed_form(device):
    n, w, h = 16, 65, 65
    sc, bounds, dens, tan = axis_scene(n, w, h, c=10)
    vol = _upload_volume(dens, tan, (n, n, n), device)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(bounds, (n, n, n))
    rp = params.render(raycast_steps=128.0)
    g = _render(vol, (n, n, n), cam, light, sp, rp, w, h)
    rep = compare_raymarch(g, dict(density=dens, tangents=tan), (n, n, n), cam, light, sp, rp, w, h)
    assert rep["fragments"] > 100


@pytest.fixture(scope="module")
def c0_volume(device):
    s = synth.strands_10k()
    res = (128, 128, 128)
    b, vol = D.style_volume(s, res, device, segments=False)
    return s, res, b, vol


@pytest.mark.parametrize("azimuth", [0.0, 90.0, 215.0])
def test_10k_strands_512(device, c0_volume, azimuth):
    s, res, b, vol = c0_volume
    w = h = 512
    sc = synth.make_scene(b, w, h, azimuth_deg=azimuth)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(raycast_steps=256.0)
    g = _render(vol, res, cam, light, sp, rp, w, h)
    rep = compare_raymarch(g, _np_vol(vol), res, cam, light, sp, rp, w, h)
    print("10k strands", azimuth, rep)
    assert rep["fragments"] > 20000


@pytest.mark.parametrize("overrides", [dict(shading_model=2), dict(shading_model=3), dict(adsm_on=0),
                                       dict(raycast_steps=100.0), dict(isosurface=0.5), dict(renderer=3)])
def test_10k_strands_parameter_variants(device, c0_volume, overrides):
    s, res, b, vol = c0_volume
    w, h = 160, 120
    sc = synth.make_scene(b, w, h, azimuth_deg=40.0)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(**dict(dict(raycast_steps=256.0), **overrides))
    g = _render(vol, res, cam, light, sp, rp, w, h)
    compare_raymarch(g, _np_vol(vol), res, cam, light, sp, rp, w, h)


@pytest.mark.parametrize("res", [(96, 64, 80), (37, 130, 21)])
def test_uneven_volume_resolution(device, res):
    """Per-axis resolutions (texel scaling, brick counts, cell strides and the
    support bounds each taken per axis): parity with the oracle, and the
    occupancy skip bit-identical to the plain march."""
    s = synth.strands_10k()
    b, vol = D.style_volume(s, res, device, segments=True)
    w, h = 160, 120
    for az in (30.0, 250.0):
        sc = synth.make_scene(b, w, h, azimuth_deg=az)
        cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
        for steps in (256.0, 100.0):
            rp = params.render(raycast_steps=steps)
            g = _render(vol, res, cam, light, sp, rp, w, h)
            rep = compare_raymarch(g, _np_vol(vol), res, cam, light, sp, rp, w, h)
            assert rep["fragments"] > 1000, (res, az, steps, rep)
            plain = _render(dict(vol, occupancy=None), res, cam, light, sp, rp, w, h)
            for k in ("rgba", "depth", "info"):
                assert np.array_equal(g[k], plain[k]), (res, az, steps, k)


def test_depth_buffer_breaks_the_march(device, c0_volume):
    s, res, b, vol = c0_volume
    w, h = 128, 128
    sc = synth.make_scene(b, w, h)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(raycast_steps=256.0)
    free = _render(vol, res, cam, light, sp, rp, w, h)
    d = free["depth"].copy()
    # occluders a little behind the surface on the left half, in front on the right
    db = np.where(np.arange(w)[None, :] < w // 2, d + 2e-4, d - 2e-4).astype(np.float32)
    db[free["info"][..., 0] != 1] = 1.0
    g = _render(vol, res, cam, light, sp, rp, w, h, depth_buffer=D.to_dev(db.reshape(-1), device))
    rep = compare_raymarch(g, _np_vol(vol), res, cam, light, sp, rp, w, h, depth_buffer=db.reshape(-1))
    assert (g["info"][..., 0] == 1).sum() < (free["info"][..., 0] == 1).sum()
    print("depth buffer", rep)
    # the jumps over empty brick regions stop at the first sample of the
    # depth test: bit-identical to the march without them
    plain = _render(dict(vol, occupancy=None), res, cam, light, sp, rp, w, h,
                    depth_buffer=D.to_dev(db.reshape(-1), device))
    for k in ("rgba", "depth", "info"):
        assert np.array_equal(g[k], plain[k]), k


def test_136k_strands_full_scale(device):
    """1920x1080, 512 steps, 256^3 from 1.6M segments: the GPU renders the
    whole frame and the oracle checks every pixel of it."""
    s = synth.strands_136k()
    res = (256, 256, 256)
    b, vol = D.style_volume(s, res, device, segments=True)
    w, h = 1920, 1080
    sc = synth.make_scene(b, w, h)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(raycast_steps=512.0)
    g = _render(vol, res, cam, light, sp, rp, w, h)
    covered = np.nonzero(g["info"].reshape(-1, 4)[:, 0] == 1)[0]
    rep = compare_raymarch(g, _np_vol(vol), res, cam, light, sp, rp, w, h)
    print("136k strands", rep, "covered", len(covered))
    assert len(covered) > 300000


def test_float_density_volume(device, c0_volume):
    """The float32 density path (used by the differentiable form)."""
    s, res, b, vol = c0_volume
    f32 = vol["density"].to(torch.float32) / 255.0
    fvol = dict(density_f32=f32, tangents=vol["tangents"])
    w, h = 128, 128
    sc = synth.make_scene(b, w, h, azimuth_deg=120.0)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(raycast_steps=256.0)
    g = _render(fvol, res, cam, light, sp, rp, w, h)
    compare_raymarch(g, dict(tangents=vol["tangents"].cpu().numpy()), res, cam, light, sp, rp, w, h,
                     density_f32=f32.cpu().numpy())


def test_invalid_arguments(device, c0_volume):
    s, res, b, vol = c0_volume
    sc = synth.make_scene(b, 8, 8)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    with pytest.raises(capi.VkhrError):
        D.raymarch(vol, res, cam, light, sp, params.render(raycast_steps=0.0), 8, 8)
    with pytest.raises(capi.VkhrError):
        D.raymarch(dict(tangents=vol["tangents"]), res, cam, light, sp, params.render(), 8, 8)


def test_occupancy_skip_is_bit_exact(device):
    """Samples in empty bricks skip their gathers: the frame, depth and
    surface decisions are bit-identical to the march without the occupancy
    (1920x1080, 512 steps, 256^3 from 1.6M segments)."""
    s = synth.strands_136k()
    res = (256, 256, 256)
    b, vol = D.style_volume(s, res, device, segments=True)
    assert vol.get("occupancy") is not None
    w, h = 1920, 1080
    for az in (0.0, 130.0):
        sc = synth.make_scene(b, w, h, azimuth_deg=az)
        cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
        rp = params.render(raycast_steps=512.0)
        a = D.raymarch(vol, res, cam, light, sp, rp, w, h, want_depth=True, want_info=True)
        plain = dict(vol, occupancy=None)
        o = D.raymarch(plain, res, cam, light, sp, rp, w, h, want_depth=True, want_info=True)
        for k in ("rgba", "depth", "info"):
            assert torch.equal(a[k], o[k]), (az, k)
        assert int((a["info"].reshape(-1, 4)[:, 0] == 1).sum()) > 100000
