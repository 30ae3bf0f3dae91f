"""Pins the float64 rasterizer / resolve oracle (oracle/raster_oracle.cc):
  * an orthographic "camera" maps world x, y to window pixels, so the
    fragment sets of a horizontal and a 45-degree line follow by hand and
    by a numpy brute force of the declared line rule;
  * shadow-map depth along a line is the linear interpolation of its ends;
  * PCF of a constant depth map is the single-tap closed form;
  * the resolve of <= 16 fragments is the textbook back-to-front "over"
    fold, identical colours blend (1-a)^n, and at most 16 + 1024 nodes of a
    list are blended (resolve.comp:12-13).
"""
import math

import numpy as np
import pytest

from oracle import raster as OR
from vkhr_amd import capi, params, synth

W, H = 32, 24


def ortho_camera(width=W, height=H):
    """view = I, projection maps world (x, y) to window (x, y), z to depth."""
    c = capi.Camera()
    proj = np.array([[2.0 / width, 0, 0, -1], [0, 2.0 / height, 0, -1], [0, 0, 1, 0], [0, 0, 0, 1]])
    c.view[:] = [float(v) for v in np.eye(4).T.reshape(-1)]
    c.projection[:] = [float(v) for v in proj.T.reshape(-1)]
    c.position[:] = [width / 2.0, height / 2.0, -100.0]
    c.look_at_distance = 100.0
    c.znear, c.zfar = 0.1, 1000.0
    c.resolution[:] = [float(width), float(height)]
    return c


def ortho_light(scale=1.0 / 16.0):
    l = capi.Light()
    m = np.array([[scale, 0, 0, 0], [0, scale, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
    l.matrix[:] = [float(v) for v in m.T.reshape(-1)]
    l.origin[:] = [16.0, 100.0, 0.0]
    l.intensity[:] = [0.43, 0.38, 0.28, 0.0]
    return l


def one_segment(a, b, thick=0.042):
    xyz = np.array([a, b], np.float32)
    tan = np.tile(np.array([[1.0, 0.0, 0.0]], np.float32), (2, 1))
    th = np.array([thick, thick], np.float32)
    return xyz, tan, th, np.array([0, 1], np.uint32)


def strand_params(width=1.0, alpha=0.3):
    b = dict(origin=np.zeros(3, np.float32), size=np.full(3, 32.0, np.float32), radius=np.float32(55.4),
             volume=np.float32(32768.0))
    return params.strand(b, (8, 8, 8), strand_radius=width, hair_opacity=alpha)


def test_horizontal_line_fragments_by_hand():
    xyz, tan, th, idx = one_segment([2.0, 5.5, 0.5], [10.0, 5.5, 0.5])
    rp = params.render(renderer=0, adsm_on=0, shading_model=2)      # no LAO, no ADSM
    o = OR.raster(xyz, tan, th, idx, ortho_camera(), ortho_light(), strand_params(), rp)
    assert sorted(o["pixel"].tolist()) == [5 * W + x for x in range(2, 10)]
    assert np.allclose(o["depth"], 0.5)
    # gpaa distance 0 -> coverage = alpha * thickness / 0.042
    assert o["kept"].all()
    assert np.allclose(o["rgba"][:, 3], 0.3 * 0.042 / 0.042, atol=1e-12)


def _brute_force_rule(a, b, lw, width, height):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    v = b - a
    l2 = v @ v
    out = []
    for py in range(height):
        for px in range(width):
            c = np.array([px + 0.5, py + 0.5]) - a
            dp = c @ v
            cr = v[0] * c[1] - v[1] * c[0]
            if 0 <= dp < l2 and cr * cr <= (lw / 2) ** 2 * l2:
                out.append(py * width + px)
    return sorted(out)


@pytest.mark.parametrize("a,b,lw", [((3.2, 2.7), (17.9, 14.1), 1.0), ((20.0, 3.0), (6.5, 21.2), 1.0),
                                    ((1.1, 11.3), (30.7, 12.9), 2.5), ((5.0, 5.0), (5.0, 5.0), 1.0)])
def test_line_rule_matches_brute_force(a, b, lw):
    xyz, tan, th, idx = one_segment([a[0], a[1], 0.25], [b[0], b[1], 0.75])
    rp = params.render(renderer=0, adsm_on=0, shading_model=2)
    o = OR.raster(xyz, tan, th, idx, ortho_camera(), ortho_light(), strand_params(width=lw), rp)
    assert sorted(o["pixel"].tolist()) == _brute_force_rule(a, b, lw, W, H)
    # depth: linear in t along the window-space segment
    if len(o["pixel"]):
        px, py = o["pixel"] % W + 0.5, o["pixel"] // W + 0.5
        v = np.subtract(b, a)
        t = ((px - a[0]) * v[0] + (py - a[1]) * v[1]) / (v @ v)
        assert np.allclose(o["depth"], 0.25 + t * 0.5, atol=1e-12)


def test_shadow_depth_keeps_the_nearest():
    m = np.array([[2.0 / W, 0, 0, -1], [0, 2.0 / H, 0, -1], [0, 0, 1, 0], [0, 0, 0, 1]]).T.reshape(-1)
    xyz = np.array([[2.0, 5.5, 0.6], [10.0, 5.5, 0.2], [2.0, 5.5, 0.3], [10.0, 5.5, 0.9]], np.float32)
    idx = np.array([0, 1, 2, 3], np.uint32)
    d = OR.shadow_depth(xyz, idx, m, 1.0, (W, H))
    x = np.arange(2, 10) + 0.5
    t = (x - 2.0) / 8.0
    z = xyz[:, 2].astype(np.float64)            # the float32 vertex depths
    want = np.minimum(z[0] + t * (z[1] - z[0]), z[2] + t * (z[3] - z[2])).astype(np.float32)
    assert np.array_equal(d[5, 2:10], want)
    assert (np.delete(d.reshape(-1), 5 * W + np.arange(2, 10)) == 1.0).all()


def test_pcf_of_a_constant_shadow_map():
    xyz, tan, th, idx = one_segment([2.0, 5.5, 0.5], [10.0, 5.5, 0.5])
    alpha = 0.3
    for depth_in_map, light_depth in [(0.5, 0.5), (0.4999, 0.5), (0.45, 0.5), (0.6, 0.5)]:
        sm = np.full((16, 16), depth_in_map, np.float32)
        rp = params.render(renderer=0, shading_model=2)               # ADSM only
        o = OR.raster(xyz, tan, th, idx, ortho_camera(), ortho_light(), strand_params(alpha=alpha), rp,
                      shadow_map=sm)
        diff = max(light_depth - float(np.float32(depth_in_map)), 0.0)
        n = diff * 15000.0 + (1.0 if diff > 1e-5 else 0.0)
        a32 = float(np.float32(alpha))           # hair_opacity is a float32 parameter
        assert np.allclose(o["rgba"][:, :3], (1 - a32) ** n, rtol=1e-10), (depth_in_map, o["rgba"][0])


def _nodes_from(frags):
    """A one-pixel PPLL: frags = list of (rgba8 tuple, depth), inserted in
    order (so the last one is the head)."""
    nodes = []
    prev = 0xFFFFFFFF
    for (c, d) in frags:
        col = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24)
        nodes += [col, int(np.float32(d).view(np.uint32)), prev]
        prev = len(nodes) // 3 - 1
    return np.array([prev], np.uint32), np.array(nodes, np.uint32)


def test_resolve_up_to_16_is_sorted_over():
    rng = np.random.default_rng(1)
    for n in [1, 5, 16]:
        cols = rng.integers(0, 256, (n, 4))
        depths = rng.permutation(np.linspace(0.1, 0.9, n))
        heads, nodes = _nodes_from([(tuple(c), d) for c, d in zip(cols, depths)])
        bg = np.array([0.2, 0.4, 0.6, 1.0], np.float32)
        out = OR.resolve(heads, nodes, 1, 1, background=bg)[0, 0]
        want = bg.astype(np.float64)
        for i in np.argsort(-depths):        # back to front
            f = cols[i] / 255.0
            want = want * (1 - f[3]) + f * f[3]
        assert np.allclose(out, want, atol=1e-12)


@pytest.mark.parametrize("n,blended", [(30, 30), (16 + 1100, 16 + 1024)])
def test_resolve_identical_colours_and_fragment_cap(n, blended):
    c = (200, 100, 50, 64)
    heads, nodes = _nodes_from([(c, 0.5 + 1e-4 * i) for i in range(n)])
    out = OR.resolve(heads, nodes, 1, 1)[0, 0]
    a = 64 / 255.0
    f = np.array(c) / 255.0
    want = np.ones(4) * (1 - a) ** blended + f * (1 - (1 - a) ** blended)
    assert np.allclose(out, want, atol=1e-12)


def test_resolve_empty_pixel_is_background():
    out = OR.resolve(np.array([0xFFFFFFFF], np.uint32), np.zeros(3, np.uint32), 1, 1)
    assert (out == 1.0).all()
