"""Builds the C-ABI parameter blocks (camera, light, strand and renderer
parameters) from a synthetic scene, in the layouts of the reference's
uniform blocks.  The same bytes feed the kernels and the oracle."""
from __future__ import annotations

import numpy as np

from . import capi, synth


def _cm(m: np.ndarray):
    """Row-major math matrix -> column-major float[16] (glm memory order)."""
    return [float(x) for x in np.asarray(m, np.float64).T.reshape(-1)]


def camera(scene: synth.Scene) -> capi.Camera:
    c = capi.Camera()
    c.view[:] = _cm(scene.view)
    c.projection[:] = _cm(scene.projection)
    c.position[:] = [float(x) for x in scene.eye]
    c.look_at_distance = float(scene.distance)
    c.znear, c.zfar = float(scene.znear), float(scene.zfar)
    c.resolution[:] = [float(scene.width), float(scene.height)]
    return c


def light(scene: synth.Scene) -> capi.Light:
    l = capi.Light()
    l.vector[:] = [float(x) for x in scene.light_vector]
    l.intensity[:] = [float(x) for x in scene.light_intensity] + [0.0]
    l.matrix[:] = _cm(scene.light_matrix)
    l.origin[:] = [float(x) for x in scene.light_origin]
    l.znear, l.zfar = 1.0, float(2.0 * scene.distance)
    return l


def strand(bounds: dict, res, **overrides) -> capi.StrandParams:
    h = dict(synth.HAIR_DEFAULTS, **overrides)
    s = capi.StrandParams()
    s.volume_bounds = capi.make_aabb(bounds)
    s.volume_resolution[:] = [float(r) for r in res]
    s.strand_radius = h["strand_radius"]
    s.hair_color[:] = list(h["hair_color"])
    s.hair_opacity = h["hair_opacity"]
    s.hair_shininess = h["hair_shininess"]
    s.strand_ratio = h["strand_ratio"]
    return s


def render(**overrides) -> capi.RenderParams:
    p = dict(synth.RENDER_DEFAULTS, **overrides)
    r = capi.RenderParams()
    for k, v in p.items():
        setattr(r, k, v)
    return r
