"""Seeded synthetic hair styles and the scene set-up (camera, light, parameters).

The reference renders assets (`share/styles/*.hair`) that are not available, so
every workload here is generated: strands rooted area-uniformly on a unit
scalp hemisphere, grown along the surface normal with a gravity bend and
per-vertex Gaussian curl noise.  All arithmetic is float64 and rounded once to
float32, so a seed fixes the bytes on every machine (numpy's PCG64 stream is
platform independent).

The scene matrices follow the reference's conventions:
  * camera: glm::perspective (right-handed, depth 0..1) with the Vulkan Y flip
    (`src/vkhr/scene_graph/camera.cc:225-249`), glm::lookAt view matrix;
  * light: a directional light whose "spotlight origin" sits at the camera's
    look-at point plus direction x camera distance, looking at the look-at
    point (`src/vkhr/scene_graph/light_source.cc:112-139`), with a 45 degree
    square perspective of near 1 and far 2 x distance
    (`light_source.cc:61-69`, `src/vkhr/scene_graph.cc:124-127`), sampled
    through the bias matrix (`light_source.cc:145-150`).
Matrices are stored column-major (glm's memory order), float32.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Defaults of Interface::Parameters (include/vkhr/rasterizer/interface.hh:86-112),
# with the raymarch step count of the benchmarks (include/vkhr/benchmark.hh:20).
RENDER_DEFAULTS = dict(
    shading_model=0,          # Kajiya-Kay
    adsm_kernel_size=3,
    adsm_sampling_type=0,
    adsm_stride_size=4,
    adsm_on=1,
    ctsm_kernel_size=5,
    ctsm_sampling_type=1,
    ctsm_bias=0.00005,
    ctsm_on=1,
    shadow_technique=1,
    isosurface=0.115,
    raycast_steps=512.0,
    occlusion_radius=2.5,
    ao_exponent=10.0,
    ao_clamp=0.16,
    lod_magnified_distance=400.0,
    renderer=2,               # Raymarcher
    lod_minified_distance=800.0,
    benchmarking=0,
)

# Hair appearance.  The .hair assets carry these in their header; the synthetic
# styles fix them here (hair_shininess = 80 as in
# src/vkhr/rasterizer/hair_style.cc:59).
HAIR_DEFAULTS = dict(
    strand_radius=1.0,         # line width in pixels (header default_thickness)
    hair_color=(0.32, 0.228, 0.128),
    hair_opacity=0.3,          # header default_transparency -> hair_alpha
    hair_shininess=80.0,
    strand_ratio=1.0,
)

LIGHT_DIRECTION = (1.0, 2.0, 1.0)           # share/scenes/ponytail.vkhr:11
LIGHT_INTENSITY = (0.430, 0.380, 0.280)     # share/scenes/ponytail.vkhr:12-14
CAMERA_FOV_DEG = 25.0                       # share/scenes/ponytail.vkhr:3
CAMERA_NEAR = 0.1
CAMERA_FAR = 1000.0


@dataclass
class Style:
    """One hair style: flat float32 xyz vertices plus per-strand segment counts."""
    vertices: np.ndarray          # (V, 3) float32
    segments: np.ndarray          # (n_strands,) uint16, segments per strand
    thickness: np.ndarray         # (V,) float32
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_strands(self) -> int:
        return int(self.segments.shape[0])

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def strand_offsets(self) -> np.ndarray:
        off = np.zeros(self.n_strands + 1, dtype=np.int64)
        off[1:] = np.cumsum(self.segments.astype(np.int64) + 1)
        return off.astype(np.uint32)

    def indices(self) -> np.ndarray:
        """Line-list index pairs (v, v+1) within each strand, strand ends skipped
        (src/vkhr/scene_graph/hair_style.cc:196-213)."""
        off = self.strand_offsets.astype(np.int64)
        seg = self.segments.astype(np.int64)
        n_seg = int(seg.sum())
        strand_of_seg = np.repeat(np.arange(self.n_strands), seg)
        local = np.arange(n_seg) - np.repeat(np.cumsum(seg) - seg, seg)
        first = off[:-1][strand_of_seg] + local
        out = np.empty((n_seg, 2), dtype=np.uint32)
        out[:, 0] = first
        out[:, 1] = first + 1
        return out.reshape(-1)


def hemisphere_style(n_strands: int, verts_per_strand, seed: int,
                     segment_length: float = 0.05, curl_sigma: float = 0.01,
                     gravity: float = 0.9, thickness=None) -> Style:
    """Strands rooted area-uniformly on the unit hemisphere y >= 0.

    verts_per_strand: an int (all strands equal) or a (lo, hi) pair drawn
    uniformly per strand.  thickness: None gives the reference's
    generate_thickness(0.042) profile (tip vertex 0, hair_style.cc:155-169);
    a (lo, hi) pair draws U[lo, hi] per vertex.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    y = rng.uniform(0.0, 1.0, n_strands)
    phi = rng.uniform(0.0, 2.0 * math.pi, n_strands)
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    normal = np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)

    if isinstance(verts_per_strand, (tuple, list)):
        lo, hi = verts_per_strand
        nv = rng.integers(lo, hi + 1, n_strands)
    else:
        nv = np.full(n_strands, int(verts_per_strand), dtype=np.int64)
    nv = nv.astype(np.int64)
    total = int(nv.sum())
    strand = np.repeat(np.arange(n_strands), nv)
    start = np.cumsum(nv) - nv
    local = np.arange(total) - np.repeat(start, nv)
    frac = local / np.repeat(np.maximum(nv - 1, 1), nv)

    n = normal[strand]
    d = n.copy()
    d[:, 1] -= gravity * frac
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    step = segment_length * d + rng.normal(0.0, curl_sigma, (total, 3))
    step[local == 0] = 0.0
    pos = np.cumsum(step, axis=0)
    # restart the running sum at each strand root
    base = pos[start] - normal
    pos -= np.repeat(base, nv, axis=0)

    if thickness is None:
        th = np.full(total, 0.042)
        th[(local == nv[strand] - 1)] = 0.0
    else:
        th = rng.uniform(thickness[0], thickness[1], total)

    return Style(vertices=np.ascontiguousarray(pos.astype(np.float32)),
                 segments=(nv - 1).astype(np.uint16),
                 thickness=th.astype(np.float32), seed=seed,
                 meta=dict(n_strands=n_strands, verts_per_strand=verts_per_strand,
                           segment_length=segment_length, curl_sigma=curl_sigma,
                           gravity=gravity))


# --- the named workloads -----------------------------------------------------

def strands_10k(seed: int = 0) -> Style:
    """10,000 strands x 16 vertices (the reference's CPU-runnable case)."""
    return hemisphere_style(10_000, 16, seed)


def strands_136k(seed: int = 1) -> Style:
    """136,320 strands x 13 vertices, thickness U[0.5, 1.0] (ponytail-sized)."""
    return hemisphere_style(136_320, 13, seed, thickness=(0.5, 1.0))


def batch_styles(n_styles: int = 32, seed: int = 2, lo: int = 5_000, hi: int = 140_000):
    """Styles of uneven size: strands log-uniform in [lo, hi], 8-32 vertices
    per strand, volume edge 64/128/256 by strand-count tertile."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), n_styles))).astype(np.int64)
    t1, t2 = np.quantile(counts, [1.0 / 3.0, 2.0 / 3.0])
    styles = []
    for i, c in enumerate(counts):
        s = hemisphere_style(int(c), (8, 32), seed * 1000 + i)
        s.meta["volume"] = 64 if c <= t1 else (128 if c <= t2 else 256)
        styles.append(s)
    return styles, dict(tertile_thresholds=[float(t1), float(t2)])


# --- scene matrices ------------------------------------------------------------

def _normalize(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.sqrt(np.dot(v, v))


def look_at(eye, center, up) -> np.ndarray:
    """Right-handed look-at view matrix, column-major 4x4 (as glm::lookAt)."""
    eye = np.asarray(eye, np.float64)
    f = _normalize(np.asarray(center, np.float64) - eye)
    s = _normalize(np.cross(f, up))
    u = np.cross(s, f)
    m = np.eye(4)
    m[0:3, 0] = [s[0], u[0], -f[0]]
    m[0:3, 1] = [s[1], u[1], -f[1]]
    m[0:3, 2] = [s[2], u[2], -f[2]]
    m[0:3, 3] = [-np.dot(s, eye), -np.dot(u, eye), np.dot(f, eye)]
    return m  # row-major math matrix M (so clip = M @ x)


def perspective_zo(fovy, aspect, znear, zfar) -> np.ndarray:
    """Right-handed perspective with depth in [0, 1] (GLM_FORCE_DEPTH_ZERO_TO_ONE)."""
    t = math.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = zfar / (znear - zfar)
    m[3, 2] = -1.0
    m[2, 3] = -(zfar * znear) / (zfar - znear)
    return m


BIAS = np.array([[0.5, 0.0, 0.0, 0.5],
                 [0.0, 0.5, 0.0, 0.5],
                 [0.0, 0.0, 1.0, 0.0],
                 [0.0, 0.0, 0.0, 1.0]])


def aabb_of(vertices: np.ndarray):
    """AABB with min/max initialised at the origin (hair_style.cc:215-255).
    float32 semantics: exact min/max, size = max - min, radius = |size|."""
    v = np.asarray(vertices, np.float32)
    mn = np.minimum(v.min(axis=0), np.float32(0.0)) if len(v) else np.zeros(3, np.float32)
    mx = np.maximum(v.max(axis=0), np.float32(0.0)) if len(v) else np.zeros(3, np.float32)
    size = (mx - mn).astype(np.float32)
    d = np.float32(size[0] * size[0]) + np.float32(size[1] * size[1])
    d = np.float32(d + np.float32(size[2] * size[2]))
    radius = np.float32(np.sqrt(d))
    vol = np.float32(np.float32(size[0] * size[1]) * size[2])
    return dict(origin=mn.astype(np.float32), size=size, radius=radius, volume=vol)


@dataclass
class Scene:
    width: int
    height: int
    view: np.ndarray             # 4x4 row-major math matrices (float32 values)
    projection: np.ndarray
    eye: np.ndarray
    look_at: np.ndarray
    distance: float
    light_origin: np.ndarray
    light_vp: np.ndarray         # unbiased light view-projection (shadow-map pass)
    light_matrix: np.ndarray     # bias x light view-projection (sampling)
    light_intensity: np.ndarray
    light_vector: np.ndarray
    znear: float = CAMERA_NEAR
    zfar: float = CAMERA_FAR


def make_scene(aabb: dict, width: int, height: int, distance_scale: float = 1.8,
               azimuth_deg: float = 0.0, fov_deg: float = CAMERA_FOV_DEG) -> Scene:
    """Camera on a ring around the AABB centre at distance_scale x AABB radius."""
    centre = aabb["origin"].astype(np.float64) + 0.5 * aabb["size"].astype(np.float64)
    dist = distance_scale * float(aabb["radius"])
    a = math.radians(azimuth_deg)
    eye = centre + dist * np.array([math.sin(a), 0.0, math.cos(a)])
    f32 = lambda x: np.asarray(x, np.float64).astype(np.float32).astype(np.float64)
    eye, centre = f32(eye), f32(centre)
    view = f32(look_at(eye, centre, (0.0, 1.0, 0.0)))
    proj = perspective_zo(math.radians(fov_deg), width / height, CAMERA_NEAR, CAMERA_FAR)
    proj[1, 1] *= -1.0                      # Vulkan Y flip (camera.cc:229)
    proj = f32(proj)
    ldir = _normalize(LIGHT_DIRECTION)
    real_dist = float(np.sqrt(np.sum((eye - centre) ** 2)))
    lorigin = f32(centre + ldir * real_dist)
    lview = look_at(lorigin, centre, (0.0, 1.0, 0.0))
    lproj = perspective_zo(math.radians(45.0), 1.0, 1.0, 2.0 * real_dist)
    lproj[1, 1] *= -1.0
    lvp = f32(lproj @ lview)
    return Scene(width=width, height=height, view=view, projection=proj, eye=eye,
                 look_at=centre, distance=real_dist, light_origin=lorigin,
                 light_vp=lvp, light_matrix=f32(BIAS @ lvp),
                 light_intensity=np.array(LIGHT_INTENSITY, np.float64),
                 light_vector=f32(np.append(ldir, 0.0)))
