"""Oracle (test infrastructure only) of the volume mip / filter operations:
Volume::downsample (include/vkhr/scene_graph/hair_style.hh:228-257) with the
max / min / mean reductions, and filter_volume
(share/shaders/volumes/sample_volume.glsl:12-35) at texel centres, in
numpy float64."""
from __future__ import annotations

import math

import numpy as np

MAX, MIN, MEAN = 0, 1, 2


def downsample(vol: np.ndarray, res, filt: int) -> np.ndarray:
    """vol: flat uint8 nx*ny*nz (x fastest) -> flat (nx/2)(ny/2)(nz/2)."""
    nx, ny, nz = res
    v = np.asarray(vol, np.uint32).reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2)
    if filt == MAX:
        r = v.max(axis=(1, 3, 5))
    elif filt == MIN:
        r = v.min(axis=(1, 3, 5))
    else:
        r = v.sum(axis=(1, 3, 5)) // 8
    return r.astype(np.uint8).reshape(-1)


def gaussian_weights(kernel_width: int) -> np.ndarray:
    """The shader's weights over the kernel_width^3 offsets, [z, y, x]:
    1/(2 pi s^2) * e^(-(x^2+y^2+z^2)/2 * s^2), s = kernel_width/2/2.4."""
    s2 = ((kernel_width / 2.0) / 2.4) ** 2
    r = kernel_width // 2
    o = np.arange(-r, r + 1, dtype=np.float64)
    z, y, x = np.meshgrid(o, o, o, indexing="ij")
    return 1.0 / (2.0 * math.pi * s2) * np.exp(-1.0 * (x * x + y * y + z * z) / 2.0 * s2)


def filter_volume(vol: np.ndarray, res, kernel_width: int) -> np.ndarray:
    """Gaussian-weighted mean of the texels around every texel centre (zero
    border; the total weight includes the taps outside the volume)."""
    nx, ny, nz = res
    v = np.asarray(vol, np.float64).reshape(nz, ny, nx)
    r = kernel_width // 2
    w = gaussian_weights(kernel_width)
    pad = np.zeros((nz + 2 * r, ny + 2 * r, nx + 2 * r))
    pad[r:r + nz, r:r + ny, r:r + nx] = v
    out = np.zeros_like(v)
    for dz in range(kernel_width):
        for dy in range(kernel_width):
            for dx in range(kernel_width):
                out += w[dz, dy, dx] * pad[dz:dz + nz, dy:dy + ny, dx:dx + nx]
    return (out / w.sum()).reshape(-1)
