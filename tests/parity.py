"""Shared parity checks of GPU renders against the float64 oracle."""
from __future__ import annotations

import numpy as np

from oracle import raymarch as OR

# The raymarch bar: max abs error per channel against the float64 oracle.
RAYMARCH_TOL = 1e-4
# Near-tie allowance: a pixel may take the other branch of the iso test only
# when the oracle's accumulated density is within this of the iso value at
# some nonzero sample (FP32 trilinear vs FP64 cannot order such ties), and
# such pixels must stay rare.
TIE_MARGIN = 1e-5
MAX_TIE_FRACTION = 1e-3


def compare_raymarch(gpu: dict, vol_np: dict, res, cam, light, strand, render, width, height, pixels=None,
                     depth_buffer=None, density_f32=None):
    """gpu: dict of numpy arrays rgba (H,W,4), depth (H,W), info (H,W,4).
    Returns a report dict; asserts the parity bar."""
    rgba = gpu["rgba"].reshape(-1, 4).astype(np.float64)
    depth = gpu["depth"].reshape(-1).astype(np.float64)
    info = gpu["info"].reshape(-1, 4)
    if pixels is None:
        pixels = np.arange(width * height, dtype=np.int64)
    o = OR.raymarch(vol_np.get("density"), vol_np.get("tangents"), res, cam, light, strand, render, width, height,
                    pixels=pixels, density_f32=density_f32, depth_buffer=depth_buffer)
    g_info = info[pixels]
    frag_g, frag_o = g_info[:, 0] == 1, o["info"][:, 0] == 1
    frag_mismatch = int((frag_g != frag_o).sum()) + int(((g_info[:, 0] == 0) != (o["info"][:, 0] == 0)).sum())
    both = frag_g & frag_o
    decision_diff = both & (np.any(g_info[:, 1:] != o["info"][:, 1:], axis=1))
    n_ties = int(decision_diff.sum())
    if n_ties:
        margins = o["diag"][decision_diff, 2]
        assert (margins < TIE_MARGIN).all(), f"decision flips away from a tie: margins {np.sort(margins)[-5:]}"
    assert n_ties <= max(2, MAX_TIE_FRACTION * both.sum()), f"{n_ties} near-tie flips of {both.sum()} fragments"
    want = o["rgba"].copy()
    want_depth = o["depth"].copy()
    if n_ties:
        sel = np.nonzero(decision_diff)[0]
        forced = g_info[sel, 0:4].astype(np.int32)
        f = OR.raymarch(vol_np.get("density"), vol_np.get("tangents"), res, cam, light, strand, render, width,
                        height, pixels=pixels[sel], density_f32=density_f32, depth_buffer=depth_buffer,
                        forced=forced)
        want[sel] = f["rgba"]
        want_depth[sel] = f["depth"]
    assert frag_mismatch <= max(2, 1e-4 * len(pixels)), f"{frag_mismatch} fragment-mask mismatches"
    ok = frag_g == frag_o
    err = np.abs(rgba[pixels][ok] - want[ok])
    max_err = float(err.max()) if err.size else 0.0
    derr = np.abs(depth[pixels][ok] - want_depth[ok])
    max_derr = float(derr.max()) if derr.size else 0.0
    report = dict(pixels=len(pixels), fragments=int(both.sum()), ties=n_ties, frag_mismatch=frag_mismatch,
                  max_abs_err=max_err, max_depth_err=max_derr,
                  p9999_err=float(np.quantile(err.max(axis=1), 0.9999)) if err.size else 0.0)
    assert max_err <= RAYMARCH_TOL, report
    assert max_derr <= 1e-6, report
    return report
