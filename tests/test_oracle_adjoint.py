"""Pins the float64 analytic gradient of the differentiable raymarch
(oracle/raymarch_oracle.cc: oracle_raymarch_grad) against central finite
differences of the oracle's own float64 forward, on a small smooth volume:
density voxels and the three parameters (absorption, ADSM thickness,
Kajiya-Kay exponent).  The forward is piecewise smooth (isosurface choices,
clamps), so a voxel whose difference quotient crosses a kink is allowed as a
rare outlier."""
import numpy as np
import pytest

from oracle import raymarch as OR
from vkhr_amd import params, synth

N = 12
W = H = 28


def small_scene(seed=3):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.0, 1.0, (N + 4, N + 4, N + 4))
    # smooth it so the trilinear field has moderate gradients
    for ax in range(3):
        d = (np.roll(d, 1, ax) + d + np.roll(d, -1, ax)) / 3.0
    dens = (0.12 * d[2:-2, 2:-2, 2:-2]).astype(np.float32).reshape(-1)
    tan = rng.integers(-127, 128, (N ** 3, 4)).astype(np.int8)
    tan[:, 3] = 0
    bounds = dict(origin=np.array([-1.0, -0.8, -1.1], np.float32), size=np.array([2.0, 1.6, 2.2], np.float32))
    bounds["radius"] = np.float32(np.linalg.norm(bounds["size"]))
    bounds["volume"] = np.float32(np.prod(bounds["size"]))
    sc = synth.make_scene(bounds, W, H, distance_scale=1.4, azimuth_deg=25.0)
    return dens, tan.reshape(-1), bounds, sc


def loss_and_grad(dens, tan, bounds, sc, diff, target):
    res = (N, N, N)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(bounds, res)
    rp = params.render(raycast_steps=48.0)
    rgba, _, _ = OR.raymarch_grad(tan, dens, res, cam, light, sp, rp, diff, W, H)
    g = 2.0 * (rgba - target)
    rgba2, dd, dp = OR.raymarch_grad(tan, dens, res, cam, light, sp, rp, diff, W, H,
                                     d_rgba=g.reshape(-1, 4).astype(np.float32))
    return float(((rgba - target) ** 2).sum()), dd, dp


def forward_loss(dens, tan, bounds, sc, diff, target):
    res = (N, N, N)
    rgba, _, _ = OR.raymarch_grad(tan, dens, res, params.camera(sc), params.light(sc), params.strand(bounds, res),
                                  params.render(raycast_steps=48.0), diff, W, H)
    return float(((rgba - target) ** 2).sum())


@pytest.fixture(scope="module")
def problem():
    dens, tan, bounds, sc = small_scene()
    diff = np.array([0.3, 11.0, 80.0], np.float32)
    rng = np.random.default_rng(9)
    noisy = np.clip(dens * (1 + 0.1 * rng.standard_normal(dens.shape)), 0, None).astype(np.float32)
    res = (N, N, N)
    target, _, _ = OR.raymarch_grad(tan, noisy, res, params.camera(sc), params.light(sc),
                                    params.strand(bounds, res), params.render(raycast_steps=48.0), diff, W, H)
    L, dd, dp = loss_and_grad(dens, tan, bounds, sc, diff, target)
    return dens, tan, bounds, sc, diff, target, L, dd, dp


def test_density_gradient_matches_central_differences(problem):
    dens, tan, bounds, sc, diff, target, L, dd, dp = problem
    assert L > 0 and np.abs(dd).max() > 0
    cand = np.nonzero(np.abs(dd) > 1e-3 * np.abs(dd).max())[0]
    rng = np.random.default_rng(4)
    picks = rng.choice(cand, min(40, len(cand)), replace=False)
    bad = 0
    for v in picks:
        up, dn = dens.copy(), dens.copy()
        up[v] = np.float32(dens[v] + 1e-5)
        dn[v] = np.float32(dens[v] - 1e-5)
        fd = (forward_loss(up, tan, bounds, sc, diff, target) - forward_loss(dn, tan, bounds, sc, diff, target)) / \
            (float(up[v]) - float(dn[v]))
        if abs(fd - dd[v]) > 1e-4 * np.abs(dd).max() + 1e-3 * abs(dd[v]):
            bad += 1
    assert bad <= 2, f"{bad} of {len(picks)} voxels off the difference quotient"


@pytest.mark.parametrize("k,h", [(0, 1e-4), (1, 1e-3), (2, 1e-2)])
def test_parameter_gradients_match_central_differences(problem, k, h):
    dens, tan, bounds, sc, diff, target, L, dd, dp = problem
    up, dn = diff.copy(), diff.copy()
    up[k] = np.float32(diff[k] + h)
    dn[k] = np.float32(diff[k] - h)
    fd = (forward_loss(dens, tan, bounds, sc, up, target) - forward_loss(dens, tan, bounds, sc, dn, target)) / \
        (float(up[k]) - float(dn[k]))
    assert fd == pytest.approx(dp[k], rel=2e-3, abs=1e-6), (fd, dp[k])
