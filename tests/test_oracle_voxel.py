"""Pins the voxelization oracle (oracle/voxel_oracle.cc) before it is trusted.

The reference ships no tests or fixtures, so the oracle is pinned by
  * hand-derived answers of hair_style.cc:257-357 on tiny grids (unit voxels,
    so every sample position is exact and the answer follows by hand),
  * invariants (count totals, saturating view, geometry of the DDA samples
    against exact real-number positions),
  * an independent float32 numpy walk of the same loop on random inputs,
  * committed fixtures (tests/golden/voxel_small.npz) that freeze the
    oracle's own outputs on small random styles.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import voxel as OV
from vkhr_amd import synth

GOLDEN = Path(__file__).resolve().parent / "golden"
UNIT4 = dict(origin=np.zeros(3, np.float32), size=np.full(3, 4.0, np.float32),
             radius=np.float32(math.sqrt(48.0)), volume=np.float32(64.0))


def lin(x, y, z, n=4):
    return x + n * y + n * n * z


def test_aabb_contains_origin_quirk():
    v = np.array([[1, 2, 3], [2, 3, 4]], np.float32)
    b = OV.aabb(v)
    assert b["origin"].tolist() == [0, 0, 0]
    assert b["size"].tolist() == [2, 3, 4]
    assert b["volume"] == 24.0
    assert b["radius"] == np.float32(math.sqrt(29.0))
    v = np.array([[-1, -2, -3], [-0.5, -1, -1]], np.float32)
    b = OV.aabb(v)
    assert b["origin"].tolist() == [-1, -2, -3]
    assert b["size"].tolist() == [1, 2, 3]


def test_vertex_voxels_by_hand():
    v = np.array([[0.5, 0.5, 0.5], [3.5, 0.5, 0.5], [0.5, 3.5, 0.5], [0.5, 0.5, 3.5],
                  [4.0, 4.0, 4.0], [2.0, 1.0, 3.0]], np.float32)
    r = OV.voxelize(v, UNIT4, (4, 4, 4))
    want = np.zeros(64, np.uint32)
    for x, y, z in [(0, 0, 0), (3, 0, 0), (0, 3, 0), (0, 0, 3), (3, 3, 3), (2, 1, 3)]:
        want[lin(x, y, z)] += 1
    assert (r["counts"] == want).all()
    assert (r["dens_u8"] == want).all()


@pytest.mark.parametrize("root,tip,voxels", [
    # integer length: ceil(3) = 3 samples from the root; the tip voxel is not visited
    ((0.5, 1.5, 2.5), (3.5, 1.5, 2.5), [(0, 1, 2), (1, 1, 2), (2, 1, 2)]),
    # fractional length 2.5 -> 3 samples
    ((0.25, 0.5, 0.5), (2.75, 0.5, 0.5), [(0, 0, 0), (1, 0, 0), (2, 0, 0)]),
    # diagonal: steps = 2 along x, y advances 0.5 per sample
    ((0.5, 0.5, 0.5), (2.5, 1.5, 0.5), [(0, 0, 0), (1, 1, 0)]),
    # negative direction along z
    ((1.5, 2.5, 3.5), (1.5, 2.5, 0.5), [(1, 2, 3), (1, 2, 2), (1, 2, 1)]),
    # zero length contributes nothing
    ((1.5, 1.5, 1.5), (1.5, 1.5, 1.5), []),
    # a sample beyond the last voxel clamps to res-1
    ((3.5, 0.5, 0.5), (5.5, 0.5, 0.5), [(3, 0, 0), (3, 0, 0)]),
])
def test_segment_dda_by_hand(root, tip, voxels):
    v = np.array([root, tip], np.float32)
    r = OV.voxelize(v, UNIT4, (4, 4, 4), indices=np.array([0, 1], np.uint32))
    want = np.zeros(64, np.uint32)
    for x, y, z in voxels:
        want[lin(x, y, z)] += 1
    assert (r["counts"] == want).all(), np.nonzero(r["counts"])


def test_saturation_and_first_255_tangent_rule():
    v = np.tile(np.array([[1.5, 1.5, 1.5]], np.float32), (300, 1))
    t = np.zeros_like(v)
    t[:, 0] = 1.0
    t[255:, 0] = -1.0          # contributions after the 255th are dropped by the reference
    r = OV.voxelize(v, UNIT4, (4, 4, 4), tangents=t, want_sums=True)
    k = lin(1, 1, 1)
    assert r["counts"][k] == 300 and r["dens_u8"][k] == 255
    assert r["sums_ref"][3 * k] == 255.0
    assert r["sums_all"][3 * k] == 255.0 - 45.0


def test_quantize_tangents_by_hand():
    sums = np.array([3.0, 0, 0, -0.5, 0.25, 0.999, 0, 0, 0], np.float32)
    den = np.array([3, 1, 0], np.uint32)
    q = OV.quantize_tangents(sums, den).reshape(-1, 4)
    assert q[0].tolist() == [127, 0, 0, 0]
    assert q[1].tolist() == [-63, 31, 126, 0]   # trunc(-63.5), trunc(31.75), trunc(126.873)
    assert q[2].tolist() == [0, 0, 0, 0]         # empty voxel: 0/0 -> 0


def test_normalize_closed_forms():
    ident = np.arange(256, dtype=np.uint8)
    assert (OV.normalize(ident) == ident).all()
    d = np.array([0, 1, 2, 15, 7], np.uint8)          # max 15 -> scale 17 exactly
    assert OV.normalize(d).tolist() == [0, 17, 34, 255, 119]
    d = np.array([5, 6, 10], np.uint8)                # scale 51
    assert OV.normalize(d).tolist() == [0, 51, 255]
    assert OV.normalize(np.full(5, 9, np.uint8)).tolist() == [0] * 5


def test_indices_and_tangents_by_hand():
    off = np.array([0, 3, 4, 6], np.uint32)  # strands of 3, 1 and 2 vertices
    assert OV.generate_indices(off).tolist() == [0, 1, 1, 2, 4, 5]
    xyz = np.array([[0, 0, 0], [3, 4, 0], [3, 4, 2], [9, 9, 9], [1, 1, 1], [1, 1, 3]], np.float32)
    t = OV.generate_tangents(xyz, off)
    assert np.allclose(t[0], [0.6, 0.8, 0]) and (t[1] == [0, 0, 1]).all() and (t[2] == t[1]).all()
    assert (t[3] == 0).all()          # single-vertex strand: zero tangent
    assert (t[4] == [0, 0, 1]).all() and (t[5] == t[4]).all()


def _numpy_dda(xyz, idx, b, res):
    """Independent float32 walk of hair_style.cc:311-329 (numpy scalars, no FMA)."""
    f = np.float32
    res = np.array(res, np.float32)
    vs = (b["size"] / res).astype(np.float32)
    counts = np.zeros(int(np.prod(res)), np.uint64)
    w, h = int(res[0]), int(res[1])
    for i in range(0, len(idx) - 1, 2):
        root = ((xyz[idx[i]] - b["origin"]) / vs).astype(np.float32)
        tip = ((xyz[idx[i + 1]] - b["origin"]) / vs).astype(np.float32)
        d = (tip - root).astype(np.float32)
        steps = f(max(abs(d[0]), abs(d[1]), abs(d[2])))
        d = (d / steps).astype(np.float32) if steps > 0 else d
        while steps > f(0):
            steps = f(steps - f(1))
            v = np.minimum(np.floor(root), res - f(1))
            v = np.maximum(v, f(0))
            counts[int(v[0]) + w * (int(v[1]) + h * int(v[2]))] += 1
            root = (root + d).astype(np.float32)
    return counts


def test_segments_match_independent_numpy_walk():
    s = synth.hemisphere_style(60, (2, 9), seed=11)
    b = OV.aabb(s.vertices)
    for res in [(16, 16, 16), (32, 24, 20)]:
        r = OV.voxelize(s.vertices, b, res, indices=s.indices())
        assert (r["counts"] == _numpy_dda(s.vertices, s.indices(), b, res)).all()


def test_segment_total_is_sum_of_ceil_linf_and_samples_lie_on_segment():
    s = synth.hemisphere_style(200, (2, 12), seed=5)
    b = OV.aabb(s.vertices)
    res = np.array([32, 32, 32])
    idx = s.indices()
    r = OV.voxelize(s.vertices, b, tuple(res), indices=idx)
    vs = b["size"] / res.astype(np.float32)
    root = ((s.vertices[idx[0::2]] - b["origin"]) / vs).astype(np.float32)
    tip = ((s.vertices[idx[1::2]] - b["origin"]) / vs).astype(np.float32)
    linf = np.abs(tip - root).max(axis=1)
    assert r["counts"].sum() == np.ceil(linf).astype(np.int64).sum()
    # Exact-real positions root + k*dir hit the same voxels except within a
    # rounding distance of voxel faces.
    want = np.zeros(res.prod(), np.int64)
    ambiguous = 0
    for a, t, L in zip(root.astype(np.float64), tip.astype(np.float64), linf.astype(np.float64)):
        if L == 0:
            continue
        for k in range(int(math.ceil(L))):
            p = a + k * (t - a) / L
            frac = np.abs(p - np.round(p))
            if (frac < 1e-4).any():
                ambiguous += 1
            v = np.clip(np.floor(p), 0, res - 1).astype(int)
            want[v[0] + res[0] * (v[1] + res[1] * v[2])] += 1
    diff = np.abs(want - r["counts"].astype(np.int64)).sum()
    assert diff <= 2 * ambiguous
    assert ambiguous < 0.01 * r["counts"].sum()


def test_vertex_total_and_saturating_view():
    s = synth.strands_10k()
    b = OV.aabb(s.vertices)
    r = OV.voxelize(s.vertices, b, (8, 8, 8))
    assert r["counts"].sum() == s.n_vertices
    assert (r["dens_u8"] == np.minimum(r["counts"], 255)).all()
    assert (r["counts"] > 255).any()   # the coarse grid saturates: the view is exercised


def test_golden_fixtures_unchanged():
    """Frozen oracle outputs (tests/golden/make_voxel_fixtures.py)."""
    z = np.load(GOLDEN / "voxel_small.npz")
    for k in range(int(z["n_cases"])):
        xyz, idx, res = z[f"xyz{k}"], z[f"idx{k}"], tuple(int(v) for v in z[f"res{k}"])
        b = OV.aabb(xyz)
        assert np.array_equal(OV.aabb_array(b), z[f"aabb{k}"])
        seg = OV.voxelize(xyz, b, res, indices=idx)
        vtx = OV.voxelize(xyz, b, res)
        assert np.array_equal(seg["counts"], z[f"seg{k}"])
        assert np.array_equal(vtx["counts"], z[f"vtx{k}"])


def test_openmp_count_variant_matches():
    s = synth.strands_10k()
    b = OV.aabb(s.vertices)
    idx = s.indices()
    want = OV.voxelize(s.vertices, b, (64, 64, 64), indices=idx)["counts"]
    OV.set_threads(4)
    got = OV.voxelize_counts_omp(s.vertices, b, (64, 64, 64), idx)
    OV.set_threads(1)
    assert np.array_equal(got, want)
