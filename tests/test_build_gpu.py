"""The one-call volume build (vkhr_volume_build_segments: LDS-binned
voxelization + finalize) against the oracle's counts and, output for
output, against the two-call path vkhr_voxelize_segments +
vkhr_volume_finalize: counts, uint8 density, quantized tangents, float
view, corner-packed cells and the brick occupancy / distance field, all
bit-identical.  Long segments exercise runs over many bricks and, past the
bin capacity, the workspace path the build switches to on the device."""
import numpy as np
import pytest
import torch

from oracle import voxel as OV
from vkhr_amd import capi, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu

KEYS = ("counts", "density", "tangents", "density_f32", "cells", "occupancy")


def _two_call(xyz, b, res, idx, tan, normalize, want_tangents=True):
    counts, sums = D.voxelize(xyz, b, res, indices=idx, tangents=tan, want_sums=True, want_counts=True)
    vol = D.finalize(counts, res, sums=sums, normalize=normalize, want_f32=True, want_cells=True)
    vol["counts"] = counts
    if not want_tangents:
        vol["tangents"] = None
    return vol


def _compare(got, want, label):
    for k in KEYS:
        if want.get(k) is None:
            assert got.get(k) is None, (label, k)
            continue
        a, b = got[k], want[k]
        if not torch.equal(a, b):
            n = int((a != b).sum())
            raise AssertionError(f"{label}: {k} differs in {n} of {a.numel()} elements")


def _build(xyz, b, res, idx, tan, normalize, want_tangents=True):
    return D.build_volume(xyz, b, res, idx, tan, normalize=normalize, want_counts=True,
                          want_tangents=want_tangents, want_f32=True, want_cells=True)


def _style_inputs(style, device):
    xyz = D.to_dev(style.vertices, device)
    off = D.to_dev(style.strand_offsets, device)
    return xyz, D.generate_indices(off), D.generate_tangents(xyz, off), D.aabb(xyz)


@pytest.mark.parametrize("normalize", [True, False])
def test_10k_strands_128_matches_two_call_path_and_oracle(device, normalize):
    s = synth.strands_10k()
    res = (128, 128, 128)
    xyz, idx, tan, b = _style_inputs(s, device)
    got = _build(xyz, b, res, idx, tan, normalize)
    _compare(got, _two_call(xyz, b, res, idx, tan, normalize), f"10k/128 normalize={normalize}")
    ref = OV.voxelize(s.vertices, OV.aabb(s.vertices), res, indices=s.indices())
    assert np.array_equal(got["counts"].cpu().numpy().view(np.uint32), ref["counts"])


def test_136k_strands_256_matches_oracle_and_two_call_path(device):
    s = synth.strands_136k()
    res = (256, 256, 256)
    xyz, idx, tan, b = _style_inputs(s, device)
    # the load-time spatial order and the generated order give the same volume
    got = _build(xyz, b, res, idx, tan, True)
    got_sorted = _build(xyz, b, res, D.sort_segments(xyz, idx, b), tan, True)
    want = _two_call(xyz, b, res, idx, tan, True)
    _compare(got, want, "136k/256")
    _compare(got_sorted, want, "136k/256 sorted")
    ref = OV.voxelize(s.vertices, OV.aabb(s.vertices), res, indices=s.indices())
    assert np.array_equal(got["counts"].cpu().numpy().view(np.uint32), ref["counts"])


@pytest.mark.parametrize("res", [(100, 37, 65), (40, 56, 48), (9, 130, 17)])
def test_uneven_sizes(device, res):
    s = synth.hemisphere_style(3000, (2, 20), seed=21)
    xyz, idx, tan, b = _style_inputs(s, device)
    _compare(_build(xyz, b, res, idx, tan, True), _two_call(xyz, b, res, idx, tan, True), f"uneven {res}")


def test_without_tangent_volume(device):
    s = synth.hemisphere_style(4000, 16, seed=2)
    res = (64, 64, 64)
    xyz, idx, tan, b = _style_inputs(s, device)
    got = _build(xyz, b, res, idx, None, True, want_tangents=False)
    _compare(got, _two_call(xyz, b, res, idx, tan, True, want_tangents=False), "no tangents")


@pytest.mark.parametrize("normalize", [True, False])
def test_without_cells(device, normalize):
    """No cells: the brick pass writes the uint8 view into the output and
    the normalize runs in place, as in vkhr_volume_finalize."""
    s = synth.hemisphere_style(4000, 16, seed=12)
    res = (64, 48, 80)
    xyz, idx, tan, b = _style_inputs(s, device)
    got = D.build_volume(xyz, b, res, idx, tan, normalize=normalize, want_counts=True, want_f32=True,
                         want_cells=False)
    want = _two_call(xyz, b, res, idx, tan, normalize)
    want["cells"] = want["occupancy"] = None
    _compare(got, want, f"no cells, normalize={normalize}")


def _brick_runs(xyz, idx, b, res):
    """Bricks entered per pair: the float32 walk of hair_style.cc:311-329,
    vectorized over the pairs (numpy float32, no contraction)."""
    f = np.float32
    r = np.array(res, np.float32)
    vs = (b["size"] / r).astype(np.float32)
    root = ((xyz[idx[0::2]] - b["origin"]) / vs).astype(np.float32)
    tip = ((xyz[idx[1::2]] - b["origin"]) / vs).astype(np.float32)
    d = (tip - root).astype(np.float32)
    steps = np.abs(d).max(axis=1).astype(np.float32)
    d = np.where(steps[:, None] > 0, d / np.where(steps > 0, steps, f(1))[:, None], d).astype(np.float32)
    nb = (np.array(res) + 7) // 8
    last = np.full(len(root), -1, np.int64)
    runs = 0
    while (steps > 0).any():
        on = steps > 0
        steps = np.where(on, steps - f(1), steps).astype(np.float32)
        v = np.maximum(np.minimum(np.floor(root), r - f(1)), f(0)).astype(np.int64) >> 3
        k = v[:, 0] + nb[0] * (v[:, 1] + nb[1] * v[:, 2])
        runs += int((on & (k != last)).sum())
        last = np.where(on, k, last)
        root = np.where(on[:, None], root + d, root).astype(np.float32)
    return runs


@pytest.mark.parametrize("pairs,overflow", [(2000, False), (20000, True)])
def test_long_segments(device, pairs, overflow):
    """Random pairs across the whole box: each visits ~16 bricks.  With 2,000
    pairs the bins hold every run (runs over many bricks, the walk's early
    end after its brick); with 20,000 they pass the capacity (4 per pair +
    65536) and the build takes the workspace path."""
    rng = np.random.default_rng(pairs)
    verts = rng.uniform(-1.0, 1.0, size=(4096, 3)).astype(np.float32)
    idx = rng.integers(0, len(verts), size=2 * pairs).astype(np.uint32)
    res = (128, 128, 128)
    b = OV.aabb(verts)
    runs = _brick_runs(verts, idx, b, res)
    assert (runs > 4 * pairs + 65536) == overflow, runs
    xyz = D.to_dev(verts, device)
    tan = D.to_dev(rng.uniform(-1, 1, size=(len(verts), 3)).astype(np.float32), device)
    di = D.to_dev(idx, device)
    bd = D.aabb(xyz)
    got = _build(xyz, bd, res, di, tan, True)
    _compare(got, _two_call(xyz, bd, res, di, tan, True), f"long segments x{pairs}")
    ref = OV.voxelize(verts, b, res, indices=idx)
    assert np.array_equal(got["counts"].cpu().numpy().view(np.uint32), ref["counts"])


def test_buffers_reused_across_calls(device):
    """A second call into the same buffers (smaller style, same volume)
    leaves nothing of the first: every voxel is written by its brick."""
    res = (64, 64, 64)
    big = synth.hemisphere_style(6000, 16, seed=8)
    small = synth.hemisphere_style(300, 8, seed=9)
    xyz, idx, tan, b = _style_inputs(big, device)
    out = _build(xyz, b, res, idx, tan, True)
    xyz2, idx2, tan2, b2 = _style_inputs(small, device)
    ws = torch.empty(2 * 64 ** 3, dtype=torch.int64, device=device)
    again = D.build_volume(xyz2, b2, res, idx2, tan2, normalize=True, want_counts=True, want_f32=True,
                           workspace=ws, out=out)
    _compare(again, _two_call(xyz2, b2, res, idx2, tan2, True), "reused buffers")


def test_empty_and_degenerate_inputs(device):
    res = (16, 16, 16)
    b = dict(origin=np.zeros(3, np.float32), size=np.ones(3, np.float32), radius=np.float32(1.7),
             volume=np.float32(1.0))
    xyz = torch.tensor([[0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.5, 0.5, 0.5]], dtype=torch.float32, device=device)
    tan = torch.zeros_like(xyz)
    # no pairs at all, then a zero-length pair and a malformed one (index past the vertices)
    for pairs in ([], [0, 1], [0, 1, 2, 99]):
        idx = torch.tensor(pairs, dtype=torch.int32, device=device)
        got = _build(xyz, b, res, idx, tan, True)
        _compare(got, _two_call(xyz, b, res, idx, tan, True), f"pairs {pairs}")
        assert int(got["density"].sum()) == 0   # a constant volume normalizes to 0


def test_runs_split_past_1023_samples(device):
    """Vertices far outside the given box: the walk is clamped to the grid's
    edge, so one pair stays in one brick for ~80,000 samples.  The count
    and fill passes split such runs into bin entries of at most 1023
    samples; the volume equals the two-call path's and the oracle's counts
    (voxels past 32767 samples get tangent 0 on both paths)."""
    res = (16, 16, 16)
    b = dict(origin=np.zeros(3, np.float32), size=np.ones(3, np.float32), radius=np.float32(1.7),
             volume=np.float32(1.0))
    verts = np.array([[-5000.0, -5.0, 0.3], [5000.0, -5.0, 0.3], [0.2, 0.3, 0.4], [0.7, 0.6, 0.5],
                      [0.5, 3000.0, 0.5], [0.5, -3000.0, 0.6]], np.float32)
    tans = np.array([[1, 0, 0], [1, 0, 0], [0.6, 0.8, 0], [0.6, 0.8, 0], [0, -1, 0], [0, -1, 0]], np.float32)
    idx = np.array([0, 1, 2, 3, 4, 5], np.uint32)
    xyz, tan, di = D.to_dev(verts, device), D.to_dev(tans, device), D.to_dev(idx, device)
    got = _build(xyz, b, res, di, tan, True)
    _compare(got, _two_call(xyz, b, res, di, tan, True), "clamped runs")
    ref = OV.voxelize(verts, b, res, indices=idx)
    counts = got["counts"].cpu().numpy().view(np.uint32)
    assert np.array_equal(counts, ref["counts"])
    assert counts.max() > 32767


def test_invalid_arguments_are_errors(device):
    buf = torch.zeros(1 << 16, dtype=torch.int64, device=device)
    p = capi.ptr(buf)
    ab = capi.make_aabb(dict(origin=[0, 0, 0], size=[1, 1, 1], radius=1.7, volume=1.0))
    import ctypes as C
    s = capi.stream_of()
    args = [p, 4, p, 4, p, C.byref(ab), 8, 8, 8, 1, None, p, p, None, None, None]
    with pytest.raises(capi.VkhrError) as e:   # no workspace
        capi.call("vkhr_volume_build_segments", *args, None, s)
    assert e.value.code == -1
    with pytest.raises(capi.VkhrError) as e:   # tangent volume without tangents
        capi.call("vkhr_volume_build_segments", p, 4, p, 4, None, C.byref(ab), 8, 8, 8, 1, None, p, p, None,
                  None, None, p, s)
    assert e.value.code == -1
    with pytest.raises(capi.VkhrError) as e:   # 2^24 voxels exceeded
        capi.call("vkhr_volume_build_segments", p, 4, p, 4, p, C.byref(ab), 512, 512, 512, 1, None, p, p, None,
                  None, None, p, s)
    assert e.value.code == -3
