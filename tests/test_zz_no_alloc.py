"""The allocation-free finalize (vkhr_volume_finalize_noalloc) against
vkhr_volume_finalize, through the C ABI: two calls in a row, the second on
a smaller volume whose min/max differ from the first's (so min/max slots
left over from the first call would change the normalize), every output
bit-identical, and no per-stream scratch area created by the variant."""
import numpy as np
import pytest
import torch

from vkhr_amd import capi, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu

CASES = (
    # (strands, vertices per strand, seed, volume, segment voxelization)
    (10_000, 16, 3, (128, 128, 128), True),
    (400, 12, 5, (40, 56, 48), False),
)


def _voxelized(n, v, seed, res, segments, device):
    s = synth.hemisphere_style(n, v, seed=seed)
    xyz = D.to_dev(s.vertices, device)
    off = D.to_dev(s.strand_offsets, device)
    tan = D.generate_tangents(xyz, off)
    b = D.aabb(xyz)
    idx = D.generate_indices(off) if segments else None
    _, sums = D.voxelize(xyz, b, res, indices=idx, tangents=tan, want_sums=True, want_counts=False)
    return sums


def _u8_range(sums):
    counts = (sums.view(-1, 2)[:, 1] & 0xFFFFFFFF).clamp(max=255)
    return int(counts.min()), int(counts.max())


def test_noalloc_finalize_matches_finalize_on_two_calls(device):
    stream = torch.cuda.Stream()
    ranges = []
    for n, v, seed, res, segments in CASES:
        sums = _voxelized(n, v, seed, res, segments, device)
        ranges.append(_u8_range(sums))
        for normalize in (True, False):
            ref = D.finalize(None, res, sums=sums, normalize=normalize, want_f32=True, want_cells=True)
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                got = D.finalize(None, res, sums=sums, normalize=normalize, want_f32=True, want_cells=True,
                                 noalloc=True)
            stream.synchronize()
            for key, want in ref.items():
                assert want is not None and torch.equal(got[key], want), (res, normalize, key)
        # the variant kept no per-stream area for its stream
        assert capi.lib.vkhr_scratch_bytes(capi.stream_of(stream)) == 0
    # the second call's min/max differ from the first's: stale slots would show
    assert ranges[0] != ranges[1], ranges
    assert ranges[1][1] < 255


def test_noalloc_finalize_rejects_what_finalize_rejects(device):
    buf = torch.zeros(1 << 12, dtype=torch.int64, device=device)
    p = capi.ptr(buf)
    with pytest.raises(capi.VkhrError) as e:
        capi.call("vkhr_volume_finalize_noalloc", p, p, 0xFFFFFFFF, 8, 8, 1, p, p, p, p, p, capi.stream_of())
    assert e.value.code == -3
    with pytest.raises(capi.VkhrError) as e:
        capi.call("vkhr_volume_finalize_noalloc", p, p, 8, 8, 8, 1, None, p, None, None, None, capi.stream_of())
    assert e.value.code == -1
