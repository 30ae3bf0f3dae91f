"""The batched form: 32 synthetic styles of uneven size (5,000-140,000
strands log-uniform, 8-32 vertices per strand, 64^3/128^3/256^3 volumes by
strand-count tertile) voxelized, finalized and raymarched at 960x540 by one
graph-captured launch; every output bit-identical to the per-style
sequence, and one style per volume size against the oracle (voxelization and
image).  Also: graph replays reproduce the first launch, and the tuning
table's shapes are valid."""
import numpy as np
import pytest
import torch

from vkhr_amd import capi, synth
from vkhr_amd.batch import Batch, StyleBuffers, load_tuning

pytestmark = pytest.mark.gpu
W, H = 960, 540


@pytest.fixture(scope="module")
def batch32(device):
    styles, meta = synth.batch_styles()
    bufs = [StyleBuffers(s, s.meta["volume"], W, H, device, azimuth=11.25 * i) for i, s in enumerate(styles)]
    return bufs


def _plain_outputs(b):
    """The per-style sequence's outputs for style b (with a tangent volume
    when b has one)."""
    nv = b.n ** 3
    dev = b.xyz.device
    tangents = b.acc is not None
    out = dict(counts=torch.empty(nv, dtype=torch.int32, device=dev),
               acc=torch.empty(2 * nv, dtype=torch.int64, device=dev) if tangents else None,
               density=torch.empty(nv, dtype=torch.uint8, device=dev),
               tangents=torch.empty(4 * nv, dtype=torch.int8, device=dev) if tangents else None,
               cells=torch.empty((b.n + 1) ** 3, dtype=torch.int64, device=dev),
               occupancy=torch.empty_like(b.occupancy),
               rgba=torch.empty((H, W, 4), dtype=torch.float32, device=dev))
    b.run_plain(out)
    return out


@pytest.mark.parametrize("tuning", [None, "table", (16, 4, 2, 1), (32, 2, 8, 1), (8, 8, 0, 2), (16, 4, 4, 5),
                                    (8, 8, 0, 32)])
def test_batch_matches_per_style_bit_for_bit(batch32, tuning):
    """Every launch shape and pipeline depth (groups of styles, up to one
    style per group) gives the per-style outputs bit for bit."""
    if tuning == "table":
        tuning = load_tuning("batch")
        if tuning is None:
            pytest.skip("no tuning table")
    elif tuning is not None:
        t = capi.Tuning()
        t.tile_w, t.tile_h, t.blocks_per_cu, t.groups = tuning
        tuning = t
    batch = Batch(batch32, W, H, tuning=tuning)
    batch.launch()
    torch.cuda.synchronize()
    first = [b.rgba.clone() for b in batch32[:4]]
    for b in batch32:
        ref = _plain_outputs(b)
        for k, mine in [("counts", b.counts), ("density", b.density), ("tangents", b.tangents),
                        ("cells", b.cells), ("occupancy", b.occupancy), ("rgba", b.rgba)]:
            assert torch.equal(mine, ref[k]), (k, b.n)
        del ref
    # replays of the captured graph give the same frame
    batch.launch()
    batch.launch()
    torch.cuda.synchronize()
    for b, f in zip(batch32[:4], first):
        assert torch.equal(b.rgba, f)
    assert sum(int((b.rgba[..., 3] > 0).sum()) for b in batch32) > 1_000_000


def test_batch_matches_per_style_one_call_build(batch32):
    """The per-style sequence with the one-call volume build gives the
    batched outputs too (density, tangents, cells, occupancy, image)."""
    batch = Batch(batch32, W, H, tuning=load_tuning("batch"))
    batch.launch()
    torch.cuda.synchronize()
    for b in batch32[::3]:
        out = _plain_outputs(b)   # buffers shaped like the style's own
        for k in ("counts", "density", "tangents", "cells", "occupancy", "rgba"):
            out[k].zero_()
        b.run_plain(out, one_call=True)
        for k, mine in [("counts", b.counts), ("density", b.density), ("tangents", b.tangents), ("cells", b.cells),
                        ("occupancy", b.occupancy), ("rgba", b.rgba)]:
            assert torch.equal(mine, out[k]), (k, b.n)


def test_batch_volumes_and_images_match_the_oracle(batch32):
    """The batch's outputs against the oracle directly, at the batch's own
    shapes, for the first style of each volume size (64^3, 128^3, 256^3):
    its counts equal the oracle's voxelization of the style, its uint8
    density the oracle's normalize of them, its tangents the reference's
    quantization within 1 LSB (no voxel of these styles saturates); and at
    960x540, 512 steps, the per-style raymarch of the batch's volume (with
    the depth and iso-decision outputs the parity check reads) equals the
    batch's image bit for bit, which the float64 oracle, run on the batch's
    volumes, checks at 6,000 random pixels plus 2,000 covered ones."""
    from oracle import voxel as OV
    from tests.parity import compare_raymarch
    from vkhr_amd import device as D
    styles, _ = synth.batch_styles()
    batch = Batch(batch32, W, H, tuning=load_tuning("batch"))
    batch.launch()
    torch.cuda.synchronize()
    rng = np.random.default_rng(11)
    for n in (64, 128, 256):
        i = next(k for k, x in enumerate(batch32) if x.n == n)
        b, st = batch32[i], styles[i]
        res = (n, n, n)
        ref = OV.voxelize(st.vertices, OV.aabb(st.vertices), res, indices=st.indices(),
                          tangents=OV.generate_tangents(st.vertices, st.strand_offsets), want_sums=True)
        assert int(ref["counts"].max()) <= 255
        assert np.array_equal(b.counts.cpu().numpy().view(np.uint32), ref["counts"]), n
        assert np.array_equal(b.density.cpu().numpy(), OV.normalize(ref["dens_u8"])), n
        q_ref = OV.quantize_tangents(ref["sums_ref"], ref["dens_u8"].astype(np.uint32)).reshape(-1, 4)
        q = b.tangents.cpu().numpy().reshape(-1, 4)
        assert np.abs(q.astype(np.int32) - q_ref.astype(np.int32)).max() <= 1, n
        vol = dict(density=b.density, tangents=b.tangents, cells=b.cells, occupancy=b.occupancy)
        g = D.raymarch(vol, res, b.cam, b.light, b.sp, b.rp, W, H, want_depth=True, want_info=True)
        assert torch.equal(g["rgba"], b.rgba), n
        g = {k: (v.cpu().numpy() if v is not None else None) for k, v in g.items()}
        covered = np.nonzero(g["info"].reshape(-1, 4)[:, 0] == 1)[0]
        assert len(covered) > 2000, (n, len(covered))
        pix = np.unique(np.concatenate([rng.integers(0, W * H, 6000), rng.choice(covered, 2000, replace=False)]))
        vol_np = dict(density=b.density.cpu().numpy(), tangents=b.tangents.cpu().numpy())
        rep = compare_raymarch(g, vol_np, res, b.cam, b.light, b.sp, b.rp, W, H, pixels=pix)
        print(n, len(st.strand_offsets) - 1, "strands", rep)


def _long_segment_style(n_pairs, seed):
    """Strands of one segment between random points of the box: each visits
    ~16 bricks at 128^3, so 20,000 of them pass the bin capacity (4 runs per
    pair + 65536) and the build takes the workspace path."""
    rng = np.random.default_rng(seed)
    verts = rng.uniform(-1.0, 1.0, size=(2 * n_pairs, 3)).astype(np.float32)
    return synth.Style(vertices=verts, segments=np.ones(n_pairs, np.uint16),
                       thickness=np.ones(2 * n_pairs, np.float32), seed=seed)


@pytest.mark.parametrize("groups", [1, 3])
def test_batch_counts_only_and_overflow_styles(device, groups):
    """Styles the 32-style batch does not hold: one with counts only (no
    tangent workspace or tangent volume: the counts are the overflow
    accumulator, shaded with the ADSM-only model), one whose runs overflow
    the bins (the workspace path, with tangents), one with no segments, and
    a plain one, in one batch; every output equals the per-style sequence."""
    from vkhr_amd import params
    styles = [(synth.hemisphere_style(3000, 16, seed=5), 64, False), (_long_segment_style(20000, 3), 128, True),
              (synth.hemisphere_style(50, 4, seed=7), 64, True), (synth.hemisphere_style(2000, 12, seed=6), 64, True)]
    bufs = []
    for i, (st, n, tangents) in enumerate(styles):
        b = StyleBuffers(st, n, W, H, device, azimuth=30.0 * i)
        if not tangents:
            b.acc, b.tangents = None, None
            b.rp = params.render(raycast_steps=512.0, shading_model=1)
        bufs.append(b)
    bufs[2].idx = bufs[2].idx[:0]
    t = capi.Tuning()
    t.tile_w, t.tile_h, t.blocks_per_cu, t.groups = 8, 8, 0, groups
    batch = Batch(bufs, W, H, tuning=t)
    batch.launch()
    batch.launch()
    torch.cuda.synchronize()
    for b in bufs:
        ref = _plain_outputs(b)
        checks = [("counts", b.counts), ("density", b.density), ("cells", b.cells), ("occupancy", b.occupancy),
                  ("rgba", b.rgba)] + ([("tangents", b.tangents)] if b.tangents is not None else [])
        for k, mine in checks:
            assert torch.equal(mine, ref[k]), (k, b.n)
    assert int((bufs[1].counts > 0).sum()) > 100_000


def test_invalid_batch_arguments(batch32):
    bad = capi.Tuning()
    bad.tile_w, bad.tile_h, bad.blocks_per_cu = 8, 4, 4    # 32 pixels per wave
    with pytest.raises(capi.VkhrError):
        Batch(batch32[:2], W, H, tuning=bad)
    b = Batch(batch32[:2], W, H)
    import ctypes as C
    assert capi.lib.vkhr_batch_launch(b.handle, None) == -1   # graph capture needs a created stream
