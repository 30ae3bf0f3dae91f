"""GPU parity of the voxelization path against the strict-IEEE oracle:
counts bit-exact (uint32 and the reference's saturating uint8 view), AABB,
indices and tangents bit-exact, tangent volumes within 1 LSB on unsaturated
voxels, normalize and the corner-packed layout exact."""
import math

import numpy as np
import pytest
import torch

from oracle import voxel as OV
from vkhr_amd import capi, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu


def _check_counts(style, res, device, segments=True):
    xyz = D.to_dev(style.vertices, device)
    b = OV.aabb(style.vertices)
    idx = style.indices()
    tan_ref = OV.generate_tangents(style.vertices, style.strand_offsets)
    ref = OV.voxelize(style.vertices, b, res, indices=idx if segments else None, tangents=tan_ref,
                      want_sums=True)
    counts, sums = D.voxelize(xyz, b, res, indices=D.to_dev(idx, device) if segments else None,
                              tangents=D.to_dev(tan_ref, device), want_sums=True)
    got = counts.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, ref["counts"]), f"{(got != ref['counts']).sum()} voxels differ"
    assert np.array_equal(np.minimum(got, 255).astype(np.uint8), ref["dens_u8"])
    return ref, counts, sums


def test_prep_bit_exact(device):
    s = synth.strands_10k()
    xyz = D.to_dev(s.vertices, device)
    off = D.to_dev(s.strand_offsets, device)
    b = D.aabb(xyz)
    r = OV.aabb(s.vertices)
    for k in ("origin", "size"):
        assert np.array_equal(b[k], r[k]), k
    assert b["radius"] == r["radius"] and b["volume"] == r["volume"]
    assert np.array_equal(D.generate_indices(off).cpu().numpy().view(np.uint32), OV.generate_indices(s.strand_offsets))
    t = D.generate_tangents(xyz, off).cpu().numpy()
    assert np.array_equal(t.view(np.uint32), OV.generate_tangents(s.vertices, s.strand_offsets).view(np.uint32))


def test_prep_uneven_strands(device):
    s = synth.hemisphere_style(3000, (1, 32), seed=9)   # includes single-vertex strands
    xyz = D.to_dev(s.vertices, device)
    off = D.to_dev(s.strand_offsets, device)
    assert np.array_equal(D.generate_indices(off).cpu().numpy().view(np.uint32), s.indices())
    t = D.generate_tangents(xyz, off).cpu().numpy()
    assert np.array_equal(t.view(np.uint32), OV.generate_tangents(s.vertices, s.strand_offsets).view(np.uint32))


@pytest.mark.parametrize("segments", [False, True])
def test_10k_strands_128(device, segments):
    _check_counts(synth.strands_10k(), (128, 128, 128), device, segments)


def test_136k_strands_segments_256(device):
    _check_counts(synth.strands_136k(), (256, 256, 256), device, True)


def test_golden_fixtures(device):
    from pathlib import Path
    z = np.load(Path(__file__).resolve().parent / "golden" / "voxel_small.npz")
    for k in range(int(z["n_cases"])):
        xyz, idx, res = z[f"xyz{k}"], z[f"idx{k}"], tuple(int(v) for v in z[f"res{k}"])
        b = OV.aabb(xyz)
        c, _ = D.voxelize(D.to_dev(xyz, device), b, res, indices=D.to_dev(idx, device))
        assert np.array_equal(c.cpu().numpy().view(np.uint32), z[f"seg{k}"])
        c, _ = D.voxelize(D.to_dev(xyz, device), b, res)
        assert np.array_equal(c.cpu().numpy().view(np.uint32), z[f"vtx{k}"])


@pytest.mark.parametrize("edge", [64, 128, 256])
def test_batch_generator_styles(device, edge):
    styles, _ = synth.batch_styles(n_styles=6, seed=4, lo=2000, hi=20000)
    for s in styles[:3]:
        _check_counts(s, (edge, edge, edge), device, True)


def test_saturation_and_degenerate_segments(device):
    unit = dict(origin=np.zeros(3, np.float32), size=np.full(3, 4.0, np.float32),
                radius=np.float32(math.sqrt(48.0)), volume=np.float32(64.0))
    pts = np.tile(np.array([[1.5, 1.5, 1.5], [1.5, 1.5, 3.5]], np.float32), (400, 1))
    idx = np.arange(len(pts), dtype=np.uint32)
    pts[-2:] = [[2.5, 2.5, 2.5], [2.5, 2.5, 2.5]]    # zero-length segment
    ref = OV.voxelize(pts, unit, (4, 4, 4), indices=idx)
    c, _ = D.voxelize(D.to_dev(pts, device), unit, (4, 4, 4), indices=D.to_dev(idx, device))
    got = c.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, ref["counts"]) and got.max() > 255


def test_empty_inputs(device):
    unit = dict(origin=np.zeros(3, np.float32), size=np.ones(3, np.float32), radius=np.float32(1.7),
                volume=np.float32(1))
    xyz = torch.zeros((1, 3), dtype=torch.float32, device=device)
    idx = torch.zeros(0, dtype=torch.int32, device=device)
    c, _ = D.voxelize(xyz, unit, (8, 8, 8), indices=idx)
    assert int(c.sum()) == 0
    b = D.aabb(torch.zeros((0, 3), dtype=torch.float32, device=device))
    assert b["size"].tolist() == [0, 0, 0]


def test_invalid_arguments_are_errors(device):
    xyz = torch.zeros((4, 3), dtype=torch.float32, device=device)
    flat = dict(origin=np.zeros(3, np.float32), size=np.array([1, 0, 1], np.float32), radius=np.float32(1.4),
                volume=np.float32(0))
    with pytest.raises(capi.VkhrError) as e:
        D.voxelize(xyz, flat, (8, 8, 8))
    assert e.value.code == -1
    ok = dict(flat, size=np.ones(3, np.float32))
    with pytest.raises(capi.VkhrError) as e:
        D.voxelize(xyz, ok, (512, 256, 256), counts=torch.empty(1, dtype=torch.int32, device=device))
    assert e.value.code == -3
    bad_idx = D.to_dev(np.array([0, 99, 1, 2], np.uint32), device)   # pair (0, 99) is out of range
    c, _ = D.voxelize(xyz + 0.5, ok, (2, 2, 2), indices=bad_idx)
    assert int(c.sum()) == 0  # (0,99) skipped; (1,2) has zero length


def test_finalize_matches_oracle(device):
    s = synth.strands_10k()
    res = (128, 128, 128)
    ref, counts, sums = _check_counts(s, res, device, True)
    vol = D.finalize(counts, res, sums=sums, normalize=True, want_f32=True, want_cells=True)
    u8 = ref["dens_u8"]
    assert np.array_equal(vol["density"].cpu().numpy(), OV.normalize(u8))
    assert np.array_equal(vol["density_f32"].cpu().numpy(), OV.normalize(u8).astype(np.float32) / np.float32(255))
    # tangents: unsaturated voxels within 1 LSB of the reference's quantization
    q_ref = OV.quantize_tangents(ref["sums_ref"], u8.astype(np.uint32)).reshape(-1, 4).astype(np.int32)
    q = vol["tangents"].cpu().numpy().reshape(-1, 4).astype(np.int32)
    unsat = ref["counts"] <= 255
    assert np.abs(q[unsat] - q_ref[unsat]).max() <= 1
    assert (q[:, 3] == 0).all()
    # corner-packed cells from the normalized density
    d = OV.normalize(u8).reshape(128, 128, 128)   # [z, y, x]
    pad = np.zeros((130, 130, 130), np.uint8)
    pad[1:129, 1:129, 1:129] = d
    cells = vol["cells"].cpu().numpy().view(np.uint64).reshape(129, 129, 129)
    want = np.zeros_like(cells)
    for b in range(8):
        dx, dy, dz = b & 1, (b >> 1) & 1, b >> 2
        want |= pad[dz:dz + 129, dy:dy + 129, dx:dx + 129].astype(np.uint64) << np.uint64(8 * b)
    assert np.array_equal(cells, want)


def test_finalize_without_normalize(device):
    s = synth.strands_10k()
    res = (64, 64, 64)
    ref, counts, _ = _check_counts(s, res, device, True)
    vol = D.finalize(counts, res, normalize=False, want_cells=False)
    assert np.array_equal(vol["density"].cpu().numpy(), ref["dens_u8"])


def test_packed_counts_path_matches(device):
    """With the tangent accumulator and no count volume the counts stay
    packed in the accumulator's second word; finalize unpacks them and gives
    the same volumes as from the unpacked counts (cells also without the
    normalize)."""
    s = synth.strands_10k()
    res = (128, 128, 128)
    ref, counts, sums = _check_counts(s, res, device, True)
    xyz = D.to_dev(s.vertices, device)
    off = D.to_dev(s.strand_offsets, device)
    idx = D.generate_indices(off)
    tan = D.generate_tangents(xyz, off)
    none, packed = D.voxelize(xyz, D.aabb(xyz), res, indices=idx, tangents=tan, want_sums=True, want_counts=False)
    assert none is None
    unpacked = (packed.view(-1, 2)[:, 1] & 0xFFFFFFFF).cpu().numpy().astype(np.uint32)
    assert np.array_equal(unpacked, ref["counts"])
    for normalize in (True, False):
        a = D.finalize(counts, res, sums=sums, normalize=normalize, want_f32=True, want_cells=True)
        b = D.finalize(None, res, sums=packed, normalize=normalize, want_f32=True, want_cells=True)
        for k in ("density", "tangents", "density_f32", "cells"):
            assert torch.equal(a[k], b[k]), (normalize, k)
    # the cells of the unnormalized view hold the saturating uint8 counts
    c = D.finalize(None, res, sums=packed, normalize=False, want_cells=True)["cells"].cpu().numpy().view(np.uint64)
    u8 = ref["dens_u8"].reshape(128, 128, 128)
    assert int(c.reshape(129, 129, 129)[5, 7, 9] & np.uint64(0xFF)) == int(u8[4, 6, 8])


def _occupancy_from_cells(cells: np.ndarray, res) -> np.ndarray:
    """The bitmask and distance nibbles of vkhr_volume_finalize's occupancy
    buffer (its working bytes after them are not part of the result): one
    bit per brick of 8x8x8 cells, set when a cell of the brick is nonzero,
    then 4 bits per brick, the chessboard distance in bricks to the nearest
    occupied brick capped at 15 (0 for an occupied one).  The distances come
    from scipy's chamfer transform, not the kernels' separable passes."""
    from scipy import ndimage
    nx, ny, nz = res
    c = cells.view(np.uint64).reshape(nz + 1, ny + 1, nx + 1) != 0
    bx, by, bz = (nx + 8) // 8, (ny + 8) // 8, (nz + 8) // 8
    pad = np.zeros((bz * 8, by * 8, bx * 8), bool)
    pad[:nz + 1, :ny + 1, :nx + 1] = c
    occupied = pad.reshape(bz, 8, by, 8, bx, 8).any(axis=(1, 3, 5))
    bricks = occupied.reshape(-1)
    words = np.zeros((bricks.size + 31) // 32, np.uint32)
    idx = np.nonzero(bricks)[0]
    np.bitwise_or.at(words, idx >> 5, (np.uint32(1) << (idx & 31).astype(np.uint32)))
    if occupied.any():
        dist = ndimage.distance_transform_cdt(~occupied, metric="chessboard").reshape(-1)
    else:
        dist = np.full(bricks.size, 15)
    dist = np.minimum(dist, 15).astype(np.uint32)
    nib = np.zeros((bricks.size + 7) // 8, np.uint32)
    b = np.arange(bricks.size)
    np.bitwise_or.at(nib, b >> 3, dist << (4 * (b & 7)).astype(np.uint32))
    return np.concatenate([words, nib])


def _cells_from_density(dens: np.ndarray, res) -> np.ndarray:
    """The corner-packed cells from the uint8 density: cell (i, j, k), i in
    [-1, nx-1], byte b = texel (i + (b & 1), j + (b >> 1 & 1), k + (b >> 2)),
    0 outside the volume."""
    nx, ny, nz = res
    pad = np.zeros((nz + 2, ny + 2, nx + 2), np.uint8)
    pad[1:nz + 1, 1:ny + 1, 1:nx + 1] = dens.reshape(nz, ny, nx)
    want = np.zeros((nz + 1, ny + 1, nx + 1), np.uint64)
    for b in range(8):
        dx, dy, dz = b & 1, (b >> 1) & 1, b >> 2
        want |= pad[dz:dz + nz + 1, dy:dy + ny + 1, dx:dx + nx + 1].astype(np.uint64) << np.uint64(8 * b)
    return want.reshape(-1)


@pytest.mark.parametrize("res", [(40, 72, 9), (37, 20, 65), (64, 8, 200)])
def test_cells_exact_at_uneven_sizes(device, res):
    """Cells and occupancy at uneven sizes: partial 64-column groups (x 40,
    37, 64 + the border column), partial brick rows, thin and deep volumes
    (several cell slabs along z), x sizes that are and are not multiples
    of 4."""
    s = synth.strands_10k()
    _, counts, sums = _check_counts(s, res, device, True)
    vol = D.finalize(counts, res, sums=sums, normalize=True, want_cells=True)
    cells = vol["cells"].cpu().numpy().view(np.uint64)
    want = _cells_from_density(vol["density"].cpu().numpy(), res)
    assert np.array_equal(cells, want)
    occ = _occupancy_from_cells(want, res)
    assert np.array_equal(vol["occupancy"].cpu().numpy().view(np.uint32)[:occ.size], occ)
    assert (want != 0).sum() > 100


@pytest.mark.parametrize("res", [(128, 128, 128), (40, 72, 9)])
def test_occupancy_bits(device, res):
    s = synth.strands_10k()
    _, counts, sums = _check_counts(s, res, device, True)
    for normalize in (True, False):
        vol = D.finalize(counts, res, sums=sums, normalize=normalize, want_cells=True)
        got = vol["occupancy"].cpu().numpy().view(np.uint32)
        assert got.size == D.occupancy_words(res)
        want = _occupancy_from_cells(vol["cells"].cpu().numpy(), res)
        assert np.array_equal(got[:want.size], want)
        n_bricks = ((res[0] + 8) // 8) * ((res[1] + 8) // 8) * ((res[2] + 8) // 8)
        bits = want[:(n_bricks + 31) // 32]
        assert 0 < int(np.unpackbits(bits.view(np.uint8)).sum()) < n_bricks
        # empty bricks at several distances (the 17^3-brick volume reaches 4)
        nib = want[(n_bricks + 31) // 32:]
        d = (nib[np.arange(n_bricks) >> 3] >> (4 * (np.arange(n_bricks) & 7)).astype(np.uint32)) & 15
        assert d.max() >= (4 if res[0] == 128 else 2), np.unique(d)


def test_finalize_on_concurrent_streams(device):
    """Two finalizes (each with its min/max reduction and normalize) in flight
    on two streams at once give the volumes of serial calls: each stream has
    its own scratch area."""
    a_style, b_style = synth.strands_10k(), synth.hemisphere_style(300, 20, seed=17)
    res = (128, 128, 128)
    ra, ca, sa = _check_counts(a_style, res, device, True)
    rb, cb, sb = _check_counts(b_style, res, device, True)
    # different maxima: a shared min/max would normalize one of them wrongly
    assert ra["dens_u8"].max() != rb["dens_u8"].max()
    want_a = D.finalize(ca, res, sums=sa, normalize=True, want_cells=True)
    want_b = D.finalize(cb, res, sums=sb, normalize=True, want_cells=True)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        with torch.cuda.stream(s1):
            got_a = D.finalize(ca, res, sums=sa, normalize=True, want_cells=True)
        with torch.cuda.stream(s2):
            got_b = D.finalize(cb, res, sums=sb, normalize=True, want_cells=True)
        torch.cuda.synchronize()
        for want, got in ((want_a, got_a), (want_b, got_b)):
            for k in ("density", "tangents", "cells", "occupancy"):
                assert torch.equal(want[k], got[k]), k
