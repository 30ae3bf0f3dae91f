"""The .hair file side of the host HairStyle (vkhr_amd/host/hair_file.cc,
driven by vkhr_amd/hair_tool; CPU only) against the oracle restatement of
the format (oracle/hair_format.py):
  * load + save reproduces the file byte for byte (all optional arrays,
    header defaults, information, bounding box);
  * the size identity 128 + 12 V + 12 V + 8 S of the reference's ponytail
    asset (V = 1,772,160 vertices, S = 1,635,840 segments -> 55,618,688 bytes,
    the size in its Git-LFS pointer), a file without a segment array;
  * shuffle / reduce against the restated xorshift64 draw loop (same strand
    order, the reference's strand count for ratio < 1);
  * the error state of every failing read;
  * generate_bounding_box (origin-started min / max) and get_bounding_box."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import hair_format as HF
from oracle import voxel as OV
from vkhr_amd import synth

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "vkhr_amd" / "hair_tool"

ERR = dict(none=0, opening=1, reading_header=2, signature=4, reading_segments=5, reading_vertices=6,
           reading_thickness=7, reading_tangents=10, reading_indices=11, invalid_format=19)


def tool(*args):
    p = subprocess.run([str(TOOL), *map(str, args)], capture_output=True, text=True, timeout=120)
    return p.returncode, json.loads(p.stdout.strip().splitlines()[-1])


def small_style(n_strands=200, seed=7):
    rng = np.random.default_rng(seed)
    segs = rng.integers(1, 20, n_strands).astype(np.uint16)
    v = int(segs.sum()) + n_strands
    arrays = dict(segments=segs, vertices=rng.normal(size=(v, 3)).astype(np.float32),
                  thickness=rng.uniform(0.5, 1.0, v).astype(np.float32),
                  transparency=rng.uniform(0, 1, v).astype(np.float32),
                  color=rng.uniform(0, 1, (v, 3)).astype(np.float32),
                  tangents=rng.normal(size=(v, 3)).astype(np.float32))
    idx, k = [], 0
    for s in segs:
        for _ in range(int(s)):
            idx += [k, k + 1]
            k += 1
        k += 1
    arrays["indices"] = np.array(idx, np.uint32)
    return arrays


def test_xorshift64_known_answers():
    """Marsaglia's xorshift64 with shifts (13, 7, 17), worked by hand: from 1,
    x ^= x << 13 gives 8193, x ^= x >> 7 gives 8257, x ^= x << 17 gives
    8257 + 8257 * 2^17 = 1082269761; and a zero state stays zero."""
    assert HF.xorshift64(1) == 1082269761
    assert HF.xorshift64(0) == 0
    # the state is 64 bits wide: from 2^63 both left shifts fall off the top
    # and the right shift adds 2^56
    assert HF.xorshift64(1 << 63) == (1 << 63) | (1 << 56)


def test_roundtrip_is_byte_identical(tmp_path):
    arrays = small_style()
    data = HF.write(arrays, 200, default_segment_count=0, default_thickness=0.042, default_transparency=0.3,
                    default_color=(0.32, 0.228, 0.128), information=b"synthetic style", bbox=((-3, -3, -3), (3, 3, 3)))
    src, dst = tmp_path / "a.hair", tmp_path / "b.hair"
    src.write_bytes(data)
    rc, info = tool("copy", src, dst)
    assert rc == 0 and info["ok"] and info["error"] == 0
    assert dst.read_bytes() == data
    assert info["size"] == len(data)
    back = HF.read(dst.read_bytes())
    for k, a in arrays.items():
        assert np.array_equal(back[k].reshape(a.shape), a), k
    assert info["has_bounding_box"] and np.array_equal(np.float32(info["bbox_max"]), np.float32([3, 3, 3]))


def test_ponytail_size_identity(tmp_path):
    s = synth.strands_136k()
    assert s.n_strands == 136_320 and s.n_vertices == 1_772_160
    verts = s.vertices.reshape(-1, 3)
    offs = s.strand_offsets.astype(np.uint32)
    tan = OV.generate_tangents(s.vertices, offs).reshape(-1, 3)
    idx = OV.generate_indices(offs)
    data = HF.write(dict(vertices=verts, tangents=tan, indices=idx), 136_320, default_segment_count=12)
    assert len(data) == 128 + 12 * 1_772_160 + 12 * 1_772_160 + 8 * 1_635_840 == 55_618_688
    src, dst = tmp_path / "ponytail_like.hair", tmp_path / "copy.hair"
    src.write_bytes(data)
    rc, info = tool("copy", src, dst)
    assert rc == 0 and info["strands"] == 136_320 and info["vertices"] == 1_772_160
    assert info["segments"] == 1_635_840 and info["size"] == 55_618_688
    assert dst.read_bytes() == data


@pytest.mark.parametrize("with_segments", [True, False])
def test_shuffle_matches_oracle(tmp_path, with_segments):
    arrays = small_style(150, seed=3)
    if not with_segments:  # uniform strands announced by the default segment count
        arrays = small_style(150, seed=3)
        segs = np.full(150, 9, np.uint16)
        v = 150 * 10
        arrays = {k: (a[:v] if k not in ("segments", "indices") else a) for k, a in arrays.items()}
        arrays.pop("segments")
        arrays.pop("indices")
    src, dst = tmp_path / "a.hair", tmp_path / "b.hair"
    src.write_bytes(HF.write(arrays, 150, default_segment_count=9))
    seed = 0x9E3779B97F4A7C15
    rc, info = tool("shuffle", src, dst, seed)
    assert rc == 0, info
    want, _ = HF.reduce(arrays, 150, 9, 1.0, seed)
    got = HF.read(dst.read_bytes())
    assert len(got["segments"]) == 150
    for k in ("segments", "vertices", "thickness", "tangents", "transparency", "color", "indices"):
        if k in want:
            assert np.array_equal(got[k].reshape(want[k].shape), want[k]), k
    # a permutation of whole strands
    assert sorted(map(tuple, got["vertices"].reshape(-1, 3).tolist())) == \
        sorted(map(tuple, arrays["vertices"].reshape(-1, 3).tolist()))


@pytest.mark.parametrize("ratio", [0.5, 0.9])
def test_reduce_keeps_the_reference_strand_count(tmp_path, ratio):
    arrays = small_style(100, seed=5)
    src, dst = tmp_path / "a.hair", tmp_path / "b.hair"
    src.write_bytes(HF.write(arrays, 100))
    rc, info = tool("reduce", src, dst, 42, ratio)
    assert rc == 0
    want, _ = HF.reduce(arrays, 100, 0, ratio, 42)
    # (100 - ceil(100 r)) - 1 strands: the reference's pre-decremented loop count
    assert info["strands"] == len(want["segments"]) == (100 - int(np.ceil(100 * np.float32(ratio)))) - 1
    got = HF.read(dst.read_bytes())
    assert np.array_equal(got["vertices"].reshape(-1, 3), want["vertices"])


def test_error_states(tmp_path):
    rc, info = tool("info", tmp_path / "missing.hair")
    assert rc == 1 and info["error"] == ERR["opening"]
    good = HF.write(small_style(20), 20)
    cases = {
        "reading_header": good[:100],
        "signature": b"HAIX" + good[4:],
        "reading_segments": good[:128 + 10],
        "reading_vertices": good[:128 + 40 + 100],
        "reading_indices": good[:-4],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.hair"
        p.write_bytes(data)
        rc, info = tool("info", p)
        assert rc == 1 and info["error"] == ERR[name], (name, info)
    # a file without vertices is not a valid style
    p = tmp_path / "empty.hair"
    p.write_bytes(HF.write(dict(vertices=np.zeros((0, 3), np.float32)), 0))
    rc, info = tool("info", p)
    assert rc == 1 and info["error"] == ERR["invalid_format"]
    # fewer vertices than strands leaves no index count
    p = tmp_path / "bad_counts.hair"
    data = bytearray(HF.write(dict(vertices=np.zeros((4, 3), np.float32), indices=np.zeros(2, np.uint32)), 2))
    data[4:8] = (9).to_bytes(4, "little")  # strand count 9 > 4 vertices
    p.write_bytes(bytes(data))
    rc, info = tool("info", p)
    assert rc == 1 and info["error"] == ERR["reading_indices"]


def test_bounding_box_in_the_header(tmp_path):
    arrays = small_style(50, seed=11)
    arrays["vertices"] += np.float32(5.0)  # entirely on the positive side: the origin still in the box
    src, dst = tmp_path / "a.hair", tmp_path / "b.hair"
    src.write_bytes(HF.write(arrays, 50))
    rc, info = tool("bbox", src, dst)
    assert rc == 0 and info["has_bounding_box"]
    f32 = lambda x: np.asarray(x, np.float32)  # noqa: E731  (%.9g round-trips float32)
    lo, hi = HF.bounding_box(arrays["vertices"])
    assert np.array_equal(f32(info["bbox_min"]), f32(lo)) and np.array_equal(f32(info["bbox_max"]), f32(hi))
    assert lo == (0.0, 0.0, 0.0)
    ref = OV.aabb(arrays["vertices"])
    assert np.array_equal(f32(info["aabb"]["origin"]), f32(ref["origin"]))
    assert np.array_equal(f32(info["aabb"]["size"]), f32(ref["size"]))
    assert f32(info["aabb"]["radius"]) == f32(ref["radius"])
    assert f32(info["aabb"]["volume"]) == f32(ref["volume"])
    back = HF.read(dst.read_bytes())
    assert back["field"] >> 7 & 1
