"""Oracle (test infrastructure only) of the .hair file side: an independent
numpy/struct restatement of the format and of the strand shuffle.

Format (include/vkhr/scene_graph/hair_style.hh:147-174; field order
src/vkhr/scene_graph/hair_style.cc:24-70, 576-665): a 128-byte header --
"HAIR", strand count, vertex count (uint32), a 32-bit field word (bit 0
segments, 1 vertices, 2 thickness, 3 transparency, 4 colour, 5 tangents,
6 indices, 7 bounding box: the declaration order of the reference's bit
field, allocated from the least significant bit), default segment count
(uint32), default thickness and transparency (float), default colour
(3 float), 64 bytes of information, bounding-box min and max (3 float
each) -- then the arrays present: segments (uint16 per strand), vertices
(3 float), thickness (float), transparency (float), colour (3 float),
tangents (3 float) per vertex, and 2 (vertices - strands) uint32 indices.

Shuffle / reduce (hair_style.cc:371-449) with xorshift64 (:679-685),
restated in Python integers."""
from __future__ import annotations

import math
import struct

import numpy as np

HEADER = struct.Struct("<4sIIIIff3f64s3f3f")
assert HEADER.size == 128

BITS = dict(segments=0, vertices=1, thickness=2, transparency=3, color=4, tangents=5, indices=6, bounding_box=7)
ORDER = [("segments", np.uint16, 1), ("vertices", np.float32, 3), ("thickness", np.float32, 1),
         ("transparency", np.float32, 1), ("color", np.float32, 3), ("tangents", np.float32, 3),
         ("indices", np.uint32, 1)]


def write(arrays: dict, strand_count: int, default_segment_count: int = 0, default_thickness: float = 0.0,
          default_transparency: float = 0.0, default_color=(0.0, 0.0, 0.0), information: bytes = b"",
          bbox=None) -> bytes:
    """The bytes of a .hair file holding `arrays` (name -> numpy array)."""
    field = 0
    for name, _, _ in ORDER:
        if name in arrays and len(arrays[name]):
            field |= 1 << BITS[name]
    lo, hi = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
    if bbox is not None:
        field |= 1 << BITS["bounding_box"]
        lo, hi = bbox
    vertex_count = len(arrays["vertices"])
    head = HEADER.pack(b"HAIR", strand_count, vertex_count, field, default_segment_count, default_thickness,
                       default_transparency, *default_color, information.ljust(64, b"\0")[:64], *lo, *hi)
    body = b"".join(np.ascontiguousarray(arrays[n], dtype=t).tobytes()
                    for n, t, _ in ORDER if n in arrays and len(arrays[n]))
    return head + body


def read(data: bytes) -> dict:
    """Header fields and arrays of a .hair file's bytes."""
    sig, strands, vertices, field, dsc, dth, dtr, c0, c1, c2, info, x0, y0, z0, x1, y1, z1 = HEADER.unpack_from(data)
    out = dict(signature=sig, strand_count=strands, vertex_count=vertices, field=field,
               default_segment_count=dsc, bbox=((x0, y0, z0), (x1, y1, z1)))
    off = HEADER.size
    for name, t, w in ORDER:
        if not field >> BITS[name] & 1:
            continue
        if name == "segments":
            n = strands
        elif name == "indices":
            n = 2 * (vertices - (strands if not field & 1 else len(out["segments"])))
        else:
            n = vertices * w
        a = np.frombuffer(data, dtype=t, count=n, offset=off)
        off += a.nbytes
        out[name] = a.reshape(-1, w) if w > 1 else a
    return out


def xorshift64(state: int) -> int:
    m = (1 << 64) - 1
    x = state
    x ^= (x << 13) & m
    x ^= x >> 7
    x ^= (x << 17) & m
    return x


def reduce(arrays: dict, strand_count: int, default_segment_count: int, ratio: float, seed: int):
    """HairStyle::reduce: returns (new arrays, final seed).  Strands are drawn
    at random (swap with the last, pop); the count keeps the reference's
    loop arithmetic (ratio 1 -> every strand)."""
    f32 = np.float32
    n = strand_count
    left = int(f32(n) - f32(math.ceil(f32(n) * f32(ratio))))  # unsigned - float, truncated to unsigned
    has_seg = "segments" in arrays and len(arrays["segments"])
    segs = [int(s) for s in arrays["segments"]] if has_seg else None
    counts = segs if has_seg else [default_segment_count] * n
    offs = [0] * n
    for i in range(1, n):
        offs[i] = offs[i - 1] + counts[i - 1] + 1
    out = {k: [] for k in ("segments", "vertices", "thickness", "tangents", "transparency", "color")}
    left = (left - 1) & 0xFFFFFFFF  # --strands_left on an unsigned
    while left and offs:
        seed = xorshift64(seed)
        r = float(seed) / float((1 << 64) - 1)
        pick = min(int(r * float(len(offs))), len(offs) - 1)
        count = default_segment_count
        if has_seg:
            count = segs[pick]
            segs[pick], segs[-1] = segs[-1], segs[pick]
            segs.pop()
        out["segments"].append(count)
        a = offs[pick]
        for k in ("vertices", "thickness", "tangents", "transparency", "color"):
            if k in arrays and len(arrays[k]):
                out[k].append(np.asarray(arrays[k])[a:a + count + 1])
        offs[pick], offs[-1] = offs[-1], offs[pick]
        offs.pop()
        left = (left - 1) & 0xFFFFFFFF
    res = dict(segments=np.array(out["segments"], np.uint16))
    for k in ("vertices", "thickness", "tangents", "transparency", "color"):
        if out[k]:
            res[k] = np.concatenate(out[k])
    idx, v = [], 0
    for c in res["segments"]:
        for _ in range(int(c)):
            idx += [v, v + 1]
            v += 1
        v += 1
    res["indices"] = np.array(idx, np.uint32)
    return res, seed


def bounding_box(vertices: np.ndarray):
    """generate_bounding_box: min / max started at the origin (float32)."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    lo = np.minimum(v.min(axis=0), np.float32(0)) if len(v) else np.zeros(3, np.float32)
    hi = np.maximum(v.max(axis=0), np.float32(0)) if len(v) else np.zeros(3, np.float32)
    return tuple(float(x) for x in lo), tuple(float(x) for x in hi)
