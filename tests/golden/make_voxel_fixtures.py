"""Writes tests/golden/voxel_small.npz: small seeded styles (<= 1k segments,
16^3-32^3 grids) with the voxel counts of the voxelization oracle, so that a
later change to the oracle or to the generator shows up as a fixture diff.

    python tests/golden/make_voxel_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import voxel as OV  # noqa: E402
from vkhr_amd import synth  # noqa: E402

CASES = [  # (strands, vertices per strand, seed, grid)
    (40, 16, 101, (16, 16, 16)),
    (90, (2, 12), 102, (32, 32, 32)),
    (64, 5, 103, (24, 16, 32)),
]


def main():
    out = {"n_cases": np.int64(len(CASES))}
    for k, (n, vps, seed, res) in enumerate(CASES):
        s = synth.hemisphere_style(n, vps, seed)
        b = OV.aabb(s.vertices)
        idx = s.indices()
        assert len(idx) // 2 <= 1000
        out[f"xyz{k}"] = s.vertices
        out[f"idx{k}"] = idx
        out[f"res{k}"] = np.array(res, np.int64)
        out[f"aabb{k}"] = OV.aabb_array(b)
        out[f"seg{k}"] = OV.voxelize(s.vertices, b, res, indices=idx)["counts"]
        out[f"vtx{k}"] = OV.voxelize(s.vertices, b, res)["counts"]
    np.savez_compressed(Path(__file__).resolve().parent / "voxel_small.npz", **out)


if __name__ == "__main__":
    main()
