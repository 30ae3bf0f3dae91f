"""Benchmark of the hybrid-hair hot path on one MI355X.

Workload (N=1): 136,320 strands x 13 vertices (1,635,840 segments), segment
voxelization into 256^3 uint32 counts + finalize (saturate, normalize,
corner-packed cells), then the volume raymarch at 1920x1080 with 512 steps,
ADSM and AO.  Inputs are resident in HBM before the timed region; a "step"
is one pass: voxelize -> finalize -> raymarch.

Prints one JSON line (see the keys below).  --gpus N>1 is not supported: the
path runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline_voxelize(style, bounds, res, budget_s=12.0):
    """The strict-IEEE restatement of HairStyle::voxelize_segments, single
    thread (as the reference's DDA loop), on a bounded prefix of the segments."""
    import numpy as np
    from oracle import voxel as OV
    idx = style.indices()
    n_seg = len(idx) // 2
    take = n_seg
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < budget_s:
        OV.voxelize(style.vertices, bounds, res, indices=idx[: 2 * take])
        reps += 1
    dt = time.perf_counter() - t0
    return dict(value=reps * take / dt, unit="segments/s", cores=1, kind="port",
                sample=f"{reps} x voxelize_segments over all {take} segments, 256^3, 1 thread, "
                       f"{os.cpu_count()} host cores present")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described, not built)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from vkhr_amd import synth
    from vkhr_amd import device as D

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    style = synth.config1()
    res = (256, 256, 256)
    xyz = D.to_dev(style.vertices, dev)
    off = D.to_dev(style.strand_offsets, dev)
    idx = D.generate_indices(off)
    tan = D.generate_tangents(xyz, off)
    bounds = D.aabb(xyz)
    n_seg = idx.shape[0] // 2
    counts = torch.empty(res[0] * res[1] * res[2], dtype=torch.int32, device=dev)

    def step():
        c, _ = D.voxelize(xyz, bounds, res, indices=idx, counts=counts)
        vol = D.finalize(c, res, normalize=True, want_cells=True)
        return vol

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # voxelize alone (the dominant kernel of this leg), HIP events on its stream
    vx0, vx1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    vx0.record()
    for _ in range(args.steps):
        D.voxelize(xyz, bounds, res, indices=idx, counts=counts)
    vx1.record()
    torch.cuda.synchronize()
    t_vox = vx0.elapsed_time(vx1) / args.steps * 1e-3

    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    t_step = ev0.elapsed_time(ev1) / args.steps * 1e-3

    samples = int(counts.to(torch.int64).sum().item())
    # algorithmic bytes of one voxelize launch: clear 4N^3, indices 8/segment,
    # two vertices 24/segment, one 4-byte atomic RMW (read+write) per sample
    n_vox = res[0] * res[1] * res[2]
    vox_bytes = 4 * n_vox + 8 * n_seg + 24 * n_seg + 8 * samples
    achieved = vox_bytes / t_vox / 1e9
    out = {
        "metric": "segments voxelized/s + raymarch Mrays/s at 256^3; adjoint/forward time ratio",
        "value": n_seg / t_vox,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded hemisphere hair)",
        "config": {"workload": "136,320 strands x 13 vertices, segment voxelization 256^3 + finalize",
                   "segments": n_seg, "voxel_samples": samples, "volume": "256^3"},
        "segments_per_s": n_seg / t_vox,
        "voxelize_ms": t_vox * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    if not args.no_cpu_baseline:
        b = {k: (np.array(v) if hasattr(v, "__len__") else v) for k, v in bounds.items()}
        out["cpu_baseline"] = cpu_baseline_voxelize(style, b, res)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
