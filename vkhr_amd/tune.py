"""Autotuner of the batched launch: times the whole graph (voxelize ->
finalize -> cells -> raymarch of all styles) for each launch shape and
writes the fastest into vkhr_amd/tuning.json under a workload key.

Shapes: the wave tile of the raymarch (tile_w x tile_h pixels, 64 per wave;
a block is 2x2 wave tiles) and the raymarch launch form (blocks_per_cu = 0:
one block per (style, tile) work item; n > 0: n persistent blocks per
compute unit pulling work items from an atomic queue).  Every shape gives
bit-identical images (tests/test_batch_gpu.py), so the choice is by time
alone.

    python -m vkhr_amd.tune            # the 32-style 960x540 workload, key "batch"
"""
from __future__ import annotations

import argparse
import json
import time

import torch

from . import capi, synth
from .batch import TUNING_PATH, Batch, StyleBuffers

TILES = [(8, 8), (16, 4), (4, 16), (32, 2)]
BLOCKS_PER_CU = [0, 2, 4, 8]
GROUPS = [1, 2, 3, 4, 6, 8]


def time_shape(bufs, width, height, tile, bpc, reps, groups=1):
    t = capi.Tuning()
    t.tile_w, t.tile_h, t.blocks_per_cu, t.groups = tile[0], tile[1], bpc, groups
    batch = Batch(bufs, width, height, tuning=t)
    batch.launch()  # capture + first replay
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(batch.stream)
    for _ in range(reps):
        batch.launch(wait=False)
    e1.record(batch.stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del batch
    return ms


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--key", default="batch")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--width", type=int, default=960)
    ap.add_argument("--height", type=int, default=540)
    ap.add_argument("--out", default=str(TUNING_PATH))
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    styles, _ = synth.batch_styles()
    bufs = [StyleBuffers(s, s.meta["volume"], args.width, args.height, dev, azimuth=11.25 * i, with_counts=False)
            for i, s in enumerate(styles)]
    results = []
    for tile in TILES:
        for bpc in BLOCKS_PER_CU:
            t0 = time.time()
            ms = time_shape(bufs, args.width, args.height, tile, bpc, args.reps)
            results.append(dict(tile_w=tile[0], tile_h=tile[1], blocks_per_cu=bpc, ms=round(ms, 4)))
            print(f"tile {tile[0]}x{tile[1]} blocks/CU {bpc}: {ms:.3f} ms ({time.time() - t0:.1f} s)", flush=True)
    best = min(results, key=lambda r: r["ms"])
    # then the pipeline depth for the best launch shape
    for g in GROUPS:
        t0 = time.time()
        ms = time_shape(bufs, args.width, args.height, (best["tile_w"], best["tile_h"]), best["blocks_per_cu"],
                        args.reps, g)
        results.append(dict(tile_w=best["tile_w"], tile_h=best["tile_h"], blocks_per_cu=best["blocks_per_cu"],
                            groups=g, ms=round(ms, 4)))
        print(f"groups {g}: {ms:.3f} ms ({time.time() - t0:.1f} s)", flush=True)
    best = min(results, key=lambda r: r["ms"])
    try:
        with open(args.out) as f:
            table = json.load(f)
    except FileNotFoundError:
        table = {}
    table[args.key] = dict(tile_w=best["tile_w"], tile_h=best["tile_h"], blocks_per_cu=best["blocks_per_cu"],
                           groups=best.get("groups", 1), ms=best["ms"], workload=f"{len(bufs)} styles, {args.width}x{args.height}",
                           device=torch.cuda.get_device_name(0), sweep=results)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    print("best", best, "->", args.out)


if __name__ == "__main__":
    main()
