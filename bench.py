"""Benchmark of the hybrid-hair hot path on one MI355X.

Workload (N=1): 136,320 strands x 13 vertices (1,635,840 segments) ->
segment voxelization into 256^3 with the tangent sums -> finalize
(saturating view, normalize, tangent quantization, corner-packed cells,
brick occupancy) in one call of the LDS-binned build -> volume raymarch at
1920x1080, 512 steps, Kajiya-Kay + ADSM + AO.  Inputs (vertices, indices in
the load-time spatial order, tangents, camera) are resident in HBM before
the timed region; a "step" is one pass build -> raymarch.

value = covered rays per second of the whole step (Mrays/s); the per-kernel
throughputs (raymarch Mrays/s, segments voxelized/s) are reported beside it.
The roofline is for the dominant kernel (the raymarch), on algorithmic
gathered bytes: 8 bytes per trilinear tap (8 R8 texels), 32 bytes per
tangent tap, 16 bytes of output per covered ray.

--gpus N > 1 is refused: the path runs on one GPU.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import re
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "segments voxelized/s + raymarch Mrays/s at 256^3; adjoint/forward time ratio"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=8.0, help="seconds per CPU-baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint and batched legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 256 MiB copy kernel three times (the reference for the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE traffic counters) and exit")
    ap.add_argument("--summarize-profile", metavar="ROOT",
                    help="write the raymarch counter summary from rocprofv3 CSV runs under ROOT (no GPU) and exit")
    ap.add_argument("--out", default=str(PROFILE_SUMMARY), help="summary file of --summarize-profile")
    ap.add_argument("--batched-only", action="store_true",
                    help="only time the 32-style batched leg against its per-style sequences (for a kernel "
                         "profile of that leg), print its JSON object and exit")
    ap.add_argument("--strand-passes-only", action="store_true",
                    help="only time the strand passes (shadow map, colour pass, resolve) at the bench's workload "
                         "(for a kernel profile of them), print their JSON object and exit")
    return ap.parse_args()


def copy_calibration():
    import ctypes as C
    import torch
    lib = C.CDLL(str(ROOT / "vkhr_amd" / "libvkhr_microbench.so"))
    lib.vkhr_mb_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    src = torch.ones(256 << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(3):
        assert lib.vkhr_mb_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), src.numel(), s) == 0
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": "3 x vkhr_mb_copy of 268435456 bytes"}))


def time_adjoint(dev, reps=10):
    """Differentiable form: 128^3 float32 normalized density from the
    10,000-strand style, 512x512, 256 steps; returns (forward ms, adjoint ms)."""
    import torch
    from vkhr_amd import params, synth
    from vkhr_amd import device as D
    res, w, h = (128, 128, 128), 512, 512
    s = synth.strands_10k()
    b, vol = D.style_volume(s, res, dev, segments=False, want_f32=True)
    fvol = dict(density_f32=vol["density_f32"], tangents=vol["tangents"])
    sc = synth.make_scene(b, w, h)
    cam, light, sp = params.camera(sc), params.light(sc), params.strand(b, res)
    rp = params.render(raycast_steps=256.0)
    diff = D.diff_params(0.3, 11.0, 80.0)
    rgba, rec = D.raymarch_record(fvol, res, cam, light, sp, rp, diff, w, h)
    g = torch.ones_like(rgba) * 1e-3
    gd = torch.zeros(res[0] ** 3, dtype=torch.float32, device=dev)
    gp = torch.zeros(3, dtype=torch.float32, device=dev)
    for _ in range(3):
        D.raymarch_record(fvol, res, cam, light, sp, rp, diff, w, h, record=rec, out=rgba)
        D.raymarch_adjoint(fvol, res, cam, light, sp, rp, diff, w, h, rec, g, gd, gp)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    for _ in range(reps):
        D.raymarch_record(fvol, res, cam, light, sp, rp, diff, w, h, record=rec, out=rgba)
    e[1].record()
    e[2].record()
    for _ in range(reps):
        D.raymarch_adjoint(fvol, res, cam, light, sp, rp, diff, w, h, rec, g, gd, gp)
    e[3].record()
    torch.cuda.synchronize()
    return e[0].elapsed_time(e[1]) / reps, e[2].elapsed_time(e[3]) / reps


def atomic_peak(dev, n_ops=1 << 26):
    """Device-atomic throughput measured by the build's microbenchmark
    (libvkhr_microbench.so): uniform-random uint32 adds over 64 MiB (one
    line request each) and 16-byte pairs like the voxelizer's (two 8-byte
    adds per line request, random lines); returns requests/s for both."""
    import ctypes as C
    import torch
    lib = C.CDLL(str(ROOT / "vkhr_amd" / "libvkhr_microbench.so"))
    lib.vkhr_mb_atomics.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]
    buf = torch.zeros(16 << 20, dtype=torch.int32, device=dev)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, kind, pattern in (("u32_uniform", 0, 0), ("u32_stream", 0, 3)):
        lib.vkhr_mb_atomics(kind, pattern, C.c_void_p(buf.data_ptr()), buf.numel(), n_ops, s)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        assert lib.vkhr_mb_atomics(kind, pattern, C.c_void_p(buf.data_ptr()), buf.numel(), n_ops, s) == 0
        e1.record()
        torch.cuda.synchronize()
        out[name] = n_ops / (e0.elapsed_time(e1) * 1e-3)
    return out


def time_raster(dev, style, xyz, idx, tan, vol, res, scene, cam, light, bounds, reps=5):
    """The strand passes at the same workload: 1024^2 shadow depth, strand
    raster into a 2^25-node PPLL (line-coverage AA, KK, PCF ADSM, LAO) at
    1920x1080, k=16 resolve; returns (shadow ms, raster ms, resolve ms,
    fragments)."""
    import numpy as np
    import torch
    from vkhr_amd import params
    from vkhr_amd import device as D
    W, H = scene.width, scene.height
    th = D.to_dev(style.thickness, dev)
    light_vp = np.asarray(scene.light_vp, np.float64).T.reshape(-1).astype(np.float32)
    sp = params.strand(bounds, res, strand_radius=1.0)
    rp = params.render(renderer=0)
    ppll = D.PPLL(W, H, 1 << 25, dev)
    smap = torch.empty((1024, 1024), dtype=torch.float32, device=dev)
    out = torch.empty((H, W, 4), dtype=torch.float32, device=dev)

    def shadow():
        D.shadow_depth(xyz, idx, light_vp, 1.0, (1024, 1024), out=smap)

    def raster():
        ppll.clear()
        D.raster_strands(xyz, tan, th, idx, cam, light, sp, rp, vol, res, smap, ppll)

    def resolve():
        ppll.resolve(out=out)

    shadow()
    raster()
    resolve()
    torch.cuda.synchronize()
    times = []
    for fn in (shadow, raster, resolve):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / reps)
    return times[0], times[1], times[2], min(ppll.used(), ppll.node_count)


def raster_result(t_sh, t_ra, t_re, n_frag, n_seg, W, H, atomic_peak_per_s) -> dict:
    """The strand passes' JSON object (times from time_raster)."""
    # algorithmic work per fragment: 9 shadow taps x 4 B, 8 AO taps x 8
    # texels x 1 B, the 12-byte node; per segment the two vertices'
    # position, tangent and thickness; one head exchange per fragment
    ra_bytes = (36 + 64 + 12) * n_frag + 2 * (12 + 12 + 4) * n_seg
    re_bytes = (4 + 16) * W * H + 12 * n_frag
    out = {"workload": f"{n_seg:,} segments, {W}x{H}, 1024^2 shadow map, k=16, 2^25-node PPLL",
           "shadow_ms": t_sh, "raster_ms": t_ra, "resolve_ms": t_re, "fragments": n_frag,
           "fragments_per_s": n_frag / (t_ra * 1e-3),
           "algorithmic_bytes": ra_bytes, "achieved_gbs": ra_bytes / (t_ra * 1e-3) / 1e9,
           "hbm_frac": ra_bytes / (t_ra * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "head_exchanges_per_s": n_frag / (t_ra * 1e-3),
           # the resolve: per pixel a 4-byte head and a 16-byte colour out, per
           # node 12 bytes, each node a dependent load of the pixel's list
           "resolve_bytes": re_bytes, "resolve_achieved_gbs": re_bytes / (t_re * 1e-3) / 1e9,
           "resolve_hbm_frac": re_bytes / (t_re * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if atomic_peak_per_s:
        out["atomic_frac"] = n_frag / (t_ra * 1e-3) / atomic_peak_per_s
    return out


def strand_passes_only(dev) -> dict:
    """The strand passes alone at the bench's workload (136k strands, the
    256^3 volume for the LAO, 1080p), for a kernel profile of them."""
    from vkhr_amd import params, synth
    from vkhr_amd import device as D
    style = synth.strands_136k()
    res = (256, 256, 256)
    xyz = D.to_dev(style.vertices, dev)
    off = D.to_dev(style.strand_offsets, dev)
    tan = D.generate_tangents(xyz, off)
    bounds = D.aabb(xyz)
    idx_mid = D.sort_segments(xyz, D.generate_indices(off), bounds)
    vol = D.build_volume(xyz, bounds, res, idx_mid, tan, normalize=True, want_cells=True)
    scene = synth.make_scene(bounds, 1920, 1080)
    cam, light = params.camera(scene), params.light(scene)
    t_sh, t_ra, t_re, n_frag = time_raster(dev, style, xyz, idx_mid, tan, vol, res, scene, cam, light, bounds)
    return raster_result(t_sh, t_ra, t_re, n_frag, idx_mid.shape[0] // 2, scene.width, scene.height, None)


def time_trace(dev, style, xyz, idx, tan, scene, bounds, frames=4):
    """The strand ray tracer at the same workload: BVH build over the
    1,635,840 segments (radius 0.004 x thickness), then samples of the
    reference's Raytracer frame at 1920x1080 -- shaded (primary, shadow and
    AO rays) and AO only; returns a dict of times and ray rates."""
    import math

    import torch
    from vkhr_amd import params, synth
    from vkhr_amd import device as D
    from vkhr_amd.trace import METHODS, Bvh, viewing_plane
    W, H = scene.width, scene.height
    rad = D.to_dev(style.thickness, dev) * 0.004
    bvh = Bvh(xyz, rad, idx)   # warm: the first build also initialises the sort
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        bvh = Bvh(xyz, rad, idx)
    t_build = (time.perf_counter() - t0) / 3
    plane = viewing_plane(scene.eye, scene.look_at, (0, 1, 0), math.radians(synth.CAMERA_FOV_DEG), W, H)
    light = params.light(scene)
    ao_radius = 2.5 * float(max(bounds["size"])) / 256.0  # the LAO's 2.5 voxels, in world units
    out = {"workload": "1,635,840 segments (flat linear curves, radius 0.004 x thickness), 1920x1080, "
                       "1 sample per pixel per frame", "build_ms": t_build * 1e3, "ao_radius": ao_radius}
    for name in ("shaded", "ambient_occlusion"):
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
        bvh.raytrace(tan, idx, plane, light, W, H, acc, method=METHODS[name], ao_radius=ao_radius, frame=0)
        torch.cuda.synchronize()
        hits = int((acc.reshape(-1, 3) != 1.0).any(dim=1).sum().item())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for f in range(frames):
            bvh.raytrace(tan, idx, plane, light, W, H, acc, method=METHODS[name], ao_radius=ao_radius, frame=f + 1)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / frames
        # rays per sample: one primary per pixel; per hit a shadow ray (shaded) and an AO ray
        rays = W * H + hits * (2 if name == "shaded" else 1)
        out[name] = {"frame_ms": ms, "hit_pixels": hits, "rays_per_frame": rays, "mrays_per_s": rays / ms / 1e3}
    return out


def time_reference_frames(dev, style, xyz, idx, tan, vol, res, bounds, reps=5):
    """The reference's published setup (ponytail, 1280x720, 1024^2 3x3 PCF
    shadow map, 256^3 volume, 512 steps; docs/vkhr.tex:352,387): the frame
    of the strand rasterizer (shadow map, PPLL clear, strands, resolve) and
    of the raymarcher (PPLL clear, raymarch into the PPLL, resolve), with
    this ponytail-sized synthetic style.  Returns ms per frame."""
    import numpy as np
    import torch
    from vkhr_amd import params, synth
    from vkhr_amd import device as D
    W, H = 1280, 720
    scene = synth.make_scene(bounds, W, H)
    cam, light = params.camera(scene), params.light(scene)
    th = D.to_dev(style.thickness, dev)
    light_vp = np.asarray(scene.light_vp, np.float64).T.reshape(-1).astype(np.float32)
    sp = params.strand(bounds, res, strand_radius=1.0)
    ppll = D.PPLL(W, H, 1 << 25, dev)
    smap = torch.empty((1024, 1024), dtype=torch.float32, device=dev)
    out = torch.empty((H, W, 4), dtype=torch.float32, device=dev)

    def raster_frame():
        D.shadow_depth(xyz, idx, light_vp, 1.0, (1024, 1024), out=smap)
        ppll.clear()
        D.raster_strands(xyz, tan, th, idx, cam, light, sp, params.render(renderer=0), vol, res, smap, ppll)
        ppll.resolve(out=out)

    def raymarch_frame():
        ppll.clear()
        D.raymarch_ppll(vol, res, cam, light, sp, params.render(renderer=2, raycast_steps=512.0), ppll)
        ppll.resolve(out=out)

    times = {}
    for name, fn in (("rasterizer", raster_frame), ("raymarcher", raymarch_frame)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / reps
    return times


def ao_ground_truth(dev, style, xyz, idx, tan, bounds, res, W=960, H=540, samples=64):
    """The ray tracer as the quality oracle of the raymarcher's local ambient
    occlusion: per pixel, the raymarcher's LAO factor at its surface point
    (from the march record) against the ray-traced AO at the strand hit
    (the unoccluded fraction of `samples` AO rays of the LAO's radius).
    Returns the Pearson correlation over pixels where both hit, and means."""
    import math

    import numpy as np
    import torch
    from vkhr_amd import params, synth
    from vkhr_amd import device as D
    from vkhr_amd.trace import METHODS, Bvh, viewing_plane
    sums = torch.empty(2 * res[0] * res[1] * res[2], dtype=torch.int64, device=dev)
    D.voxelize(xyz, bounds, res, indices=idx, tangents=tan, want_sums=True, sums=sums, want_counts=False)
    vol = D.finalize(None, res, sums=sums, normalize=True, want_f32=True, want_cells=False)
    vol = dict(density=vol["density"], cells=None, tangents=vol["tangents"], density_f32=vol["density_f32"],
               occupancy=None)
    scene = synth.make_scene(bounds, W, H)
    cam, light = params.camera(scene), params.light(scene)
    sp, rp = params.strand(bounds, res), params.render(raycast_steps=512.0)
    _, rec = D.raymarch_record(vol, res, cam, light, sp, rp, D.diff_params(0.8), W, H)
    r = rec.view(-1, 64).cpu().numpy()
    lao = r[:, 36:40].copy().view(np.float32).reshape(H, W)
    state = r[:, 60:64].copy().view(np.int32).reshape(H, W)
    rad = D.to_dev(style.thickness, dev) * 0.004
    bvh = Bvh(xyz, rad, idx)
    plane = viewing_plane(scene.eye, scene.look_at, (0, 1, 0), math.radians(synth.CAMERA_FOV_DEG), W, H)
    ao_radius = float(rp.occlusion_radius) * float(max(bounds["size"])) / res[0]
    def unoccluded(ignore_self):
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device=dev)
        for f in range(samples):
            bvh.raytrace(tan, idx, plane, light, W, H, acc, method=METHODS["ambient_occlusion"], ao_radius=ao_radius,
                         frame=f, ignore_self=ignore_self)
        a = acc[:, :, 0].cpu().numpy() / samples
        hit = a != 1.0  # a miss scores 1 on every sample
        return np.where(hit, a / 2.0, np.nan)[:, ::-1]  # the viewing plane's x runs to the left: mirror it

    rt = unoccluded(True)
    rt_ref = unoccluded(False)
    both = (state == 1) & np.isfinite(rt)
    x, y = lao[both].astype(np.float64), rt[both]
    corr = float(np.corrcoef(x, y)[0, 1]) if both.sum() > 2 else float("nan")
    return {"workload": f"{W}x{H}, {samples} AO samples per pixel, radius {ao_radius:.4f} (the LAO's "
                        f"{rp.occlusion_radius} voxels), AO rays skip the segment they start on",
            "pixels": int(both.sum()), "pearson_lao_vs_raytraced": corr,
            "mean_lao": float(x.mean()) if len(x) else None,
            "mean_raytraced_unoccluded": float(y.mean()) if len(y) else None,
            "mean_raytraced_unoccluded_reference_rule": float(np.nanmean(rt_ref[both])) if both.any() else None}


def time_adjoint_batch(dev):
    """Batched adjoint: 8 generator styles (128^3 float densities) x 8 ring
    views at 512x512, 256 steps; returns (per-render ms, graph ms) for the
    forward + L2 loss gradient + adjoint of all 64 renders."""
    import torch
    from vkhr_amd.adjoint_batch import AdjointBatch, AdjointProblem
    prob = AdjointProblem(dev)
    batch = AdjointBatch(prob)
    for _ in range(2):
        batch.launch()
        prob.run_plain()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    prob.run_plain()
    e[1].record()
    torch.cuda.synchronize()
    e[2].record(batch.stream)
    batch.launch(wait=False)
    e[3].record(batch.stream)
    torch.cuda.synchronize()
    return e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])


PROFILE_SUMMARY = ROOT / "profiles" / "raymarch_counters_136k_strands_1080p.md"
# What compiles into the raymarch kernel: its translation unit, the headers
# it includes and the build flags.  A profile summary holds the hash of these
# files as they were when it was taken, and its counters are used only for
# the same sources.
RAYMARCH_SOURCES = ("vkhr_amd/csrc/raymarch.hip", "vkhr_amd/csrc/march.hpp", "vkhr_amd/csrc/render.hpp",
                    "vkhr_amd/csrc/common.hpp")
# the lines of vkhr_amd/Makefile that set the device compile flags
MAKEFILE_FLAG_LINES = ("HIPFLAGS", "STRICT", "SRC_FAST", "             -Xclang", "             -I")
RAYMARCH_KERNEL = "raymarch_kernel<false>"
CLOCK_HZ = 2.4e9          # MI355X peak engine clock
SIMDS = 1024              # 256 CUs x 4 SIMDs
VALU_CYCLES = 4           # a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles


def source_hash(files=RAYMARCH_SOURCES) -> str:
    h = hashlib.sha256()
    for f in files:
        h.update(f.encode())
        h.update((ROOT / f).read_bytes())
    for line in (ROOT / "vkhr_amd" / "Makefile").read_text().splitlines():
        if line.startswith(MAKEFILE_FLAG_LINES):
            h.update(line.encode())
    return h.hexdigest()[:16]


def profile_counters():
    """(counters, None) from the committed rocprofv3 summary of the raymarch
    kernel when it was taken of this tree's kernel sources, else (None,
    reason)."""
    if not PROFILE_SUMMARY.exists():
        return None, f"{PROFILE_SUMMARY.name} missing"
    m = re.search(r"```json\n(.*?)\n```", PROFILE_SUMMARY.read_text(), re.S)
    if not m:
        return None, f"{PROFILE_SUMMARY.name} holds no counter block"
    data = json.loads(m.group(1))
    here = source_hash()
    if data.get("source_hash") != here:
        return None, f"the profile was taken of other kernel sources ({data.get('source_hash')} != {here})"
    return data, None


def raymarch_roofline(t_ray: float, ray_bytes: float) -> dict:
    """The raymarch kernel against the resource that binds it, VALU issue:
    its VALU wave-instructions per launch (from the profile summary of the
    same kernel sources) x 4 cycles over 1024 SIMDs x 2.4 GHz x the launch
    time measured in this run.  Beside it the measured DRAM traffic per
    launch against 8 TB/s, and the gathered-byte figure of the per-ray
    formula (every sample of every covered ray, skipped ones included),
    which is not a DRAM figure."""
    prof, why = profile_counters()
    peak = SIMDS * CLOCK_HZ / VALU_CYCLES / 1e9   # G wave-instructions per second
    out = {"bound": "valu", "kernel": RAYMARCH_KERNEL, "unit": "G VALU wave-instructions/s", "peak": peak,
           "achieved": None, "frac": None, "traffic": None, "seconds_per_launch": t_ray,
           "profile": f"profiles/{PROFILE_SUMMARY.name}", "source_hash": source_hash()}
    if prof is None:
        out["profile_note"] = why
    else:
        out["achieved"] = prof["valu_wave_instructions"] / t_ray / 1e9
        out["frac"] = out["achieved"] / peak
        out["traffic"] = prof["dram_bytes"]
        out["dram"] = {"bytes_per_launch": prof["dram_bytes"], "achieved_gbs": prof["dram_bytes"] / t_ray / 1e9,
                       "frac": prof["dram_bytes"] / t_ray / 1e9 / HBM_PEAK_GBS}
        out["profile_avg_duration_us"] = prof["avg_duration_us"]
        out["profile_valu_issue_frac"] = prof["valu_issue_frac"]
    out["gathered_bytes_effective"] = {
        "bytes_per_launch": ray_bytes, "gbs": ray_bytes / t_ray / 1e9, "frac_of_hbm": ray_bytes / t_ray / 1e9 / HBM_PEAK_GBS,
        "note": "the per-ray formula 8 B x (S primary + S + 8 ADSM/LAO taps) + 48 B over every covered ray, samples "
                "in empty brick regions included although the kernel skips them; not a DRAM figure"}
    return out


def _csv_rows(root: Path, suffix: str):
    import csv
    rows = []
    for f in sorted(root.rglob(f"*{suffix}")):
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _counter_means(root: Path) -> dict:
    """{kernel name: {counter: mean per dispatch}} from rocprofv3 --pmc CSV output."""
    per = {}
    for r in _csv_rows(root, "counter_collection.csv"):
        key = (r["Kernel_Name"], r.get("Dispatch_Id", r.get("Correlation_Id")))
        d = per.setdefault(key, {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    out = {}
    for (name, _), d in per.items():
        k = out.setdefault(name, {})
        for c, v in d.items():
            k.setdefault(c, []).append(v)
    return {n: {c: sum(v) / len(v) for c, v in d.items()} for n, d in out.items()}


def _short(name: str) -> str:
    name = re.sub(r"\(.*$", "", name.replace("(anonymous namespace)::", ""))
    return name.replace("vkhr::", "").replace("void ", "")


def summarize_profile(root: Path, out: Path):
    """Writes the profile summary of the raymarch kernel (and the step's
    other kernels) from rocprofv3 runs of bench.py under root: stats/
    (--kernel-trace --stats), fetch/ and write/ (--pmc FETCH_SIZE,
    --pmc WRITE_SIZE), sq/ (--pmc SQ counters), copy_fetch/ and copy_write/
    (the same two counters over bench.py --copy-calibration)."""
    stats = {r["Name"]: r for r in _csv_rows(root / "stats", "kernel_stats.csv")}
    fetch, write = _counter_means(root / "fetch"), _counter_means(root / "write")
    sq = _counter_means(root / "sq")
    cf, cw = _counter_means(root / "copy_fetch"), _counter_means(root / "copy_write")
    copy_name = next(n for n in cf if "copy_kernel" in n or "mb_copy" in n)
    copy_bytes = 268435456.0
    f_scale = copy_bytes / (cf[copy_name]["FETCH_SIZE"] * 1024.0)
    w_scale = copy_bytes / (cw[copy_name]["WRITE_SIZE"] * 1024.0)
    names = sorted(set(fetch) | set(sq), key=lambda n: -float(stats.get(n, {}).get("TotalDurationNs", 0)))
    rm = next(n for n in names if RAYMARCH_KERNEL in n)
    # the tree the counters were taken of: VKHR_PROFILE_TREE when the run
    # happened in a copy without git metadata, else the checkout's HEAD
    commit, dirty = os.environ.get("VKHR_PROFILE_TREE", ""), ""
    if not commit:
        try:
            commit = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "--short=12", "HEAD"], capture_output=True,
                                    text=True).stdout.strip() or "unknown"
            dirty = subprocess.run(["git", "-C", str(ROOT), "status", "--porcelain", "--", *RAYMARCH_SOURCES],
                                   capture_output=True, text=True).stdout.strip()
        except OSError:
            commit = "unknown"
    # the average over the dispatches that ran alone: in the pipelined steps
    # a raymarch overlaps the next step's volume build on the other stream,
    # and bench.py times the kernel alone
    trace = _csv_rows(root / "stats", "kernel_trace.csv")
    spans = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in trace]
    solo = [e - b for b, e, n in spans if n == rm and not any(b2 < e and e2 > b for b2, e2, n2 in spans
                                                              if (b2, e2, n2) != (b, e, n))]
    avg_all_ns = float(stats[rm]["AverageNs"]) if rm in stats else None
    avg_ns = sum(solo) / len(solo) if solo else avg_all_ns
    valu = sq[rm].get("SQ_INSTS_VALU")
    traffic = fetch[rm]["FETCH_SIZE"] * 1024.0 * f_scale + write[rm]["WRITE_SIZE"] * 1024.0 * w_scale
    block = {
        "source_hash": source_hash(), "tree": commit + ("+uncommitted kernel sources" if dirty else ""),
        "kernel": rm, "avg_duration_us": avg_ns / 1e3 if avg_ns else None, "solo_dispatches": len(solo),
        "avg_duration_all_dispatches_us": avg_all_ns / 1e3 if avg_all_ns else None,
        "valu_wave_instructions": valu, "salu_wave_instructions": sq[rm].get("SQ_INSTS_SALU"),
        "vmem_read_wave_instructions": sq[rm].get("SQ_INSTS_VMEM_RD"), "lds_wave_instructions": sq[rm].get("SQ_INSTS_LDS"),
        "waves": sq[rm].get("SQ_WAVES"),
        "fetch_kib": fetch[rm]["FETCH_SIZE"], "write_kib": write[rm]["WRITE_SIZE"],
        "fetch_scale": f_scale, "write_scale": w_scale, "dram_bytes": traffic,
        "valu_issue_frac": (valu * VALU_CYCLES / (SIMDS * CLOCK_HZ * avg_ns * 1e-9)) if (valu and avg_ns) else None,
        "dram_frac": traffic / (avg_ns * 1e-9) / (HBM_PEAK_GBS * 1e9) if avg_ns else None,
    }
    lines = [
        "# Raymarch kernel counters: 136,320-strand step (256^3, 1920x1080, 512 steps)", "",
        "Commands (one MI355X, repository root, `cd /tmp && export TMPDIR=/tmp` first; each `--pmc` pass a run of its own):",
        "", "```",
        "rocprofv3 --kernel-trace --stats -d <root>/stats -o run --output-format csv -- python3 bench.py --steps 10 --warmup 2 --no-cpu-baseline --no-extras",
        "rocprofv3 --pmc FETCH_SIZE -d <root>/fetch -o run --output-format csv -- python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-extras",
        "rocprofv3 --pmc WRITE_SIZE -d <root>/write -o run --output-format csv -- python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-extras",
        "rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_LDS SQ_BUSY_CYCLES -d <root>/sq -o run --output-format csv -- python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-extras",
        "rocprofv3 --pmc FETCH_SIZE -d <root>/copy_fetch -o run --output-format csv -- python3 bench.py --copy-calibration",
        "rocprofv3 --pmc WRITE_SIZE -d <root>/copy_write -o run --output-format csv -- python3 bench.py --copy-calibration",
        "python3 bench.py --summarize-profile <root> --out profiles/raymarch_counters_136k_strands_1080p.md",
        "```", "",
        f"Copy calibration: the 256 MiB copy kernel reads {cf[copy_name]['FETCH_SIZE']:.0f} KiB (FETCH_SIZE) and writes "
        f"{cw[copy_name]['WRITE_SIZE']:.0f} KiB (WRITE_SIZE) per dispatch for 262144 KiB each way: scale factors "
        f"{f_scale:.3f} (fetch) and {w_scale:.3f} (write).", "",
        f"VALU issue fraction = VALU wave-instructions x {VALU_CYCLES} cycles / ({SIMDS} SIMDs x {CLOCK_HZ / 1e9:.1f} GHz x "
        "average duration), the average over the dispatches that ran with no other kernel beside them (the "
        "pipelined steps overlap a raymarch with the next volume build).  `bench.py` reads the block below only while the hash of the raymarch kernel's sources "
        f"({', '.join(RAYMARCH_SOURCES)} and the device flags of vkhr_amd/Makefile) equals its `source_hash`.", "",
        "| kernel | avg us | VALU | SALU | VMEM rd | LDS | waves | FETCH KiB | WRITE KiB | HBM MB (scaled) |",
        "|---|---|---|---|---|---|---|---|---|---|"]
    for n in names:
        st, q = stats.get(n, {}), sq.get(n, {})
        fk, wk = fetch.get(n, {}).get("FETCH_SIZE"), write.get(n, {}).get("WRITE_SIZE")
        hbm = (fk * 1024 * f_scale + wk * 1024 * w_scale) / 1e6 if fk is not None and wk is not None else None

        def g(v, fmt="{:.3g}"):
            return fmt.format(v) if v is not None else "-"
        lines.append(f"| `{_short(n)[:90]}` | {g(float(st['AverageNs']) / 1e3 if st else None, '{:.1f}')} | "
                     f"{g(q.get('SQ_INSTS_VALU'))} | {g(q.get('SQ_INSTS_SALU'))} | {g(q.get('SQ_INSTS_VMEM_RD'))} | "
                     f"{g(q.get('SQ_INSTS_LDS'))} | {g(q.get('SQ_WAVES'), '{:.0f}')} | {g(fk, '{:.0f}')} | "
                     f"{g(wk, '{:.0f}')} | {g(hbm, '{:.1f}')} |")
    lines += ["", "Counter block of the raymarch kernel (per dispatch):", "", "```json", json.dumps(block, indent=1),
              "```", ""]
    out.write_text("\n".join(lines))
    print(json.dumps(block))


def time_batch(dev, reps=3):
    """Batched form: 32 styles of uneven size at 960x540; returns (per-style
    launches back to back ms, batched graph ms, sum over styles of one
    style's launches waited for, ms, per-style launches with the one-call
    volume build ms)."""
    import torch
    from vkhr_amd import synth
    from vkhr_amd.batch import Batch, StyleBuffers, load_tuning
    styles, _ = synth.batch_styles()
    bufs = [StyleBuffers(st, st.meta["volume"], 960, 540, dev, azimuth=11.25 * i, with_counts=False)
            for i, st in enumerate(styles)]
    batch = Batch(bufs, 960, 540, tuning=load_tuning("batch"))
    for _ in range(2):
        batch.launch()
        for b in bufs:
            b.run_plain()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
    e[0].record()
    for _ in range(reps):
        for b in bufs:
            b.run_plain()
    e[1].record()
    torch.cuda.synchronize()
    e[2].record(batch.stream)
    for _ in range(reps):
        batch.launch(wait=False)
    e[3].record(batch.stream)
    torch.cuda.synchronize()
    # the per-style sequences with the one-call volume build
    for b in bufs:
        b.run_plain(one_call=True)
    torch.cuda.synchronize()
    e[4].record()
    for _ in range(reps):
        for b in bufs:
            b.run_plain(one_call=True)
    e[5].record()
    torch.cuda.synchronize()
    # each style run on its own and waited for (wall clock, host issue included)
    synced = 0.0
    for b in bufs:
        t0 = time.perf_counter()
        b.run_plain()
        torch.cuda.synchronize()
        synced += time.perf_counter() - t0
    return (e[0].elapsed_time(e[1]) / reps, e[2].elapsed_time(e[3]) / reps, synced * 1e3,
            e[4].elapsed_time(e[5]) / reps)


def batched_result(t_plain, t_batch, t_synced, t_plain_build) -> dict:
    return {"workload": "32 styles, 5k-140k strands, 64^3/128^3/256^3, 960x540, 512 steps",
            "per_style_ms": t_plain, "batched_graph_ms": t_batch, "speedup": t_plain / t_batch,
            "per_style_synced_ms": t_synced, "speedup_vs_synced": t_synced / t_batch,
            "per_style_one_call_build_ms": t_plain_build,
            "speedup_vs_one_call_build": t_plain_build / t_batch,
            "note": "per_style_ms: the 32 per-style sequences (voxelize_segments, finalize, raymarch) issued back "
                    "to back; per_style_synced_ms: the sum of each style's sequence timed alone; "
                    "per_style_one_call_build_ms: the per-style sequences with the one-call volume build instead "
                    "of voxelize + finalize"}


def embree_present() -> bool:
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=10).stdout
    except Exception:
        return False
    return "libembree3" in out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may run on (the GPU box grants a share of the host)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baselines(style, bounds, res, vol_np, scene_args, covered_pixels, budget):
    """Host-core legs, each for about `budget` seconds.

    The reported value is the reference's own load-time volume build,
    restated in the oracle: HairStyle::voxelize_segments
    (src/vkhr/scene_graph/hair_style.cc:296-342) -- the serial DDA over the
    segments with the saturating uint8 densities and the first-255 tangent
    sums, then the tangent quantization over an OpenMP team (dynamic
    schedule, as the reference) -- followed by Volume::normalize (:344-357),
    in segments per second.  Beside it, not the reference's: the OpenMP
    count-only DDA and the float64 raymarch transcription (this build's
    checker; the reference has no CPU raymarcher)."""
    import numpy as np
    from oracle import raymarch as OR
    from oracle import voxel as OV
    cores = host_cores()
    idx = style.indices()
    tan = OV.generate_tangents(style.vertices, style.strand_offsets)
    n_seg = len(idx) // 2
    t_dda = t_quant = t_norm = 0.0
    reps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget or reps == 0:
        OV.set_threads(1)
        a = time.perf_counter()
        v = OV.voxelize(style.vertices, bounds, res, indices=idx, tangents=tan, want_sums=True, want_sums_all=False)
        b = time.perf_counter()
        OV.set_threads(cores)
        OV.quantize_tangents(v["sums_ref"], v["dens_u8"])
        c = time.perf_counter()
        OV.set_threads(1)
        OV.normalize(v["dens_u8"])
        d = time.perf_counter()
        t_dda, t_quant, t_norm = t_dda + (b - a), t_quant + (c - b), t_norm + (d - c)
        reps += 1
    build_rate = reps * n_seg / (t_dda + t_quant + t_norm)
    OV.set_threads(cores)
    t0, reps_omp = time.perf_counter(), 0
    while time.perf_counter() - t0 < budget / 2 or reps_omp == 0:
        OV.voxelize_counts_omp(style.vertices, bounds, res, idx)
        reps_omp += 1
    seg_rate_omp = reps_omp * n_seg / (time.perf_counter() - t0)
    rng = np.random.default_rng(0)
    order = rng.permutation(covered_pixels)
    cam, light, sp, rp, w, h = scene_args
    rates = {}
    start = 0
    for threads, chunk_n in ((1, 500), (cores, 500 * cores)):
        OV.set_threads(threads)
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget or done == 0:
            chunk = order[start + done:start + done + chunk_n]
            if len(chunk) == 0:
                break
            OR.raymarch(vol_np["density"], vol_np["tangents"], res, cam, light, sp, rp, w, h, pixels=chunk)
            done += len(chunk)
        rates[threads] = (done, done / (time.perf_counter() - t0))
        start = (start + done) % max(len(order) - chunk_n, 1)
    OV.set_threads(1)
    done_all, ray_rate = rates[cores]
    done_one, ray_rate_one = rates[1]
    return dict(
        value=build_rate, unit="segments/s", cores=cores, kind="port",
        sample=(f"HairStyle::voxelize_segments + Volume::normalize (src/vkhr/scene_graph/hair_style.cc:296-357) "
                f"restated in the oracle, {reps} passes over the {n_seg} segments into 256^3: the DDA with the "
                f"first-255 tangent sums on 1 thread (the reference's serial loop, {t_dda / reps * 1e3:.0f} ms per "
                f"pass), the tangent quantization on {cores} threads (OpenMP, dynamic schedule, as the reference; "
                f"{t_quant / reps * 1e3:.0f} ms), normalize on 1 thread ({t_norm / reps * 1e3:.0f} ms); host "
                f"{cpu_model()}, {cores} cores available"),
        reference_function="HairStyle::voxelize_segments + HairStyle::Volume::normalize "
                           "(src/vkhr/scene_graph/hair_style.cc:296-357)",
        dda_segments_per_s_1_thread=reps * n_seg / t_dda,
        segments_per_s_openmp_count_only=seg_rate_omp,
        raymarch_port={"mrays_per_s": ray_rate / 1e6, "threads": cores, "pixels": done_all,
                       "mrays_per_s_1_thread": ray_rate_one / 1e6, "pixels_1_thread": done_one,
                       "note": "this build's float64 transcription of volume.frag (its checker), on random covered "
                               "pixels of the 1920x1080 frame (512 steps, 256^3); the reference has no CPU "
                               "raymarcher.  Embree " + ("present (not timed)" if embree_present() else
                                                         "unavailable")})


def main():
    args = parse()
    if args.summarize_profile:
        summarize_profile(Path(args.summarize_profile), Path(args.out))
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this path runs on one GPU; --gpus must be 1 (the multi-GPU sharding is a "
                                   "design, not built)"}))
        sys.exit(2)
    if args.batched_only:
        import torch
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        print(json.dumps({"batched": batched_result(*time_batch(dev))}))
        return
    if args.strand_passes_only:
        import torch
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        print(json.dumps({"raster": strand_passes_only(dev)}))
        return
    import numpy as np
    import torch

    from vkhr_amd import params, synth
    from vkhr_amd import device as D

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    style = synth.strands_136k()
    res = (256, 256, 256)
    W, H, S = 1920, 1080, 512
    xyz = D.to_dev(style.vertices, dev)
    off = D.to_dev(style.strand_offsets, dev)
    tan = D.generate_tangents(xyz, off)
    bounds = D.aabb(xyz)
    # load-time layouts (permutations of the pairs: every pass gives the same
    # results on them): for the step's volume build the strands in the
    # Morton order of their roots, each strand's pairs together; for the
    # strand passes the pairs in the Morton order of their midpoints.  The
    # first calls also initialise the sorts, so warm calls are timed.
    idx_generated = D.generate_indices(off)
    idx = D.sort_strands(xyz, idx_generated, bounds)
    idx_mid = D.sort_segments(xyz, idx_generated, bounds)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        idx = D.sort_strands(xyz, idx_generated, bounds)
    torch.cuda.synchronize()
    t_sort = (time.perf_counter() - t0) / 3
    del idx_generated
    n_seg = idx.shape[0] // 2
    nvox = res[0] * res[1] * res[2]
    scene = synth.make_scene(bounds, W, H)
    cam, light = params.camera(scene), params.light(scene)
    sp, rp = params.strand(bounds, res), params.render(raycast_steps=float(S))
    rgbas = [torch.empty((H, W, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    rgba = rgbas[0]
    # the volume build's workspace: read or written only if the brick bins
    # overflow (they do not at this workload); one per buffer set
    workspaces = [torch.empty(2 * nvox, dtype=torch.int64, device=dev) for _ in range(2)]
    vols = [None, None]

    def build(k):
        """voxelize_segments + normalize + cells in one call into buffer set k."""
        vols[k] = D.build_volume(xyz, bounds, res, idx, tan, normalize=True, want_cells=True,
                                 workspace=workspaces[k], out=vols[k])
        return vols[k]

    def step():
        vol = build(0)
        D.raymarch(vol, res, cam, light, sp, rp, W, H, want_depth=True, out=rgba)

    # Two steps in flight: step k runs on stream k % 2 with its own volume
    # and image, so step k+1's volume build (LDS atomics and HBM streams)
    # overlaps step k's raymarch (VALU bound).  Both streams have the
    # default priority; the library keeps a scratch area per stream.
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]

    def pipelined_step(k):
        with torch.cuda.stream(streams[k % 2]):
            vol = build(k % 2)
            D.raymarch(vol, res, cam, light, sp, rp, W, H, want_depth=True, out=rgbas[k % 2])

    for k in range(args.warmup):
        step()
        pipelined_step(k)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]

    # the whole job: K steps, two in flight, bracketed by synchronizes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        pipelined_step(k)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    # every step's image equals the serial one
    step()
    torch.cuda.synchronize()
    images_match = bool(torch.equal(rgbas[1], rgba))

    # one step at a time (the latency of a step)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(args.steps):
        step()
    ev[1].record()
    torch.cuda.synchronize()
    t_serial = ev[0].elapsed_time(ev[1]) / args.steps * 1e-3

    # the raymarch kernel alone (events on the stream it is launched on)
    vol = vols[0]
    ev[2].record()
    for _ in range(args.steps):
        D.raymarch(vol, res, cam, light, sp, rp, W, H, want_depth=True, out=rgba)
    ev[3].record()
    # the volume build alone
    ev[4].record()
    for _ in range(args.steps):
        build(1)
    ev[5].record()
    # the two-call path it replaces: voxelize_segments into the packed
    # accumulator (zeroed first), then finalize (unpack, uint8 view,
    # tangents, normalize, cells, distance field)
    sums = workspaces[1]
    ev_two = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev_two[0].record()
    for _ in range(args.steps):
        D.voxelize(xyz, bounds, res, indices=idx, tangents=tan, want_sums=True, sums=sums, want_counts=False)
    ev_two[1].record()
    for _ in range(args.steps):
        D.finalize(None, res, sums=sums, normalize=True, want_cells=True)
    ev_two[2].record()
    torch.cuda.synchronize()
    t_ray = ev[2].elapsed_time(ev[3]) / args.steps * 1e-3
    t_build = ev[4].elapsed_time(ev[5]) / args.steps * 1e-3
    t_vox = ev_two[0].elapsed_time(ev_two[1]) / args.steps * 1e-3
    t_fin = ev_two[1].elapsed_time(ev_two[2]) / args.steps * 1e-3
    # the built volume equals the two-call path's
    two = D.finalize(None, res, sums=sums, normalize=True, want_cells=True)
    build_matches = all(bool(torch.equal(two[k], vols[1][k])) for k in ("density", "tangents", "cells", "occupancy"))

    out = D.raymarch(vol, res, cam, light, sp, rp, W, H, want_depth=True, want_info=True)
    info = out["info"].reshape(-1, 4).cpu().numpy()
    frags = int((info[:, 0] == 1).sum())
    # covered = the centre ray enters the box, so the march ran (state 1 or 2)
    covered_mask = info[:, 0] != 0
    covered = int(covered_mask.sum())
    # every covered ray marches S steps; fragments also march S ADSM steps,
    # 8 AO taps and one tangent tap
    ray_bytes = 8 * S * covered + 8 * (S + 8) * frags + 32 * frags + 16 * covered
    samples = int((sums.view(-1, 2)[:, 1] & 0xFFFFFFFF).sum().item())  # the packed counts
    # every sample of the two-call voxelizer is one device-wide line request
    # (two 8-byte adds, paired lanes); the build adds them in LDS instead
    peak = atomic_peak(dev)
    voxelize_atomics = {"requests_per_s": samples / t_vox, "peak_uniform_requests_per_s": peak["u32_uniform"],
                        "frac": samples / t_vox / peak["u32_uniform"],
                        "peak_lane_contiguous_atomics_per_s": peak["u32_stream"],
                        "note": "the two-call path's voxelize_segments (voxelize_ms, including zeroing the 256 MiB "
                                "accumulator); the one-call build issues no device-wide voxel atomics"}
    # algorithmic bytes of the volume build: 12 B per vertex, 4 B per strand
    # offset, and per voxel the 4-byte count clear, the finalize's 4-byte
    # count read and its 1-byte density write
    nvert = int(xyz.shape[0])
    vox_bytes = 12 * nvert + 4 * (style.n_strands + 1) + 9 * nvox
    vox_hbm = vox_bytes / t_build / 1e9
    voxelize_roofline = {"algorithmic_bytes": vox_bytes, "seconds": t_build, "achieved_gbs": vox_hbm,
                         "hbm_frac": vox_hbm / HBM_PEAK_GBS, "build_ms": t_build * 1e3,
                         "two_call_voxelize_ms": t_vox * 1e3, "two_call_finalize_ms": t_fin * 1e3,
                         "build_speedup_vs_two_call": (t_vox + t_fin) / t_build,
                         "build_matches_two_call": build_matches,
                         "note": "vkhr_volume_build_segments (brick-binned, voxel sums in LDS; normalize, cells and "
                                 "the brick distance field included) against the algorithmic bytes of the volume "
                                 "build; the DRAM it moves is in the profile summary"}

    result = {
        "metric": METRIC,
        "value": covered / t_step / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "serial_ms_per_step": t_serial * 1e3,
        "pipelining": {"steps_in_flight": 2, "streams": 2, "images_match_serial": images_match,
                       "note": "ms_per_step is the whole job's time over K steps with two steps in flight on two "
                               "default-priority streams (double-buffered accumulator, volume and image); "
                               "serial_ms_per_step is one step at a time on one stream"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded hemisphere hair, generated in-process)",
        "config": {"workload": "136,320 strands x 13 vertices -> segment voxelization 256^3 -> finalize -> "
                               "raymarch 1920x1080, 512 steps, Kajiya-Kay + ADSM + AO",
                   "segments": n_seg, "voxel_samples": samples, "covered_rays": covered, "fragments": frags},
        "raymarch_mrays_per_s": covered / t_ray / 1e6,
        "raymarch_ms": t_ray * 1e3,
        "segments_per_s": n_seg / t_build,
        "build_ms": t_build * 1e3,
        "voxelize_ms": t_vox * 1e3,
        "voxelize_atomics": voxelize_atomics,
        "voxelize_roofline": voxelize_roofline,
        "load": {"sort_strands_ms": t_sort * 1e3,
                 "note": "once per style at load, outside the step: the index pairs strand by strand, the strands "
                         "in the Morton order of their roots, which the step's build draws; the strand passes "
                         "(raster, shadow, tracer) draw the pairs in the Morton order of their midpoints"},
        "roofline": raymarch_roofline(t_ray, ray_bytes),
    }
    if not args.no_extras:
        t_sh, t_ra, t_re, n_frag = time_raster(dev, style, xyz, idx_mid, tan, vol, res, scene, cam, light, bounds)
        result["raster"] = raster_result(t_sh, t_ra, t_re, n_frag, n_seg, W, H, peak["u32_uniform"])
        frames = time_reference_frames(dev, style, xyz, idx_mid, tan, vol, res, bounds)
        result["reference_setup_720p"] = {
            "workload": "136,320 strands x 13 vertices, 1280x720, 1024^2 3x3 PCF shadow map, 256^3, 512 steps",
            "rasterizer_frame_ms": frames["rasterizer"], "raymarcher_frame_ms": frames["raymarcher"],
            "published": "ponytail rasterized 7 ms per frame on a Radeon Pro WX 9100 (readme.md:8); the "
                         "raymarcher about 2x faster near (docs/vkhr.tex:403) -- other hardware, not a ratio"}
        result["raytrace"] = time_trace(dev, style, xyz, idx_mid, tan, scene, bounds)
        result["ao_ground_truth"] = ao_ground_truth(dev, style, xyz, idx_mid, tan, bounds, res)
        t_f, t_b = time_adjoint(dev)
        # (t_fwd + t_bwd) / t_fwd: the forward and the backward of one
        # gradient evaluation against the forward alone (target <= 3)
        result["adjoint_fwd_ratio"] = (t_f + t_b) / t_f
        result["adjoint"] = {"workload": "128^3 float32 density, 512x512, 256 steps", "forward_ms": t_f,
                             "adjoint_ms": t_b, "bwd_over_fwd": t_b / t_f,
                             "note": "adjoint_fwd_ratio = (forward_ms + adjoint_ms) / forward_ms"}
        result["batched"] = batched_result(*time_batch(dev))
        t_plain, t_batch = time_adjoint_batch(dev)
        result["batched_adjoint"] = {"workload": "8 styles x 8 views, 128^3 float32, 512x512, 256 steps, "
                                                 "forward + L2 loss gradient + adjoint",
                                     "per_render_ms": t_plain, "batched_graph_ms": t_batch,
                                     "speedup": t_plain / t_batch}
    if not args.no_cpu_baseline:
        vol_np = dict(density=vol["density"].cpu().numpy(), tangents=vol["tangents"].cpu().numpy())
        b = {k: np.asarray(v) for k, v in bounds.items()}
        cov_pix = np.nonzero(covered_mask)[0]
        result["cpu_baseline"] = cpu_baselines(style, b, res, vol_np, (cam, light, sp, rp, W, H), cov_pix,
                                               args.cpu_budget)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
