// Device bodies of the voxelizers (src/vkhr/scene_graph/hair_style.cc:
// 257-342), shared by the per-style and the batched kernels.  Translation
// units including this header are compiled with -ffp-contract=off: the DDA
// must round exactly as the reference's float arithmetic does.
#pragma once

#include "common.hpp"

#include <cmath>

namespace vkhr {

// Longest DDA walked: 2^24 samples (a segment inside a 2^24-voxel box has
// at most 2^24 of them along its longest axis).
constexpr float kMaxSteps = 16777216.0f;

struct VoxelGrid {
    float ox, oy, oz;       // AABB origin
    float vx, vy, vz;       // voxel size = size / resolution
    float mx, my, mz;       // resolution - 1 (float, as glm::min's operand)
    uint32_t nx, ny, nz;
};

// Voxel coordinates of a sample in voxel units.
__device__ __forceinline__ uint3 voxel_coords(const VoxelGrid& g, float px, float py, float pz) {
    // glm::min(glm::floor(p), resolution - 1): the upper clamp of the
    // reference; the lower clamp at 0 only guards samples that leave the box
    // through float rounding (undefined behaviour in the reference).
    float fx = fminf(floorf(px), g.mx);
    float fy = fminf(floorf(py), g.my);
    float fz = fminf(floorf(pz), g.mz);
    int ix = fx > 0.0f ? int(fx) : 0;
    int iy = fy > 0.0f ? int(fy) : 0;
    int iz = fz > 0.0f ? int(fz) : 0;
    return make_uint3(uint32_t(ix), uint32_t(iy), uint32_t(iz));
}

__device__ __forceinline__ uint32_t voxel_of(const VoxelGrid& g, float px, float py, float pz) {
    const uint3 c = voxel_coords(g, px, py, pz);
    return c.x + g.nx * (c.y + g.ny * c.z);
}

// The DDA of one index pair (hair_style.cc:311-329): root and per-sample
// direction in voxel units and the remaining step count; the samples are
// taken while steps > 0, each one decrementing steps by 1 and then
// advancing root by the direction (advance()).
struct SegmentDda {
    float rx, ry, rz, dx, dy, dz, steps;

    __device__ __forceinline__ void advance() {
        rx += dx;
        ry += dy;
        rz += dz;
    }
};

// Vertex i of a packed float3 array as one 12-byte load (three dwords of
// 4-byte alignment) rather than three.
typedef float packed_float3 __attribute__((ext_vector_type(3), aligned(4)));

__device__ __forceinline__ float3 load_vertex(const float* __restrict__ p, uint32_t i) {
    const packed_float3 v = *reinterpret_cast<const packed_float3*>(p + 3u * i);
    return make_float3(v.x, v.y, v.z);
}

// The set-up of the reference's DDA; a malformed pair, or a non-finite or
// absurd length (vertices far outside the given box, which would never
// finish), gives no samples (steps = 0).
__device__ __forceinline__ SegmentDda segment_dda(const VoxelGrid& g, const float* __restrict__ xyz,
                                                  uint32_t n_vertices, uint32_t ia, uint32_t ib) {
    SegmentDda w = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (!(ia < n_vertices && ib < n_vertices)) return w;
    float3 a = load_vertex(xyz, ia);
    float3 b = load_vertex(xyz, ib);
    // root / tip in voxel units (hair_style.cc:312-313)
    w.rx = (a.x - g.ox) / g.vx; w.ry = (a.y - g.oy) / g.vy; w.rz = (a.z - g.oz) / g.vz;
    float tx = (b.x - g.ox) / g.vx, ty = (b.y - g.oy) / g.vy, tz = (b.z - g.oz) / g.vz;
    w.dx = tx - w.rx; w.dy = ty - w.ry; w.dz = tz - w.rz;
    // steps = compMax(abs(direction)); direction /= steps (:315-317)
    w.steps = fmaxf(fmaxf(fabsf(w.dx), fabsf(w.dy)), fabsf(w.dz));
    if (w.steps <= kMaxSteps) {
        w.dx /= w.steps;
        w.dy /= w.steps;
        w.dz /= w.steps;
    } else {
        w.steps = 0.0f;
    }
    return w;
}

// Tangent accumulation in fixed point, 2^-16 per unit, packed with the count
// in two 64-bit integer sums per voxel: word 0 = x * 2^32 + y (sign-extended
// fields), word 1 = z * 2^32 + count.  Integer atomics make the sums
// independent of the order in which threads add them; word 0 decodes exactly
// while |sum y| < 2^31, i.e. for fewer than 32768 contributions to one voxel,
// and the low half of word 1 is the exact count whatever z does.  Two
// atomics per sample instead of three (count, x/y, z): device atomics are
// bound by the number of line requests.
constexpr float kTangentScale = 65536.0f;

__device__ __forceinline__ long long tangent_fixed(float v) {
    v = fminf(fmaxf(v, -1.0f), 1.0f);
    return (long long)rintf(v * kTangentScale);  // NaN components become 0 via the clamp
}

// The two packed words of a sample with tangent t (count 1).
__device__ __forceinline__ void packed_words(float3 t, unsigned long long* w0, unsigned long long* w1) {
    long long x = tangent_fixed(t.x), y = tangent_fixed(t.y), z = tangent_fixed(t.z);
    *w0 = (unsigned long long)(x) * (1ull << 32) + (unsigned long long)(y);
    *w1 = (unsigned long long)(z) * (1ull << 32) + 1ull;
}

// Adds every lane's sample to its packed accumulator entry (target =
// &acc[2v], or null for no sample), two lanes per sample: in round r lane L
// adds word (L & 1) of the sample of lane (L >> 1) + 32r, so that the 16
// contiguous bytes of one entry travel in one line request.  Every lane of
// the wave must call it (shuffles).
__device__ __forceinline__ void add_packed_wave(long long* target, unsigned long long w0, unsigned long long w1) {
    const unsigned long long live = __ballot(target != nullptr);
    if (!live) return;
    const int lane = __lane_id();
    const unsigned long long addr = reinterpret_cast<unsigned long long>(target);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        if (!((live >> (32 * r)) & 0xffffffffull)) continue;
        const int src = (lane >> 1) + 32 * r;
        const unsigned long long a = __shfl(addr, src, kWave);
        const unsigned long long s0 = __shfl(w0, src, kWave), s1 = __shfl(w1, src, kWave);
        if (a) atomicAdd(reinterpret_cast<unsigned long long*>(a) + (lane & 1), (lane & 1) ? s1 : s0);
    }
}

// The count of voxel v from word 1 of the packed accumulator.
__device__ __forceinline__ uint32_t packed_count(const long long* __restrict__ acc, size_t v) {
    return uint32_t((unsigned long long)(acc[2 * v + 1]) & 0xffffffffull);
}

// One vertex per lane (hair_style.cc:272-281): count, or count and tangent
// packed.  Wave-cooperative: every lane calls it, has = false for none.
__device__ __forceinline__ void voxelize_vertex(const VoxelGrid& g, const float* __restrict__ xyz, bool has,
                                                uint32_t i, const float* __restrict__ tangents,
                                                uint32_t* __restrict__ counts, long long* __restrict__ sums) {
    uint32_t v = 0;
    if (has) {
        float3 p = load_xyz(xyz, i);
        // (vertex - origin) / voxel_size, per component (hair_style.cc:274)
        float px = (p.x - g.ox) / g.vx;
        float py = (p.y - g.oy) / g.vy;
        float pz = (p.z - g.oz) / g.vz;
        v = voxel_of(g, px, py, pz);
        if (!sums) atomicAdd(&counts[v], 1u);
    }
    unsigned long long w0 = 0, w1 = 0;
    if (has && sums) packed_words(load_xyz(tangents, i), &w0, &w1);
    add_packed_wave(has && sums ? &sums[2ull * v] : nullptr, w0, w1);
}

// One index pair per lane (hair_style.cc:311-329): the sequential float
// DDA.  Wave-cooperative (the lanes step together until the longest
// segment of the wave is done): every lane calls it, has = false for none.
__device__ __forceinline__ void voxelize_segment(const VoxelGrid& g, const float* __restrict__ xyz,
                                                 uint32_t n_vertices, bool has, uint32_t ia, uint32_t ib,
                                                 const float* __restrict__ tangents, uint32_t* __restrict__ counts,
                                                 long long* __restrict__ sums) {
    SegmentDda w = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    unsigned long long w0 = 0, w1 = 0;
    if (has) {
        w = segment_dda(g, xyz, n_vertices, ia, ib);
        if (sums && w.steps > 0.0f) packed_words(load_xyz(tangents, ia), &w0, &w1);
    }
    // while (steps-- > 0.0f): ceil(steps) samples, root accumulated (:319-328)
    while (__ballot(w.steps > 0.0f)) {
        const bool on = w.steps > 0.0f;
        uint32_t v = 0;
        if (on) {
            w.steps -= 1.0f;
            v = voxel_of(g, w.rx, w.ry, w.rz);
            if (!sums) atomicAdd(&counts[v], 1u);
        }
        add_packed_wave(on && sums ? &sums[2ull * v] : nullptr, w0, w1);
        if (on) w.advance();
    }
}

// Voxel grid of a style (hair_style.cc:258-268, 297-307); validates the box.
inline int make_grid(const vkhr_aabb* bounds, uint32_t nx, uint32_t ny, uint32_t nz, VoxelGrid* g) {
    VKHR_REQUIRE(bounds, VKHR_ERR_INVALID, "bounds is null");
    int rc = check_volume(nx, ny, nz, false, "voxel grid");
    if (rc) return rc;
    VKHR_REQUIRE(uint64_t(nx) * ny * nz <= (1ull << 24), VKHR_ERR_OVERFLOW,
                 "volume %ux%ux%u exceeds 2^24 voxels (float linear index no longer exact)", nx, ny, nz);
    for (int c = 0; c < 3; ++c)
        VKHR_REQUIRE(bounds->size[c] > 0.0f && std::isfinite(bounds->size[c]), VKHR_ERR_INVALID,
                     "AABB size[%d] = %g: a flat or non-finite box has no voxel grid", c, bounds->size[c]);
    g->ox = bounds->origin[0];
    g->oy = bounds->origin[1];
    g->oz = bounds->origin[2];
    // glm::vec3 size / glm::vec3 resolution (hair_style.cc:268, 307)
    g->vx = bounds->size[0] / float(nx);
    g->vy = bounds->size[1] / float(ny);
    g->vz = bounds->size[2] / float(nz);
    g->mx = float(nx) - 1.0f;
    g->my = float(ny) - 1.0f;
    g->mz = float(nz) - 1.0f;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    return VKHR_OK;
}

}  // namespace vkhr
