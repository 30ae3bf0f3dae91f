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

__device__ __forceinline__ uint32_t voxel_of(const VoxelGrid& g, float px, float py, float pz) {
    // glm::min(glm::floor(p), resolution - 1): the upper clamp of the
    // reference; the lower clamp at 0 only guards samples that leave the box
    // through float rounding (undefined behaviour in the reference).
    float fx = fminf(floorf(px), g.mx);
    float fy = fminf(floorf(py), g.my);
    float fz = fminf(floorf(pz), g.mz);
    int ix = fx > 0.0f ? int(fx) : 0;
    int iy = fy > 0.0f ? int(fy) : 0;
    int iz = fz > 0.0f ? int(fz) : 0;
    return uint32_t(ix) + g.nx * (uint32_t(iy) + g.ny * uint32_t(iz));
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

// Adds one sample to voxel v: the count alone, or count and tangent packed.
__device__ __forceinline__ void add_sample(uint32_t* __restrict__ counts, long long* __restrict__ acc, uint32_t v,
                                           float3 t) {
    if (!acc) {
        atomicAdd(&counts[v], 1u);
        return;
    }
    long long x = tangent_fixed(t.x), y = tangent_fixed(t.y), z = tangent_fixed(t.z);
    unsigned long long w0 = (unsigned long long)(x) * (1ull << 32) + (unsigned long long)(y);
    unsigned long long w1 = (unsigned long long)(z) * (1ull << 32) + 1ull;
    atomicAdd(reinterpret_cast<unsigned long long*>(&acc[2ull * v + 0ull]), w0);
    atomicAdd(reinterpret_cast<unsigned long long*>(&acc[2ull * v + 1ull]), w1);
}

// The count of voxel v from word 1 of the packed accumulator.
__device__ __forceinline__ uint32_t packed_count(const long long* __restrict__ acc, size_t v) {
    return uint32_t((unsigned long long)(acc[2 * v + 1]) & 0xffffffffull);
}

// One vertex (hair_style.cc:272-281): count + optional tangent sum.
__device__ __forceinline__ void voxelize_vertex(const VoxelGrid& g, const float* __restrict__ xyz, uint32_t i,
                                                const float* __restrict__ tangents, uint32_t* __restrict__ counts,
                                                long long* __restrict__ sums) {
    float3 p = load_xyz(xyz, i);
    // (vertex - origin) / voxel_size, per component (hair_style.cc:274)
    float px = (p.x - g.ox) / g.vx;
    float py = (p.y - g.oy) / g.vy;
    float pz = (p.z - g.oz) / g.vz;
    uint32_t v = voxel_of(g, px, py, pz);
    add_sample(counts, sums, v, sums ? load_xyz(tangents, i) : make_float3(0.0f, 0.0f, 0.0f));
}

// One index pair (hair_style.cc:311-329): the sequential float DDA.
__device__ __forceinline__ void voxelize_segment(const VoxelGrid& g, const float* __restrict__ xyz,
                                                 uint32_t n_vertices, uint32_t ia, uint32_t ib,
                                                 const float* __restrict__ tangents, uint32_t* __restrict__ counts,
                                                 long long* __restrict__ sums) {
    uint2 ab = make_uint2(ia, ib);
    if (ab.x >= n_vertices || ab.y >= n_vertices) return;  // malformed pair: skipped
    float3 a = load_xyz(xyz, ab.x);
    float3 b = load_xyz(xyz, ab.y);
    // root / tip in voxel units (hair_style.cc:312-313)
    float rx = (a.x - g.ox) / g.vx, ry = (a.y - g.oy) / g.vy, rz = (a.z - g.oz) / g.vz;
    float tx = (b.x - g.ox) / g.vx, ty = (b.y - g.oy) / g.vy, tz = (b.z - g.oz) / g.vz;
    float dx = tx - rx, dy = ty - ry, dz = tz - rz;
    // steps = compMax(abs(direction)); direction /= steps (:315-317)
    float steps = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz));
    // A non-finite or absurd length (vertices far outside the given box)
    // would never finish; such segments are skipped.
    if (!(steps <= kMaxSteps)) return;
    dx /= steps;
    dy /= steps;
    dz /= steps;
    float3 t = make_float3(0.0f, 0.0f, 0.0f);
    if (sums) t = load_xyz(tangents, ab.x);
    // while (steps-- > 0.0f): ceil(steps) samples, root accumulated (:319-328)
    while (steps > 0.0f) {
        steps -= 1.0f;
        uint32_t v = voxel_of(g, rx, ry, rz);
        add_sample(counts, sums, v, t);
        rx += dx;
        ry += dy;
        rz += dz;
    }
}

// Voxel grid of a style (hair_style.cc:258-268, 297-307); validates the box.
inline int make_grid(const vkhr_aabb* bounds, uint32_t nx, uint32_t ny, uint32_t nz, VoxelGrid* g) {
    VKHR_REQUIRE(bounds, VKHR_ERR_INVALID, "bounds is null");
    VKHR_REQUIRE(nx && ny && nz, VKHR_ERR_INVALID, "zero volume resolution");
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
