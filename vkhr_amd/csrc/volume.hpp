// Device bodies of the volume finalization (hair_style.cc:283-291,
// 331-357), shared by the per-style and the batched kernels.  Translation
// units including this header are compiled with -ffp-contract=off.
#pragma once

#include "render.hpp"

namespace vkhr {

// The count of voxel i: from counts, or, without them, from word 1 of the
// packed accumulator (voxel.hpp).
__device__ __forceinline__ uint32_t voxel_count(const uint32_t* __restrict__ counts,
                                                const long long* __restrict__ acc, size_t i) {
    if (counts) return counts[i];
    return uint32_t((unsigned long long)(acc[2 * i + 1]) & 0xffffffffull);
}

// Quantized tangent of voxel i: sum / float(count) * 127 truncated toward
// zero -- the reference's expression on sums of every contribution (the
// reference keeps only the first 255 in its serial order, so saturated
// voxels may differ).  The sums come from the fixed-point accumulator
// (voxel.hpp); beyond 32767 contributions they are not decodable and the
// tangent is 0, as for empty voxels.
__device__ __forceinline__ char4 quantize_tangent(const long long* __restrict__ acc, size_t i, uint32_t c) {
    if (!(c && c < 32768u)) return make_char4(0, 0, 0, 0);
    const long long w0 = acc[2 * i + 0], w1 = acc[2 * i + 1];
    const long long sy = (long long)(int)(unsigned)(w0 & 0xffffffffll);
    const long long sx = (w0 - sy) / (1ll << 32);
    const long long sz = w1 >> 32;  // w1 = z * 2^32 + count, 0 <= count < 2^32
    const float inv = 1.0f / 65536.0f;
    const float denom = float(c);
    const float x = float(sx) * inv / denom * 127.0f;
    const float y = float(sy) * inv / denom * 127.0f;
    const float z = float(sz) * inv / denom * 127.0f;
    auto cvt = [](float v) -> signed char {
        v = fminf(fmaxf(v, -127.0f), 127.0f);
        return static_cast<signed char>(static_cast<int>(v));
    };
    return make_char4(cvt(x), cvt(y), cvt(z), 0);
}

// First pass over voxel i: the saturating uint8 view min(count, 255)
// (stored and returned, for the min/max of the normalize) and the quantized
// tangent.
__device__ __forceinline__ uint32_t finalize_voxel(const uint32_t* __restrict__ counts,
                                                   const long long* __restrict__ acc, size_t i,
                                                   uint8_t* __restrict__ density, int8_t* __restrict__ tangents) {
    const uint32_t c = voxel_count(counts, acc, i);
    const uint32_t d = c < 255u ? c : 255u;
    density[i] = uint8_t(d);
    if (tangents) reinterpret_cast<char4*>(tangents)[i] = quantize_tangent(acc, i, c);
    return d;
}

// finalize_voxel of voxels i .. i+3 (i a multiple of 4, all < n) with
// vector loads and stores; folds their uint8 views into lo / hi.
__device__ __forceinline__ void finalize_quad(const uint32_t* __restrict__ counts, const long long* __restrict__ acc,
                                              size_t i, uint8_t* __restrict__ density,
                                              int8_t* __restrict__ tangents, uint32_t* lo, uint32_t* hi) {
    uint32_t c[4];
    if (counts) {
        const uint4 v = reinterpret_cast<const uint4*>(counts)[i / 4];
        c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = uint32_t((unsigned long long)(acc[2 * (i + k) + 1]) & 0xffffffffull);
    }
    uchar4 d;
    d.x = uint8_t(c[0] < 255u ? c[0] : 255u);
    d.y = uint8_t(c[1] < 255u ? c[1] : 255u);
    d.z = uint8_t(c[2] < 255u ? c[2] : 255u);
    d.w = uint8_t(c[3] < 255u ? c[3] : 255u);
    reinterpret_cast<uchar4*>(density)[i / 4] = d;
    if (tangents) {
        char4 q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = quantize_tangent(acc, i + k, c[k]);
        int4 packed;
        packed.x = *reinterpret_cast<int*>(&q[0]);
        packed.y = *reinterpret_cast<int*>(&q[1]);
        packed.z = *reinterpret_cast<int*>(&q[2]);
        packed.w = *reinterpret_cast<int*>(&q[3]);
        reinterpret_cast<int4*>(tangents)[i / 4] = packed;
    }
    const uint32_t mn = min(min(uint32_t(d.x), uint32_t(d.y)), min(uint32_t(d.z), uint32_t(d.w)));
    const uint32_t mx = max(max(uint32_t(d.x), uint32_t(d.y)), max(uint32_t(d.z), uint32_t(d.w)));
    *lo = mn < *lo ? mn : *lo;
    *hi = mx > *hi ? mx : *hi;
}

// HairStyle::Volume::normalize of one value: (d - min) * (255 / (max - min)),
// uint8 arithmetic then float, truncated; a constant volume becomes 0.
__device__ __forceinline__ uint32_t normalize_value(uint32_t d, uint32_t lo, float scaling, bool flat) {
    uint32_t shifted = (d - lo) & 0xffu;   // densities[i] -= data_min (uint8)
    float v = float(shifted) * scaling;    // densities[i] * scaling
    return flat ? 0u : (uint32_t(v) & 0xffu);
}

__host__ __device__ __forceinline__ float normalize_scaling(uint32_t lo, uint32_t hi) {
    return hi == lo ? 0.0f : 255.0f / float(int(hi) - int(lo));  // hair_style.cc:351
}

// The uint8 view -> the density the renderers sample: the normalize when
// it is on.
struct DensityMap {
    uint32_t lo;
    float scaling;
    bool flat;
    bool normalize;
};

__device__ __forceinline__ DensityMap density_map(const uint32_t* minmax, bool normalize) {
    DensityMap m;
    m.lo = normalize ? minmax[0] : 0u;
    const uint32_t hi = normalize ? minmax[1] : 255u;
    m.flat = hi == m.lo;
    m.scaling = normalize_scaling(m.lo, hi);
    m.normalize = normalize;
    return m;
}

__device__ __forceinline__ uint32_t texel_density(uint32_t d, const DensityMap& m) {
    return m.normalize ? normalize_value(d, m.lo, m.scaling, m.flat) : d;
}

// One tile of the corner-packed layout -- cells (i, j, k), i in [-1, nx-1],
// index (i+1) + (nx+1) * ((j+1) + (ny+1) * (k+1)), each the 8 texels of the
// trilinear footprint whose lower corner is (i,j,k) as one uint64 (byte b =
// texel (i+(b&1), j+(b>>1&1), k+(b>>2)), 0 outside the volume) -- from the
// final uint8 density: 64 cell columns i+1 in [64 xb, 64 xb + 64) (the lanes
// of a wave) x 4 cell rows j+1 in [4 yb, 4 yb + 4) (the waves of a block) at
// cell level k+1 = ck.  Lanes read consecutive texels and store consecutive
// cells (coalesced).  The bricks (8x8x8 cells) holding a nonzero cell are
// or-ed into occupancy (nullable; one word atomic per tile with any).  All
// threads of the block call it.
constexpr uint32_t kTileRows = 4;   // cell rows per tile (waves of a 256-thread block)

__device__ __forceinline__ void cells_tile(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny,
                                           uint32_t nz, uint32_t xb, uint32_t yb, uint32_t ck,
                                           unsigned long long* __restrict__ cells, uint32_t* occupancy,
                                           uint32_t* lds_flags) {
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t cxi = xb * kWave + lane;        // cell column i + 1
    const uint32_t cj = yb * kTileRows + wave;     // cell row j + 1
    unsigned long long packed = 0ull;
    if (cxi <= nx && cj <= ny) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {  // texel rows (j + (q&1), k + (q>>1))
            const int y = int(cj) - 1 + (q & 1), z = int(ck) - 1 + (q >> 1);
            if (y < 0 || z < 0 || y >= int(ny) || z >= int(nz)) continue;
            const uint8_t* src = density + uint32_t(y) * nx + uint32_t(z) * nx * ny;
            const int x = int(cxi) - 1;
            const uint32_t t0 = x >= 0 ? src[x] : 0u;
            const uint32_t t1 = cxi < nx ? src[cxi] : 0u;
            packed |= (unsigned long long)(t0 | (t1 << 8)) << (16 * q);
        }
        cells[size_t(cj + (ny + 1u) * ck) * (nx + 1u) + cxi] = packed;
    }
    if (!occupancy) return;
    // lanes 8g..8g+7 of a wave are the cells of brick 8 xb + g
    const unsigned long long nonzero = __ballot(packed != 0ull);
    uint32_t m = 0;
#pragma unroll
    for (int g = 0; g < 8; ++g) m |= ((nonzero >> (8 * g)) & 0xffull) ? (1u << g) : 0u;
    if (threadIdx.x == 0) *lds_flags = 0u;
    __syncthreads();
    if (lane == 0 && m) atomicOr(lds_flags, m);
    __syncthreads();
    if (threadIdx.x == 0 && *lds_flags) {
        // bricks 8 xb .. 8 xb + 7 of row (by, bz): bits first .. first + 7
        const uint32_t bxs = bricks_along(nx), bys = bricks_along(ny);
        const uint32_t first = bxs * ((cj >> kBrickShift) + bys * (ck >> kBrickShift)) + 8u * xb;
        const unsigned long long bits = (unsigned long long)(*lds_flags) << (first & 31u);
        if (uint32_t(bits)) atomicOr(&occupancy[first >> 5], uint32_t(bits));
        if (uint32_t(bits >> 32)) atomicOr(&occupancy[(first >> 5) + 1], uint32_t(bits >> 32));
    }
}

}  // namespace vkhr
