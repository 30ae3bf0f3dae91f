// Device bodies of the volume finalization (hair_style.cc:283-291,
// 331-357), shared by the per-style and the batched kernels.  Translation
// units including this header are compiled with -ffp-contract=off.
#pragma once

#include "common.hpp"

namespace vkhr {

// Saturating uint8 view of voxel i and its quantized tangent; returns the
// uint8 value.  Tangents: sum / float(count) * 127 truncated toward zero --
// the reference's expression on sums of every contribution (the reference
// keeps only the first 255 in its serial order, so saturated voxels may
// differ).  The sums come from the fixed-point accumulator (voxel.hpp);
// beyond 32767 contributions they are not decodable and the tangent is 0.
// With counts_from_acc the count is taken from the accumulator and stored
// to counts[i] (the batched path unpacks the counts here).
__device__ __forceinline__ uint32_t finalize_voxel(uint32_t* __restrict__ counts, const long long* __restrict__ acc,
                                                   size_t i, uint8_t* __restrict__ density,
                                                   int8_t* __restrict__ tangents, bool counts_from_acc = false) {
    uint32_t c;
    if (counts_from_acc) {
        c = uint32_t((unsigned long long)(acc[2 * i + 1]) & 0xffffffffull);
        counts[i] = c;
    } else {
        c = counts[i];
    }
    uint32_t d = c < 255u ? c : 255u;
    density[i] = uint8_t(d);
    if (tangents) {
        char4 q = make_char4(0, 0, 0, 0);
        if (c && c < 32768u) {
            long long w0 = acc[2 * i + 0], w1 = acc[2 * i + 1];
            long long sy = (long long)(int)(unsigned)(w0 & 0xffffffffll);
            long long sx = (w0 - sy) / (1ll << 32);
            const float inv = 1.0f / 65536.0f;
            float denom = float(c);
            float x = float(sx) * inv / denom * 127.0f;
            float y = float(sy) * inv / denom * 127.0f;
            long long sz = w1 >> 32;  // w1 = z * 2^32 + count, 0 <= count < 2^32
            float z = float(sz) * inv / denom * 127.0f;
            auto cvt = [](float v) -> signed char {
                v = fminf(fmaxf(v, -127.0f), 127.0f);
                return static_cast<signed char>(static_cast<int>(v));
            };
            q = make_char4(cvt(x), cvt(y), cvt(z), 0);
        }
        reinterpret_cast<char4*>(tangents)[i] = q;
    }
    return d;
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

// Cell c of the corner-packed layout: the 8 texels of the trilinear footprint
// whose lower corner is (i,j,k) in [-1, n-1]^3, zero outside the volume.
__device__ __forceinline__ unsigned long long pack_cell(const uint8_t* __restrict__ density, uint32_t nx,
                                                        uint32_t ny, uint32_t nz, size_t c) {
    uint32_t cx = nx + 1u, cy = ny + 1u;
    int i = int(c % cx) - 1;
    int j = int((c / cx) % cy) - 1;
    int k = int(c / (size_t(cx) * cy)) - 1;
    unsigned long long packed = 0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        int x = i + (b & 1), y = j + ((b >> 1) & 1), z = k + (b >> 2);
        if (x >= 0 && y >= 0 && z >= 0 && x < int(nx) && y < int(ny) && z < int(nz)) {
            unsigned long long t = density[size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * size_t(z))];
            packed |= t << (8 * b);
        }
    }
    return packed;
}

}  // namespace vkhr
