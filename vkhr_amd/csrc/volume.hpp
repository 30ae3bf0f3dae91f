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
// tangent is 0, as for empty voxels.  (sx, sy, sz): the fixed-point sums
// of the c contributions.
__device__ __forceinline__ char4 quantize_tangent_sums(long long sx, long long sy, long long sz, uint32_t c) {
    if (!(c && c < 32768u)) return make_char4(0, 0, 0, 0);
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

// The same from the two packed words of the accumulator (voxel.hpp).
__device__ __forceinline__ char4 quantize_tangent_words(long long w0, long long w1, uint32_t c) {
    if (!(c && c < 32768u)) return make_char4(0, 0, 0, 0);
    const long long sy = (long long)(int)(unsigned)(w0 & 0xffffffffll);
    const long long sx = (w0 - sy) / (1ll << 32);
    const long long sz = w1 >> 32;  // w1 = z * 2^32 + count, 0 <= count < 2^32
    return quantize_tangent_sums(sx, sy, sz, c);
}

__device__ __forceinline__ char4 quantize_tangent(const long long* __restrict__ acc, size_t i, uint32_t c) {
    if (!(c && c < 32768u)) return make_char4(0, 0, 0, 0);
    return quantize_tangent_words(acc[2 * i + 0], acc[2 * i + 1], c);
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
// vector loads and stores (the accumulator 16-byte aligned: one load per
// voxel); folds their uint8 views into lo / hi.  With acc and counts_out the
// unpacked counts are stored too.
__device__ __forceinline__ void finalize_quad(const uint32_t* __restrict__ counts, const long long* __restrict__ acc,
                                              size_t i, uint8_t* __restrict__ density,
                                              int8_t* __restrict__ tangents, uint32_t* lo, uint32_t* hi,
                                              uint32_t* __restrict__ counts_out = nullptr) {
    uint32_t c[4];
    longlong2 w[4];
    if (counts) {
        const uint4 v = reinterpret_cast<const uint4*>(counts)[i / 4];
        c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            w[k] = reinterpret_cast<const longlong2*>(acc)[i + k];
            c[k] = uint32_t((unsigned long long)(w[k].y) & 0xffffffffull);
        }
        if (counts_out) reinterpret_cast<uint4*>(counts_out)[i / 4] = make_uint4(c[0], c[1], c[2], c[3]);
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
        for (int k = 0; k < 4; ++k)
            q[k] = counts ? quantize_tangent(acc, i + k, c[k]) : quantize_tangent_words(w[k].x, w[k].y, c[k]);
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

__device__ __forceinline__ DensityMap density_map_of(uint32_t lo, uint32_t hi, bool normalize) {
    DensityMap m;
    m.lo = normalize ? lo : 0u;
    hi = normalize ? hi : 255u;
    m.flat = hi == m.lo;
    m.scaling = normalize_scaling(m.lo, hi);
    m.normalize = normalize;
    return m;
}

__device__ __forceinline__ DensityMap density_map(const uint32_t* minmax, bool normalize) {
    return density_map_of(normalize ? minmax[0] : 0u, normalize ? minmax[1] : 255u, normalize);
}

// The per-volume min/max of the finalize is kept in kMinMaxSlots slots, one
// per 64-byte line (block b updates slot b % kMinMaxSlots), so that the
// blocks' atomics do not serialize on one address.
constexpr int kMinMaxSlots = 32;
constexpr int kMinMaxStride = 16;  // words between slots
constexpr size_t kMinMaxBytes = size_t(kMinMaxSlots) * kMinMaxStride * sizeof(uint32_t);

// The passes after the uint8 view and its min/max slots (mm, nullable when
// normalize is 0): the normalize in place and the float view, then the
// corner-packed cells, their brick occupancy and the brick distance field
// (each output nullable).  Enqueued on s (volume.hip).
int finish_volume(uint8_t* density_u8, uint32_t nx, uint32_t ny, uint32_t nz, int normalize, float* density_f32,
                  uint64_t* cells, uint32_t* occupancy, const uint32_t* mm, hipStream_t s);

// The occupancy after the cells pass: the bitmask packed from the brick
// flags, then the three passes of the brick distance field (none when a
// side is too long to transform).
int finish_occupancy(uint32_t* occupancy, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t s);

// Folds (lo, hi) into a min/max slot.  A slot only moves toward the
// result, so an atomic whose value a read of the slot already reaches (or
// passes) changes nothing and is skipped -- a stale read is never more
// extreme than the slot, so at worst the atomic is issued -- which keeps
// thousands of workgroups from queueing on the same two words.
__device__ __forceinline__ void fold_minmax(uint32_t* slot, uint32_t lo, uint32_t hi) {
    if (lo < __hip_atomic_load(&slot[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&slot[0], lo);
    if (hi > __hip_atomic_load(&slot[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&slot[1], hi);
}

// Min/max over the slots; called by every lane of a wave (shuffles).
__device__ __forceinline__ DensityMap density_map_slots(const uint32_t* minmax, bool normalize) {
    const int lane = threadIdx.x % kWave;
    uint32_t lo = 255u, hi = 0u;
    if (normalize && lane < kMinMaxSlots) {
        lo = minmax[lane * kMinMaxStride];
        hi = minmax[lane * kMinMaxStride + 1];
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    return density_map_of(lo, hi, normalize);
}

__device__ __forceinline__ uint32_t texel_density(uint32_t d, const DensityMap& m) {
    return m.normalize ? normalize_value(d, m.lo, m.scaling, m.flat) : d;
}

// The brick flags the cells pass sets: one byte per brick in the last
// section of the occupancy buffer (zeroed before the pass), set to 1 with
// a plain store for a brick holding a nonzero cell.  Plain stores, unlike
// atomics on the bit words, do not queue at the memory side when the
// waves of thousands of workgroups mark the same few bricks; pack_brick_bits
// turns the flags into the bitmask afterwards, and the distance field's
// passes then reuse the bytes.
__host__ __device__ __forceinline__ uint8_t* brick_flags(uint32_t* occupancy, uint32_t nx, uint32_t ny, uint32_t nz) {
    return reinterpret_cast<uint8_t*>(occupancy + occupancy_bit_words(nx, ny, nz) + distance_words(nx, ny, nz));
}

// Bit word w of the bitmask from the brick flags (one thread per word).
__device__ __forceinline__ void pack_brick_bits(uint32_t* occupancy, uint32_t nx, uint32_t ny, uint32_t nz,
                                                uint32_t w) {
    const uint8_t* flags = brick_flags(occupancy, nx, ny, nz);
    const uint32_t n = uint32_t(brick_count(nx, ny, nz));
    uint32_t bits = 0u;
    const uint32_t b0 = 32u * w;
    if (b0 + 32u <= n) {  // the flags section is 4-byte aligned and b0 a multiple of 32
        const uint32_t* f = reinterpret_cast<const uint32_t*>(flags + b0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t v = f[k];
#pragma unroll
            for (int c = 0; c < 4; ++c) bits |= ((v >> (8 * c)) & 0xffu) ? 1u << (4 * k + c) : 0u;
        }
    } else {
        for (uint32_t b = b0; b < n; ++b) bits |= flags[b] ? 1u << (b - b0) : 0u;
    }
    occupancy[w] = bits;
}

// One block of the corner-packed layout -- cells (i, j, k), i in [-1, nx-1],
// index (i+1) + (nx+1) * ((j+1) + (ny+1) * (k+1)), each the 8 texels of the
// trilinear footprint whose lower corner is (i,j,k) as one uint64 (byte b =
// texel (i+(b&1), j+(b>>1&1), k+(b>>2)), 0 outside the volume) -- from the
// final uint8 density: cell columns i+1 in [64 xb, 64 xb + 64) (the lanes of
// a wave), cell rows j+1 in [8 by, 8 by + 8) (two per wave), cell levels
// k+1 in [slab zc, slab zc + slab) walked upwards (slab = cell_slab), so each
// level loads only its own texel slice and reuses the one below.  Lanes
// read consecutive texels and store consecutive cells.  The bricks (8x8x8
// cells) holding a nonzero cell are flagged in occupancy's brick flags
// (nullable; one store per wave, brick and brick level).  All threads of
// the block call it.
// Cell levels per block: 32, halved (down to 4) while the volume would give
// fewer than 2048 blocks, so small volumes still spread over the chip
// instead of walking long columns on a few compute units.
__host__ __device__ __forceinline__ uint32_t cell_slab(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t plane = ((nx + 1u + 63u) / 64u) * bricks_along(ny);
    uint32_t slab = 32u;
    while (slab > 4u && plane * ((nz + 1u + slab - 1u) / slab) < 2048u) slab >>= 1;
    return slab;
}

__host__ __device__ __forceinline__ uint32_t cell_blocks(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t slab = cell_slab(nx, ny, nz);
    return ((nx + 1u + 63u) / 64u) * bricks_along(ny) * ((nz + 1u + slab - 1u) / slab);
}

// kMode 0: density is the final volume.  kMode 1 and 2: density holds the
// saturating uint8 view before the normalize, and each cell (i+1, j+1, k+1)
// with i, j, k >= 0 writes its lower corner texel (i, j, k) of the final
// volume -- every texel exactly once -- to out_u8 (and out_f32, nullable),
// so the normalize needs no pass of its own: in mode 1 the normalize is the
// identity, in mode 2 every loaded texel goes through lut (256 entries,
// the map of the normalize).
template <int kMode = 0>
__device__ __forceinline__ void cells_block(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny,
                                            uint32_t nz, uint32_t block, unsigned long long* __restrict__ cells,
                                            uint32_t* occupancy, const uint8_t* lut = nullptr,
                                            uint8_t* __restrict__ out_u8 = nullptr,
                                            float* __restrict__ out_f32 = nullptr) {
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t xbs = (nx + 1u + 63u) / 64u, bys = bricks_along(ny);
    const uint32_t xb = block % xbs, by = (block / xbs) % bys, zc = block / (xbs * bys);
    const uint32_t cxi = xb * 64u + lane;                        // cell column i + 1
    const bool col = cxi <= nx;
    const uint32_t slab = cell_slab(nx, ny, nz);
    const uint32_t k0 = zc * slab, k1 = min(k0 + slab, nz + 1u);
    // the two cell rows of this wave and their texel rows j, j+1
    uint32_t cj[2];
    bool row[2];
    int ty[2][2];
    for (int r = 0; r < 2; ++r) {
        cj[r] = (by << kBrickShift) + wave + 4u * uint32_t(r);
        row[r] = cj[r] <= ny;
        ty[r][0] = int(cj[r]) - 1;
        ty[r][1] = int(cj[r]);
    }
    // texels (x-1, x) of texel row y in slice z, packed as 2 bytes (0 outside)
    auto pair = [&](int y, int z) -> uint32_t {
        if (!col || y < 0 || z < 0 || y >= int(ny) || z >= int(nz)) return 0u;
        const uint8_t* src = density + uint32_t(y) * nx + uint32_t(z) * nx * ny;
        uint32_t lo = cxi >= 1u ? src[cxi - 1u] : 0u;
        uint32_t hi = cxi < nx ? src[cxi] : 0u;
        if constexpr (kMode == 2) {  // texels outside the volume stay 0 (the border)
            lo = cxi >= 1u ? uint32_t(lut[lo]) : 0u;
            hi = cxi < nx ? uint32_t(lut[hi]) : 0u;
        }
        return lo | (hi << 8);
    };
    uint32_t below[2];  // slice k (texel z = k) of each row: rows j, j+1 -> 4 bytes
    for (int r = 0; r < 2; ++r)
        below[r] = pair(ty[r][0], int(k0) - 1) | (pair(ty[r][1], int(k0) - 1) << 16);
    uint32_t any = 0u;  // a nonzero cell in this brick level (per lane)
    // Levels go in groups of kLevelBatch whose texel loads are all issued
    // before the first is used: one memory latency per group, not per level.
    constexpr uint32_t kLevelBatch = 4;
    for (uint32_t cg = k0; cg < k1; cg += kLevelBatch) {
        uint32_t above[kLevelBatch][2];
#pragma unroll
        for (uint32_t u = 0; u < kLevelBatch; ++u)
#pragma unroll
            for (int r = 0; r < 2; ++r)
                above[u][r] = cg + u < k1 ? pair(ty[r][0], int(cg + u)) | (pair(ty[r][1], int(cg + u)) << 16) : 0u;
#pragma unroll
        for (uint32_t u = 0; u < kLevelBatch; ++u) {
            const uint32_t ck = cg + u;
            if (ck >= k1) break;  // wave-uniform
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const unsigned long long packed =
                    (unsigned long long)(below[r]) | ((unsigned long long)(above[u][r]) << 32);
                if (col && row[r]) {
                    cells[size_t(cj[r] + (ny + 1u) * ck) * (nx + 1u) + cxi] = packed;
                    any |= packed != 0ull ? 1u : 0u;
                    if constexpr (kMode != 0) {
                        // the lower corner texel (cxi-1, cj-1, ck-1): byte 0 of the cell
                        if (cxi >= 1u && cj[r] >= 1u && ck >= 1u && cxi <= nx && cj[r] <= ny && ck <= nz) {
                            const size_t t = size_t(cxi - 1u) + size_t(nx) * (size_t(cj[r] - 1u) + size_t(ny) * (ck - 1u));
                            const uint32_t v = uint32_t(packed & 0xffull);
                            out_u8[t] = uint8_t(v);
                            if (out_f32) out_f32[t] = float(v) / 255.0f;  // R8_UNORM
                        }
                    }
                }
                below[r] = above[u][r];
            }
            if (occupancy && ((ck & 7u) == 7u || ck + 1u == k1)) {  // wave-uniform: the end of a brick level
                const unsigned long long nz_lanes = __ballot(any != 0u);
                uint32_t m = 0u;
#pragma unroll
                for (int g = 0; g < 8; ++g) m |= ((nz_lanes >> (8 * g)) & 0xffull) ? (1u << g) : 0u;
                if (lane < 8u && ((m >> lane) & 1u)) {
                    // bricks 8 xb .. 8 xb + 7 of brick row (by, ck / 8), one per lane
                    const uint32_t first = bricks_along(nx) * (by + bys * (ck >> kBrickShift)) + 8u * xb;
                    brick_flags(occupancy, nx, ny, nz)[first + lane] = 1u;
                }
                any = 0u;
            }
        }
    }
}

// The same cells four per lane (quads: cell columns i+1 in [4q, 4q + 4) of
// one cell row, with one 4-byte load per texel row and level instead of
// eight byte loads), for rows of nx % 4 == 0 texels whose nx / 4 quads
// divide the wave or fill whole waves.  Per level a lane loads the texel
// words of rows j-1 and j at slice k (the slice below is kept from the
// level before) and takes texel 4q-1 from its neighbour lane's word; the
// last quad of a row also writes cell column nx.  A wave takes 64
// consecutive (row, quad) pairs of a level; workgroups of 4 waves walk a
// slab of levels, as cells_block, and flag the bricks as it does.  In
// modes 1 and 2 the lane at (row j, level k) writes texel row (j-1, k-1)
// of the final volume as one word.
__host__ __device__ __forceinline__ bool cells_by_quads(uint32_t nx) {
    const uint32_t q = nx / 4u;
    return nx % 4u == 0u && q > 0u && (64u % q == 0u || q % 64u == 0u);
}

__host__ __device__ __forceinline__ uint32_t quad_waves(uint32_t nx, uint32_t ny) {
    return ((ny + 1u) * (nx / 4u) + uint32_t(kWave) - 1u) / uint32_t(kWave);
}

__host__ __device__ __forceinline__ uint32_t quad_slab(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint32_t plane = (quad_waves(nx, ny) + 3u) / 4u;
    uint32_t slab = 32u;
    while (slab > 4u && plane * ((nz + 1u + slab - 1u) / slab) < 2048u) slab >>= 1;
    return slab;
}

// Workgroups of the cells pass (either form).
__host__ __device__ __forceinline__ uint32_t cells_pass_blocks(uint32_t nx, uint32_t ny, uint32_t nz) {
    if (!cells_by_quads(nx)) return cell_blocks(nx, ny, nz);
    const uint32_t slab = quad_slab(nx, ny, nz);
    return (quad_waves(nx, ny) + 3u) / 4u * ((nz + 1u + slab - 1u) / slab);
}

template <int kMode>
__device__ __forceinline__ uint32_t map_word(uint32_t w, const uint8_t* lut) {
    if constexpr (kMode != 2) return w;
    return uint32_t(lut[w & 0xffu]) | (uint32_t(lut[(w >> 8) & 0xffu]) << 8) |
           (uint32_t(lut[(w >> 16) & 0xffu]) << 16) | (uint32_t(lut[w >> 24]) << 24);
}

template <int kMode = 0>
__device__ __forceinline__ void cells_quads(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny,
                                            uint32_t nz, uint32_t block, unsigned long long* __restrict__ cells,
                                            uint32_t* occupancy, const uint8_t* lut = nullptr,
                                            uint8_t* __restrict__ out_u8 = nullptr,
                                            float* __restrict__ out_f32 = nullptr) {
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t nq = nx / 4u;
    const uint32_t waves = quad_waves(nx, ny), plane = (waves + 3u) / 4u;
    const uint32_t chunk = (block % plane) * 4u + wave, zc = block / plane;
    if (chunk >= waves) return;  // wave-uniform
    const uint32_t slab = quad_slab(nx, ny, nz);
    const uint32_t k0 = zc * slab, k1 = min(k0 + slab, nz + 1u);
    const uint32_t rq = chunk * uint32_t(kWave) + lane;
    const bool live = rq < (ny + 1u) * nq;
    const uint32_t cj = rq / nq, q = rq % nq;  // cell row (texel rows cj-1, cj) and quad
    const bool last = live && q + 1u == nq;   // also writes cell column nx
    const size_t row_pitch = nx, slice_pitch = size_t(nx) * ny;
    // texel words 4q .. 4q+3 of texel row y in slice z (0 outside), and
    // texel 4q-1 of the same row
    auto word = [&](int y, int z) -> uint32_t {
        if (!live || y < 0 || z < 0 || y >= int(ny) || z >= int(nz)) return 0u;
        const uint32_t w = *reinterpret_cast<const uint32_t*>(density + size_t(y) * row_pitch +
                                                               size_t(z) * slice_pitch + 4u * q);
        return map_word<kMode>(w, lut);
    };
    auto before = [&](uint32_t w, int y, int z) -> uint32_t {
        const uint32_t from_left = uint32_t(__shfl_up(int(w), 1, kWave)) >> 24;
        if (q == 0u) return 0u;
        if (lane != 0u) return from_left;
        // the wave starts inside a row (nx / 4 a multiple of 64)
        if (!live || y < 0 || z < 0 || y >= int(ny) || z >= int(nz)) return 0u;
        const uint32_t t = density[size_t(y) * row_pitch + size_t(z) * slice_pitch + 4u * q - 1u];
        return kMode == 2 ? uint32_t(lut[t]) : t;
    };
    const int y0 = int(cj) - 1, y1 = int(cj);
    uint32_t bw[2], bp[2];  // slice k0 - 1: words and texels 4q-1 of rows y0, y1
    bw[0] = word(y0, int(k0) - 1);
    bw[1] = word(y1, int(k0) - 1);
    bp[0] = before(bw[0], y0, int(k0) - 1);
    bp[1] = before(bw[1], y1, int(k0) - 1);
    uint32_t any = 0u, any_last = 0u;
    constexpr uint32_t kLevelBatch = 4;
    for (uint32_t cg = k0; cg < k1; cg += kLevelBatch) {
        uint32_t aw[kLevelBatch][2];
#pragma unroll
        for (uint32_t u = 0; u < kLevelBatch; ++u) {
            aw[u][0] = cg + u < k1 ? word(y0, int(cg + u)) : 0u;
            aw[u][1] = cg + u < k1 ? word(y1, int(cg + u)) : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kLevelBatch; ++u) {
            const uint32_t ck = cg + u;
            if (ck >= k1) break;  // wave-uniform
            uint32_t ap[2];
            ap[0] = before(aw[u][0], y0, int(ck));
            ap[1] = before(aw[u][1], y1, int(ck));
            // bytes 4q-1 .. 4q+3 of the four texel rows, as 40-bit strings
            const unsigned long long s00 = bp[0] | ((unsigned long long)bw[0] << 8);
            const unsigned long long s01 = bp[1] | ((unsigned long long)bw[1] << 8);
            const unsigned long long s10 = ap[0] | ((unsigned long long)aw[u][0] << 8);
            const unsigned long long s11 = ap[1] | ((unsigned long long)aw[u][1] << 8);
            unsigned long long c[4];
#pragma unroll
            for (uint32_t m = 0; m < 4; ++m) {
                const uint32_t sh = 8u * m;
                const uint32_t lo = uint32_t((s00 >> sh) & 0xffffull) | (uint32_t((s01 >> sh) & 0xffffull) << 16);
                const uint32_t hi = uint32_t((s10 >> sh) & 0xffffull) | (uint32_t((s11 >> sh) & 0xffffull) << 16);
                c[m] = (unsigned long long)lo | ((unsigned long long)hi << 32);
            }
            if (live) {
                unsigned long long* dst = cells + (size_t(cj) + size_t(ny + 1u) * ck) * (nx + 1u) + 4u * q;
                dst[0] = c[0];
                dst[1] = c[1];
                dst[2] = c[2];
                dst[3] = c[3];
                any |= (c[0] | c[1] | c[2] | c[3]) != 0ull ? 1u : 0u;
                if (last) {  // cell column nx: texels nx-1 and the zero border
                    const uint32_t lo = (bw[0] >> 24) | ((bw[1] >> 24) << 16);
                    const uint32_t hi = (aw[u][0] >> 24) | ((aw[u][1] >> 24) << 16);
                    const unsigned long long e = (unsigned long long)lo | ((unsigned long long)hi << 32);
                    dst[4] = e;
                    any_last |= e != 0ull ? 1u : 0u;
                }
                if constexpr (kMode != 0) {
                    // texel row (cj-1, ck-1) of the final volume, once per texel
                    if (cj >= 1u && ck >= 1u && cj <= ny && ck <= nz) {
                        const size_t t = size_t(cj - 1u) * row_pitch + size_t(ck - 1u) * slice_pitch + 4u * q;
                        *reinterpret_cast<uint32_t*>(out_u8 + t) = bw[0];
                        if (out_f32)
                            *reinterpret_cast<float4*>(out_f32 + t) =
                                make_float4(float(bw[0] & 0xffu) / 255.0f, float((bw[0] >> 8) & 0xffu) / 255.0f,
                                            float((bw[0] >> 16) & 0xffu) / 255.0f, float(bw[0] >> 24) / 255.0f);
                    }
                }
            }
            bw[0] = aw[u][0];
            bw[1] = aw[u][1];
            bp[0] = ap[0];
            bp[1] = ap[1];
            if (occupancy && ((ck & 7u) == 7u || ck + 1u == k1)) {  // wave-uniform: the end of a brick level
                // the two quads of a brick column share its bit
                any |= uint32_t(__shfl_xor(int(any), 1, kWave));
                if (live && any && (q & 1u) == 0u) {
                    const uint32_t b = (q >> 1) + bricks_along(nx) * ((cj >> kBrickShift) +
                                                                      bricks_along(ny) * (ck >> kBrickShift));
                    brick_flags(occupancy, nx, ny, nz)[b] = 1u;
                }
                if (any_last) {
                    const uint32_t b = (nx >> kBrickShift) + bricks_along(nx) * ((cj >> kBrickShift) +
                                                                                 bricks_along(ny) * (ck >> kBrickShift));
                    brick_flags(occupancy, nx, ny, nz)[b] = 1u;
                }
                any = 0u;
                any_last = 0u;
            }
        }
    }
}

// The cells pass of one workgroup in whichever form fits nx
// (cells_pass_blocks workgroups).
template <int kMode = 0>
__device__ __forceinline__ void cells_pass(const uint8_t* __restrict__ density, uint32_t nx, uint32_t ny, uint32_t nz,
                                           uint32_t block, unsigned long long* __restrict__ cells, uint32_t* occupancy,
                                           const uint8_t* lut = nullptr, uint8_t* __restrict__ out_u8 = nullptr,
                                           float* __restrict__ out_f32 = nullptr) {
    if (cells_by_quads(nx))
        cells_quads<kMode>(density, nx, ny, nz, block, cells, occupancy, lut, out_u8, out_f32);
    else
        cells_block<kMode>(density, nx, ny, nz, block, cells, occupancy, lut, out_u8, out_f32);
}

// The brick distance field of the occupancy buffer (occupancy_words): a
// chessboard distance transform of the occupied bricks, capped at
// kDistanceCap, as three passes of a one-dimensional transform (along x,
// then y, then z; for the L-infinity distance the passes compose).  A pass
// takes one line of bricks per wave: the wave loads the line into its LDS
// row, then each brick takes min over offsets |o| <= cap - 1 of max(|o|,
// value at o).  Pass 0 reads the bitmask (0 occupied, cap empty) and
// writes the line bytes; pass 1 updates the bytes in place (a line is read
// whole before it is written); pass 2 reads the bytes and or-s the nibbles
// (zeroed before the cells pass).  Lines longer than kDistanceLineMax
// bricks are not transformed: the nibbles stay 0 ("occupied"), so nothing
// is skipped.
// Lines of pass `axis`: the bricks of the other two axes.
__host__ __device__ __forceinline__ uint32_t distance_lines(uint32_t nx, uint32_t ny, uint32_t nz, int axis) {
    const uint32_t bx = bricks_along(nx), by = bricks_along(ny), bz = bricks_along(nz);
    return axis == 0 ? by * bz : axis == 1 ? bx * bz : bx * by;
}

// One line of pass `axis` per wave (inactive waves still reach the
// barrier); lds: this wave's kDistanceLineMax words.  All threads of the
// block call it.
__device__ __forceinline__ void distance_line(uint32_t* occupancy, uint32_t nx, uint32_t ny, uint32_t nz, int axis,
                                              uint32_t line, bool active, uint32_t* lds) {
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t bx = bricks_along(nx), by = bricks_along(ny), bz = bricks_along(nz);
    const uint32_t len = axis == 0 ? bx : axis == 1 ? by : bz;
    uint32_t base = 0u, stride = 1u;
    if (axis == 0) {
        base = bx * line;                                   // line = y + by * z
    } else if (axis == 1) {
        base = (line % bx) + bx * by * (line / bx);         // line = x + bx * z
        stride = bx;
    } else {
        base = line;                                        // line = x + bx * y
        stride = bx * by;
    }
    const uint32_t* bits = occupancy;
    uint32_t* nibbles = occupancy + occupancy_bit_words(nx, ny, nz);
    uint8_t* bytes = reinterpret_cast<uint8_t*>(nibbles + distance_words(nx, ny, nz));
    constexpr uint32_t kPer = kDistanceLineMax / kWave;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t e = lane + k * kWave;
        if (active && e < len) {
            const uint32_t b = base + e * stride;
            lds[e] = axis == 0 ? (((bits[b >> 5] >> (b & 31u)) & 1u) ? 0u : kDistanceCap) : uint32_t(bytes[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t e = lane + k * kWave;
        if (!(active && e < len)) continue;
        uint32_t m = lds[e];
        for (uint32_t o = 1; o < kDistanceCap && m > o; ++o) {
            if (e >= o) m = min(m, max(o, lds[e - o]));
            if (e + o < len) m = min(m, max(o, lds[e + o]));
        }
        const uint32_t b = base + e * stride;
        if (axis < 2)
            bytes[b] = uint8_t(m);
        else if (m)
            atomicOr(&nibbles[b >> 3], m << ((b & 7u) << 2));
    }
}

}  // namespace vkhr
