// Device pieces shared by the raymarcher and the strand rasterizer:
// exact trilinear density sampling, packUnorm4x8, and the per-pixel linked
// list insert with wave-aggregated node allocation.
#pragma once

#include "common.hpp"

namespace vkhr {

constexpr uint32_t kPpllNull = 0xffffffffu;

// The density volume as the kernels read it: corner-packed uint8 cells
// (one 8-byte gather per trilinear tap) or a float volume, and the brick
// distance field of the cells (4 bits per 8x8x8 cells, nullable; see
// occupancy_words).
struct DensityVolume {
    const unsigned long long* cells;
    const float* f32;
    const uint32_t* distance;
    int nx, ny, nz;
    uint32_t bx, by;         // bricks along x and y
    uint32_t dist_words;     // words of the distance nibbles
};

constexpr uint32_t kBrickShift = 3;              // 8x8x8 cells per brick
constexpr uint32_t kDistanceCap = 15;            // largest brick distance a nibble holds
constexpr uint32_t kDistanceLdsWords = 8192;     // largest distance field staged in LDS (32 KiB, 65536 bricks)

__host__ __device__ __forceinline__ uint32_t bricks_along(uint32_t n) { return (n + 8u) >> kBrickShift; }

__host__ __device__ __forceinline__ size_t brick_count(uint32_t nx, uint32_t ny, uint32_t nz) {
    return size_t(bricks_along(nx)) * bricks_along(ny) * bricks_along(nz);
}

// The occupancy buffer of a volume, in uint32 words, holds three arrays:
//   [0, bits)                  one bit per brick, set when a cell of the brick
//                              holds a nonzero texel (brick (bx,by,bz) = cell
//                              (i+1,j+1,k+1) >> 3 at index bx + BX*(by + BY*bz));
//   [bits, bits + nibbles)     4 bits per brick (brick b at bits 4(b%8) of word
//                              b/8): 0 for an occupied brick, else the
//                              chessboard distance in bricks to the nearest
//                              occupied one, capped at kDistanceCap -- every
//                              brick within distance - 1 of it is empty;
//   [bits + nibbles, end)      one byte per brick of working space for the
//                              distance passes.
__host__ __device__ __forceinline__ size_t occupancy_bit_words(uint32_t nx, uint32_t ny, uint32_t nz) {
    return (brick_count(nx, ny, nz) + 31) / 32;
}
__host__ __device__ __forceinline__ size_t distance_words(uint32_t nx, uint32_t ny, uint32_t nz) {
    return (brick_count(nx, ny, nz) + 7) / 8;
}
__host__ __device__ __forceinline__ size_t occupancy_words(uint32_t nx, uint32_t ny, uint32_t nz) {
    return occupancy_bit_words(nx, ny, nz) + distance_words(nx, ny, nz) + (brick_count(nx, ny, nz) + 3) / 4;
}

// Longest line of bricks the distance passes transform (volume.hpp): a
// volume with a longer one keeps its nibbles at 0, and nothing is skipped.
constexpr uint32_t kDistanceLineMax = 256;

__host__ __device__ __forceinline__ bool distance_field_fits(uint32_t nx, uint32_t ny, uint32_t nz) {
    return bricks_along(nx) <= kDistanceLineMax && bricks_along(ny) <= kDistanceLineMax &&
           bricks_along(nz) <= kDistanceLineMax;
}

// Copies the distance field into LDS when it fits; returns whether it did
// (otherwise no sample is skipped).  All threads of the block call it.
__device__ __forceinline__ bool stage_distance(const DensityVolume& v, uint32_t* lds) {
    const bool use = v.distance && v.cells && !v.f32 && v.dist_words <= kDistanceLdsWords;
    if (!use) return false;
    for (uint32_t i = threadIdx.x; i < v.dist_words; i += blockDim.x) lds[i] = v.distance[i];
    __syncthreads();
    return true;
}

// LDS bytes stage_distance needs for a volume (0: it stages nothing).
__host__ __forceinline__ size_t distance_lds_bytes(const DensityVolume& v) {
    return v.distance && v.cells && !v.f32 && v.dist_words <= kDistanceLdsWords ? size_t(v.dist_words) * 4 : 0;
}

// The LDS copy of the distance field a march reads (use = false: none).
struct OccupancyView {
    const uint32_t* lds;
    bool use;
};

// The distance nibble of brick b.
__device__ __forceinline__ uint32_t brick_distance(const OccupancyView& occ, uint32_t b) {
    return (occ.lds[b >> 3] >> ((b & 7u) << 2)) & 0xfu;
}

// texture(density, uvw).r at texel coordinate (x, y, z) = uvw * n - 0.5:
// exact weights (the fraction of a float64 coordinate), zero border.
__device__ __forceinline__ float density_at(const DensityVolume& v, double x, double y, double z) {
    if (!(x >= -1.0 && y >= -1.0 && z >= -1.0 && x < double(v.nx) && y < double(v.ny) && z < double(v.nz)))
        return 0.0f;
    double fx = floor(x), fy = floor(y), fz = floor(z);
    float wx = float(x - fx), wy = float(y - fy), wz = float(z - fz);
    int ix = int(fx), iy = int(fy), iz = int(fz);
    float b[8];
    if (v.f32) {
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            int tx = ix + (n & 1), ty = iy + ((n >> 1) & 1), tz = iz + (n >> 2);
            bool in = tx >= 0 && ty >= 0 && tz >= 0 && tx < v.nx && ty < v.ny && tz < v.nz;
            b[n] = in ? v.f32[size_t(tx) + size_t(v.nx) * (size_t(ty) + size_t(v.ny) * size_t(tz))] : 0.0f;
        }
    } else {
        size_t cx = size_t(v.nx) + 1, cy = size_t(v.ny) + 1;
        unsigned long long c = v.cells[size_t(ix + 1) + cx * (size_t(iy + 1) + cy * size_t(iz + 1))];
        uint32_t lo = uint32_t(c), hi = uint32_t(c >> 32);
        b[0] = float(lo & 0xffu); b[1] = float((lo >> 8) & 0xffu);
        b[2] = float((lo >> 16) & 0xffu); b[3] = float(lo >> 24);
        b[4] = float(hi & 0xffu); b[5] = float((hi >> 8) & 0xffu);
        b[6] = float((hi >> 16) & 0xffu); b[7] = float(hi >> 24);
    }
    float c00 = fmaf(wx, b[1] - b[0], b[0]);
    float c10 = fmaf(wx, b[3] - b[2], b[2]);
    float c01 = fmaf(wx, b[5] - b[4], b[4]);
    float c11 = fmaf(wx, b[7] - b[6], b[6]);
    float c0 = fmaf(wy, c10 - c00, c00);
    float c1 = fmaf(wy, c11 - c01, c01);
    float r = fmaf(wz, c1 - c0, c0);
    return v.f32 ? r : r * (1.0f / 255.0f);
}

// density_at for the march loops, whose samples lie inside the support
// interval of their ray (coordinates within a margin of [-1, n), so the
// float-to-int conversions are exact): the range test x in [-1, n) is done
// on the integer cell, floor(x) in [-1, n-1], and folded with the brick
// occupancy into one predicate (one branch around the gather).  Indices
// use 24-bit multiplies: the callers bound n + 1 and (ny + 1)(nz + 1)
// below 2^24 (and the cell count below 2^32).  kF32 selects the float
// volume.
// With the distance field staged (occ.use), a sample whose brick is empty
// (distance nibble above 0) reads 0 without its gather: every texel of its
// footprint is 0.
template <bool kF32>
__device__ __forceinline__ float march_tap(const DensityVolume& v, double x, double y, double z, OccupancyView occ) {
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int ix = int(fx), iy = int(fy), iz = int(fz);
    const uint32_t cx1 = uint32_t(ix + 1), cy1 = uint32_t(iy + 1), cz1 = uint32_t(iz + 1);
    bool live = (cx1 <= uint32_t(v.nx)) & (cy1 <= uint32_t(v.ny)) & (cz1 <= uint32_t(v.nz));
    if (occ.use) {
        const uint32_t b = (cx1 >> kBrickShift) +
                           __umul24(v.bx, (cy1 >> kBrickShift) + __umul24(v.by, cz1 >> kBrickShift));
        live = live & (brick_distance(occ, live ? b : 0u) == 0u);  // b is in range only for live cells
    }
    if (!live) return 0.0f;
    const float wx = float(x - fx), wy = float(y - fy), wz = float(z - fz);
    float b[8];
    if constexpr (kF32) {
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int tx = ix + (n & 1), ty = iy + ((n >> 1) & 1), tz = iz + (n >> 2);
            const bool in = (tx >= 0) & (ty >= 0) & (tz >= 0) & (tx < v.nx) & (ty < v.ny) & (tz < v.nz);
            b[n] = in ? v.f32[uint32_t(tx) + __umul24(uint32_t(v.nx), uint32_t(ty) + __umul24(uint32_t(v.ny),
                                                                                             uint32_t(tz)))]
                      : 0.0f;
        }
    } else {
        const uint32_t c = cx1 + __umul24(uint32_t(v.nx) + 1u, cy1 + __umul24(uint32_t(v.ny) + 1u, cz1));
        const unsigned long long q = v.cells[c];
        const uint32_t lo = uint32_t(q), hi = uint32_t(q >> 32);
        b[0] = float(lo & 0xffu); b[1] = float((lo >> 8) & 0xffu);
        b[2] = float((lo >> 16) & 0xffu); b[3] = float(lo >> 24);
        b[4] = float(hi & 0xffu); b[5] = float((hi >> 8) & 0xffu);
        b[6] = float((hi >> 16) & 0xffu); b[7] = float(hi >> 24);
    }
    const float c00 = fmaf(wx, b[1] - b[0], b[0]);
    const float c10 = fmaf(wx, b[3] - b[2], b[2]);
    const float c01 = fmaf(wx, b[5] - b[4], b[4]);
    const float c11 = fmaf(wx, b[7] - b[6], b[6]);
    const float c0 = fmaf(wy, c10 - c00, c00);
    const float c1 = fmaf(wy, c11 - c01, c01);
    const float r = fmaf(wz, c1 - c0, c0);
    return kF32 ? r : r * (1.0f / 255.0f);
}

// march_tap for the marches over the power-of-two sample grid, at texel
// coordinate + 1 (x1 = x + 1, so floor(x1) is the cell index i + 1 itself).
// Same value as march_tap at x = x1 - 1.  *region receives the radius of
// the empty region around the sample's brick (its distance nibble - 1:
// every brick within that chessboard radius is empty) when the sample's
// cell is inside the volume and its brick is empty, else -1; *brick its
// brick coordinates.
template <bool kF32>
__device__ __forceinline__ float grid_tap(const DensityVolume& v, double x1, double y1, double z1,
                                          OccupancyView occ, int* region, uint3* brick) {
    const double fx = floor(x1), fy = floor(y1), fz = floor(z1);
    const uint32_t cx1 = uint32_t(int(fx)), cy1 = uint32_t(int(fy)), cz1 = uint32_t(int(fz));
    bool live = (cx1 <= uint32_t(v.nx)) & (cy1 <= uint32_t(v.ny)) & (cz1 <= uint32_t(v.nz));
    const uint3 bk = make_uint3(cx1 >> kBrickShift, cy1 >> kBrickShift, cz1 >> kBrickShift);
    int reg = -1;
    if (occ.use) {
        const uint32_t b = bk.x + __umul24(v.bx, bk.y + __umul24(v.by, bk.z));
        const uint32_t d = brick_distance(occ, live ? b : 0u);  // b is in range only for live cells
        reg = (live & (d != 0u)) ? int(d) - 1 : -1;
        live = live & (d == 0u);
    }
    *region = reg;
    *brick = bk;
    if (!live) return 0.0f;
    const float wx = float(x1 - fx), wy = float(y1 - fy), wz = float(z1 - fz);
    float b[8];
    if constexpr (kF32) {
        const int ix = int(cx1) - 1, iy = int(cy1) - 1, iz = int(cz1) - 1;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
            const int tx = ix + (n & 1), ty = iy + ((n >> 1) & 1), tz = iz + (n >> 2);
            const bool in = (tx >= 0) & (ty >= 0) & (tz >= 0) & (tx < v.nx) & (ty < v.ny) & (tz < v.nz);
            b[n] = in ? v.f32[uint32_t(tx) + __umul24(uint32_t(v.nx), uint32_t(ty) + __umul24(uint32_t(v.ny),
                                                                                             uint32_t(tz)))]
                      : 0.0f;
        }
    } else {
        const uint32_t c = cx1 + __umul24(uint32_t(v.nx) + 1u, cy1 + __umul24(uint32_t(v.ny) + 1u, cz1));
        const unsigned long long q = v.cells[c];
        const uint32_t lo = uint32_t(q), hi = uint32_t(q >> 32);
        b[0] = float(lo & 0xffu); b[1] = float((lo >> 8) & 0xffu);
        b[2] = float((lo >> 16) & 0xffu); b[3] = float(lo >> 24);
        b[4] = float(hi & 0xffu); b[5] = float((hi >> 8) & 0xffu);
        b[6] = float((hi >> 16) & 0xffu); b[7] = float(hi >> 24);
    }
    const float c00 = fmaf(wx, b[1] - b[0], b[0]);
    const float c10 = fmaf(wx, b[3] - b[2], b[2]);
    const float c01 = fmaf(wx, b[5] - b[4], b[4]);
    const float c11 = fmaf(wx, b[7] - b[6], b[6]);
    const float c0 = fmaf(wy, c10 - c00, c00);
    const float c1 = fmaf(wy, c11 - c01, c01);
    const float r = fmaf(wz, c1 - c0, c0);
    return kF32 ? r : r * (1.0f / 255.0f);
}

// local_ambient_occlusion with kernel size 2 (8 taps at +-scale/2 texels)
// around texel coordinate s (local_ambient_occlusion.glsl:9-30).
__device__ __forceinline__ double local_ambient_occlusion(const DensityVolume& v, const double s[3], double scale,
                                                          double clamp_v, double exponent) {
    double dens = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        double ox = ((n & 1) ? 0.5 : -0.5) * scale;
        double oy = (((n >> 1) & 1) ? 0.5 : -0.5) * scale;
        double oz = ((n >> 2) ? 0.5 : -0.5) * scale;
        dens += fmin(double(density_at(v, s[0] + ox, s[1] + oy, s[2] + oz)), clamp_v);
    }
    return pow(1.0 - dens / 8.0, exponent);
}

// density_at of the corner-packed cells without branches: a sample outside
// [-1, n) gathers cell 0 and returns 0, so the gathers of a loop of taps
// can all be in flight together.  Same value as density_at inside.
__device__ __forceinline__ float density_at_cells_select(const DensityVolume& v, double x, double y, double z) {
    const bool in = (x >= -1.0) & (y >= -1.0) & (z >= -1.0) & (x < double(v.nx)) & (y < double(v.ny)) &
                    (z < double(v.nz));
    x = in ? x : -1.0;
    y = in ? y : -1.0;
    z = in ? z : -1.0;
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const float wx = float(x - fx), wy = float(y - fy), wz = float(z - fz);
    const int ix = int(fx), iy = int(fy), iz = int(fz);
    const size_t cx = size_t(v.nx) + 1, cy = size_t(v.ny) + 1;
    const unsigned long long c = v.cells[size_t(ix + 1) + cx * (size_t(iy + 1) + cy * size_t(iz + 1))];
    const uint32_t lo = uint32_t(c), hi = uint32_t(c >> 32);
    float b[8];
    b[0] = float(lo & 0xffu); b[1] = float((lo >> 8) & 0xffu);
    b[2] = float((lo >> 16) & 0xffu); b[3] = float(lo >> 24);
    b[4] = float(hi & 0xffu); b[5] = float((hi >> 8) & 0xffu);
    b[6] = float((hi >> 16) & 0xffu); b[7] = float(hi >> 24);
    const float c00 = fmaf(wx, b[1] - b[0], b[0]);
    const float c10 = fmaf(wx, b[3] - b[2], b[2]);
    const float c01 = fmaf(wx, b[5] - b[4], b[4]);
    const float c11 = fmaf(wx, b[7] - b[6], b[6]);
    const float c0 = fmaf(wy, c10 - c00, c00);
    const float c1 = fmaf(wy, c11 - c01, c01);
    const float r = fmaf(wz, c1 - c0, c0);
    return in ? r * (1.0f / 255.0f) : 0.0f;
}

// The strand rasterizer's per-fragment terms with float32 transcendentals
// (one power each; the float64 positions and sums stay).  Their relative
// error, ~1e-7, is far inside the fragment tolerance; the raymarcher keeps
// the float64 forms below.
__device__ __forceinline__ double local_ambient_occlusion_f32pow(const DensityVolume& v, const double s[3],
                                                                 double scale, double clamp_v, double exponent) {
    double dens = 0.0;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        double ox = ((n & 1) ? 0.5 : -0.5) * scale;
        double oy = (((n >> 1) & 1) ? 0.5 : -0.5) * scale;
        double oz = ((n >> 2) ? 0.5 : -0.5) * scale;
        const float d = v.f32 ? density_at(v, s[0] + ox, s[1] + oy, s[2] + oz)
                              : density_at_cells_select(v, s[0] + ox, s[1] + oy, s[2] + oz);
        dens += fmin(double(d), clamp_v);
    }
    return double(powf(float(1.0 - dens / 8.0), float(exponent)));
}

// local_ambient_occlusion_f32pow on the corner-packed cells with the taps'
// per-axis terms formed once: the 8 taps sit at s[c] +- scale/2 on each axis,
// so each axis has two positions, and with them two floors, weights, range
// tests and cell-index terms.  Every tap then adds three index terms.  The
// operations on each position, the trilinear blend and the summation order
// are those of density_at_cells_select in the generic loop, so the value is
// the same bit for bit; a tap outside [-1, n) on any axis gathers cell 0 and
// adds 0, as there.
__device__ __forceinline__ double local_ambient_occlusion_cells_f32pow(const DensityVolume& v, const double s[3],
                                                                      double scale, double clamp_v,
                                                                      double exponent) {
    const uint32_t cx = uint32_t(v.nx) + 1u, cxy = cx * (uint32_t(v.ny) + 1u);
    const int n[3] = {v.nx, v.ny, v.nz};
    const uint32_t stride[3] = {1u, cx, cxy};
    float w[3][2];
    uint32_t term[3][2];
    bool in[3][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const double p = s[c] + (e ? 0.5 : -0.5) * scale;
            in[c][e] = (p >= -1.0) & (p < double(n[c]));
            const double f = floor(p);
            w[c][e] = float(p - f);
            term[c][e] = uint32_t(int(f) + 1) * stride[c];  // used only when in range
        }
    double dens = 0.0;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int i = t & 1, j = (t >> 1) & 1, k = t >> 2;
        const bool inside = in[0][i] & in[1][j] & in[2][k];
        const uint32_t cell = inside ? term[0][i] + term[1][j] + term[2][k] : 0u;
        const unsigned long long q = v.cells[cell];
        const uint32_t lo = uint32_t(q), hi = uint32_t(q >> 32);
        float b[8];
        b[0] = float(lo & 0xffu); b[1] = float((lo >> 8) & 0xffu);
        b[2] = float((lo >> 16) & 0xffu); b[3] = float(lo >> 24);
        b[4] = float(hi & 0xffu); b[5] = float((hi >> 8) & 0xffu);
        b[6] = float((hi >> 16) & 0xffu); b[7] = float(hi >> 24);
        const float wx = w[0][i], wy = w[1][j], wz = w[2][k];
        const float c00 = fmaf(wx, b[1] - b[0], b[0]);
        const float c10 = fmaf(wx, b[3] - b[2], b[2]);
        const float c01 = fmaf(wx, b[5] - b[4], b[4]);
        const float c11 = fmaf(wx, b[7] - b[6], b[6]);
        const float c0 = fmaf(wy, c10 - c00, c00);
        const float c1 = fmaf(wy, c11 - c01, c01);
        const float r = fmaf(wz, c1 - c0, c0);
        dens += fmin(double(inside ? r * (1.0f / 255.0f) : 0.0f), clamp_v);
    }
    return double(powf(float(1.0 - dens / 8.0), float(exponent)));
}

// Kajiya-Kay (kajiya-kay.glsl:7-24); sqrt of a negative rounding residue is 0.
__device__ __forceinline__ void kajiya_kay(const double diffuse[3], const double specular[3], double p,
                                           const double t[3], const double l[3], const double e[3], double out[3],
                                           double* base_out = nullptr) {
    double ctl = t[0] * l[0] + t[1] * l[1] + t[2] * l[2];
    double cte = t[0] * e[0] + t[1] * e[1] + t[2] * e[2];
    double stl = sqrt(fmax(1.0 - ctl * ctl, 0.0)), ste = sqrt(fmax(1.0 - cte * cte, 0.0));
    double base = fmax(ctl * cte + stl * ste, 0.0);
    if (base_out) *base_out = base;
    double spec = pow(base, p);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = diffuse[c] * stl + specular[c] * spec;
}

__device__ __forceinline__ void normalize3(double v[3]) {
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double s = l > 0.0 ? 1.0 / l : 0.0;
    v[0] *= s; v[1] *= s; v[2] *= s;
}

// packUnorm4x8: round(clamp(c, 0, 1) * 255), ties to even.
__device__ __forceinline__ uint32_t pack_unorm4x8(float4 c) {
    auto q = [](float v) -> uint32_t {
        v = fminf(fmaxf(v, 0.0f), 1.0f);
        return uint32_t(rintf(v * 255.0f));
    };
    return q(c.x) | (q(c.y) << 8) | (q(c.z) << 16) | (q(c.w) << 24);
}

struct PpllView {
    uint32_t* heads;
    uint32_t* nodes;   // 3 words per node: colour, depth bits, previous
    uint32_t* counter;
    uint32_t node_count;
};

// Optional per-node capture (tests/debugging).
struct CaptureView {
    uint32_t* segment;
    uint32_t* pixel;
    float4* color;
};

// Inserts one node per lane with `want` set (ppll.glsl:23-48).  Must be
// called by every lane of the wave (it ballots): one counter atomic per
// wave, a lane's node is base + its rank among the wanting lanes.  Returns
// the node index or kPpllNull when the pool is exhausted.
__device__ __forceinline__ uint32_t ppll_insert(const PpllView& p, bool want, uint32_t pixel, float4 color,
                                                float depth) {
    unsigned long long mask = __ballot(want);
    if (mask == 0ull) return kPpllNull;
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(p.counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader, kWave);
    if (!want) return kPpllNull;
    uint32_t node = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    if (node >= p.node_count) return kPpllNull;  // pool exhausted: fragment dropped
    uint32_t* n = p.nodes + 3ull * node;
    n[0] = pack_unorm4x8(color);
    n[1] = __float_as_uint(depth);
    n[2] = atomicExch(&p.heads[pixel], node);
    return node;
}

}  // namespace vkhr
