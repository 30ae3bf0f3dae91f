// The load-time volume build in one call: HairStyle::voxelize_segments
// (src/vkhr/scene_graph/hair_style.cc:296-342) and Volume::normalize
// (:344-357), run back to back as the reference's load path does
// (src/vkhr/rasterizer/hair_style.cc:75-77), with every voxel sum kept in
// LDS instead of device-wide atomics.
//
// Bricks of 8x8x8 voxels.  The index pairs are binned by the bricks their
// samples visit -- a counting sort: count, scan, fill -- then one workgroup
// per brick walks the DDA of every pair in its bin and adds the samples
// that fall inside its brick into an LDS accumulator (count and fixed-point
// tangent sums: integer adds, so the order of the adds does not matter).
// The workgroup finalizes its brick straight from LDS: the saturating uint8
// view, the quantized tangent (the sums and the expression of
// vkhr_volume_finalize, so both give the same bits), the counts when asked
// for, and the min/max of the uint8 view.  The corner-packed cells pass
// then applies the normalize as it loads the uint8 view (writing the
// normalized volume and the float view as it goes) and sets the brick
// occupancy; the brick distance field follows as in vkhr_volume_finalize.
//
// The count and fill passes give each workgroup a contiguous range of
// pairs and count the runs of its range in an LDS hash table first, so a
// brick takes one device-wide atomic per workgroup that reaches it rather
// than one per run (keys that find no free slot fall back to a device-wide
// atomic per run).
//
// A pair's voxel coordinates are monotone in its sample index (the root is
// advanced by a fixed direction and the clamps are monotone), so the bricks
// it visits form runs that never repeat: a pair lands in a bin at most once.
//
// The bin entries live in the per-stream scratch, 4 per pair plus 64 Ki.
// If the pairs visit more bricks than that (long segments), a flag set on
// the device switches the build to the workspace path: the fill pass zeroes
// the caller's workspace instead, the brick pass voxelizes into it with the
// packed device-wide atomics of vkhr_voxelize_segments, and one more kernel
// (which returns at once while the flag is clear) finalizes from it, so the
// host never waits for the count.
//
// Compiled with -ffp-contract=off: the DDA and the quantization must round
// as the reference's float arithmetic does.
#include "volume.hpp"
#include "voxel.hpp"

#include <algorithm>

namespace vkhr {
namespace {

constexpr uint32_t kBuildBlock = 256;
constexpr uint32_t kNone = 0xffffffffu;

// Bricks of the voxel grid (8 voxels a side, the last ones partial).
struct BrickGrid {
    uint32_t bx, by, bz;
};

__host__ __device__ __forceinline__ BrickGrid brick_grid(uint32_t nx, uint32_t ny, uint32_t nz) {
    return {(nx + 7u) >> 3, (ny + 7u) >> 3, (nz + 7u) >> 3};
}

__device__ __forceinline__ uint32_t brick_index(const BrickGrid& b, uint3 v) {
    return (v.x >> 3) + b.bx * ((v.y >> 3) + b.by * (v.z >> 3));
}

// Where the build keeps its state in the per-stream scratch (32-bit words).
struct BuildLayout {
    size_t bins;        // bin sizes, one per brick
    size_t cursor;      // bin starts, then (after the fill) bin ends
    size_t status;      // [0] total entries, [1] overflow flag
    size_t minmax;      // the min/max slots (kMinMaxBytes)
    size_t raw;         // the uint8 view before the normalize (with cells), n voxels
    size_t entries;     // bin entries: the pairs' vertex indices (uint2)
    size_t words;
};

__host__ __forceinline__ BuildLayout build_layout(size_t n_bricks, size_t n_voxels, bool raw, size_t capacity) {
    auto up = [](size_t w) { return (w + 63) & ~size_t(63); };  // 256-byte aligned sections
    BuildLayout l;
    l.bins = 0;
    l.cursor = up(n_bricks);
    l.status = l.cursor + up(n_bricks);
    l.minmax = l.status + 64;
    l.raw = l.minmax + up(kMinMaxBytes / 4);
    l.entries = l.raw + (raw ? up((n_voxels + 3) / 4) : 0);
    l.words = l.entries + up(2 * capacity);
    return l;
}

struct BuildArgs {
    VoxelGrid g;
    BrickGrid b;
    const float* xyz;
    const uint32_t* indices;
    const float* tangents;          // per vertex (null: no tangent volume)
    uint32_t n_vertices, n_pairs;
    uint32_t n_bricks;
    uint32_t pairs_per_block;       // the contiguous range of a bin workgroup
    uint32_t* bins;
    uint32_t* cursor;
    uint32_t* status;
    uint32_t* minmax;
    uint2* entries;
    uint32_t capacity;
    uint32_t* occupancy;            // zeroed by the init pass (nullable)
    size_t occupancy_words;
    uint32_t* counts;               // outputs (counts and tangents nullable)
    uint8_t* density;               // the uint8 view the brick pass writes (raw, or the output)
    int8_t* tangent_i8x4;
    long long* acc;                 // workspace of the overflow path
};

// Zeroes the bin sizes, the status and the occupancy, and empties the
// min/max slots.
__global__ void __launch_bounds__(256) build_init_kernel(BuildArgs a) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t first = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t i = first; i < a.n_bricks; i += stride) a.bins[i] = 0u;
    for (size_t i = first; i < a.occupancy_words; i += stride) a.occupancy[i] = 0u;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 2) a.status[threadIdx.x] = 0u;
        if (threadIdx.x < kMinMaxSlots) {
            a.minmax[threadIdx.x * kMinMaxStride] = 255u;
            a.minmax[threadIdx.x * kMinMaxStride + 1] = 0u;
        }
    }
}

// --- the bin passes

// An LDS hash table of the bricks one workgroup's pairs reach: open
// addressing over kTableSlots slots, kTableProbes probes.  A key that finds
// no free slot among its probes never will (slots are only ever filled),
// so every lookup of it fails alike and its runs take the device-wide path.
constexpr uint32_t kTableBits = 10, kTableSlots = 1u << kTableBits, kTableProbes = 16;

struct BinTable {
    uint32_t key[kTableSlots];
    uint32_t count[kTableSlots];
    uint32_t base[kTableSlots];
};

__device__ __forceinline__ uint32_t table_home(uint32_t k) { return (k * 2654435761u) >> (32 - kTableBits); }

__device__ __forceinline__ uint32_t table_insert(BinTable& t, uint32_t k) {
    const uint32_t h = table_home(k);
    for (uint32_t p = 0; p < kTableProbes; ++p) {
        const uint32_t slot = (h + p) & (kTableSlots - 1u);
        const uint32_t prev = atomicCAS(&t.key[slot], kNone, k);
        if (prev == kNone || prev == k) return slot;
    }
    return kNone;
}

__device__ __forceinline__ uint32_t table_find(const BinTable& t, uint32_t k) {
    const uint32_t h = table_home(k);
    for (uint32_t p = 0; p < kTableProbes; ++p) {
        const uint32_t slot = (h + p) & (kTableSlots - 1u);
        const uint32_t key = t.key[slot];
        if (key == k) return slot;
        if (key == kNone) return kNone;
    }
    return kNone;
}

__device__ __forceinline__ void table_clear(BinTable& t) {
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x) {
        t.key[i] = kNone;
        t.count[i] = 0u;
    }
}

// Calls run(pair's vertex indices, brick) for each brick run of each pair
// of this workgroup's range (at most kPairsPerThread per thread: the index
// loads of all of a thread's pairs are issued before the first walk).
constexpr uint32_t kPairsPerThread = 4;

template <class Run>
__device__ __forceinline__ void walk_runs(const BuildArgs& a, Run run) {
    const uint32_t r0 = blockIdx.x * a.pairs_per_block;
    const uint32_t r1 = min(a.n_pairs, r0 + a.pairs_per_block);
    uint2 pr[kPairsPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        const uint32_t s = r0 + threadIdx.x + j * blockDim.x;
        pr[j] = s < r1 ? make_uint2(a.indices[2u * s], a.indices[2u * s + 1u]) : make_uint2(kNone, kNone);
    }
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        SegmentDda w = segment_dda(a.g, a.xyz, a.n_vertices, pr[j].x, pr[j].y);  // kNone: no samples
        uint32_t last = kNone;
        while (w.steps > 0.0f) {
            w.steps -= 1.0f;
            const uint32_t k = brick_index(a.b, voxel_coords(a.g, w.rx, w.ry, w.rz));
            if (k != last) run(pr[j], k);
            last = k;
            w.advance();
        }
    }
}

// The bin sizes: runs counted per brick in the table, then one add per
// table entry.
__global__ void __launch_bounds__(kBuildBlock) bin_count_kernel(BuildArgs a) {
    __shared__ BinTable t;
    table_clear(t);
    __syncthreads();
    walk_runs(a, [&](uint2, uint32_t k) {
        const uint32_t slot = table_insert(t, k);
        if (slot != kNone)
            atomicAdd(&t.count[slot], 1u);
        else
            atomicAdd(&a.bins[k], 1u);
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x)
        if (t.key[i] != kNone) atomicAdd(&a.bins[t.key[i]], t.count[i]);
}

// The bin entries: the same counts in the table, one range reserved per
// table entry in its brick's bin, then a second walk that places each run
// in its range (table misses take a slot of their own).  With the overflow
// flag set the workgroups zero the workspace instead.
__global__ void __launch_bounds__(kBuildBlock) bin_fill_kernel(BuildArgs a, size_t workspace_words) {
    __shared__ BinTable t;
    if (a.status[1]) {
        uint32_t* acc = reinterpret_cast<uint32_t*>(a.acc);
        const size_t stride = size_t(gridDim.x) * blockDim.x;
        for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < workspace_words; i += stride) acc[i] = 0u;
        return;
    }
    table_clear(t);
    __syncthreads();
    walk_runs(a, [&](uint2, uint32_t k) {
        const uint32_t slot = table_insert(t, k);
        if (slot != kNone) atomicAdd(&t.count[slot], 1u);
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x) {
        if (t.key[i] != kNone) t.base[i] = atomicAdd(&a.cursor[t.key[i]], t.count[i]);
        t.count[i] = 0u;
    }
    __syncthreads();
    walk_runs(a, [&](uint2 pair, uint32_t k) {
        const uint32_t slot = table_find(t, k);
        const uint32_t pos = slot != kNone ? t.base[slot] + atomicAdd(&t.count[slot], 1u) : atomicAdd(&a.cursor[k], 1u);
        if (pos < a.capacity) a.entries[pos] = pair;
    });
}

// Exclusive scan of the bin sizes into the bin starts (one block, 16 sizes
// per thread per round), the total, and the overflow flag.
__global__ void __launch_bounds__(1024) scan_kernel(BuildArgs a) {
    constexpr uint32_t kPer = 16, kThreads = 1024;
    __shared__ unsigned long long wave_sums[kThreads / kWave];
    __shared__ unsigned long long carry_lds;  // 64-bit: the total of absurd inputs may pass 2^32
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t n = a.n_bricks;
    if (threadIdx.x == 0) carry_lds = 0;
    __syncthreads();
    for (uint32_t chunk = 0; chunk < n; chunk += kPer * kThreads) {
        const uint32_t i0 = chunk + kPer * threadIdx.x;
        uint32_t v[kPer];
        if (i0 + kPer <= n) {  // the sections are 256-byte aligned and i0 a multiple of 16
#pragma unroll
            for (uint32_t q = 0; q < kPer / 4; ++q) {
                const uint4 u = reinterpret_cast<const uint4*>(a.bins + i0)[q];
                v[4 * q] = u.x; v[4 * q + 1] = u.y; v[4 * q + 2] = u.z; v[4 * q + 3] = u.w;
            }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j) v[j] = i0 + j < n ? a.bins[i0 + j] : 0u;
        }
        unsigned long long sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) sum += v[j];
        // inclusive wave scan of the per-thread sums
        unsigned long long inc = sum;
        for (int o = 1; o < kWave; o <<= 1) {
            const unsigned long long t = __shfl_up(inc, o, kWave);
            if (int(lane) >= o) inc += t;
        }
        if (lane == kWave - 1) wave_sums[wave] = inc;
        __syncthreads();
        unsigned long long before = carry_lds;
        for (uint32_t w = 0; w < wave; ++w) before += wave_sums[w];
        before += inc - sum;
        // a wrapped start only happens with the overflow flag set, and then
        // no bin is read
        uint32_t out[kPer];
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            out[j] = uint32_t(before);
            before += v[j];
        }
        if (i0 + kPer <= n) {
#pragma unroll
            for (uint32_t q = 0; q < kPer / 4; ++q)
                reinterpret_cast<uint4*>(a.cursor + i0)[q] =
                    make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kPer; ++j)
                if (i0 + j < n) a.cursor[i0 + j] = out[j];
        }
        __syncthreads();
        if (threadIdx.x == kThreads - 1) carry_lds = before;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.status[0] = uint32_t(carry_lds < 0xffffffffull ? carry_lds : 0xffffffffull);
        a.status[1] = carry_lds > a.capacity ? 1u : 0u;
    }
}

// Brick of a workgroup: groups of 8 consecutive bricks along x go to the
// same XCD (workgroups are dealt to the 8 XCDs in turn), so the rows of a
// group's output are written through one L2.  Bricks past the grid
// (groups padded to 8 along x) return kNone.
__device__ __forceinline__ uint32_t brick_of_block(const BrickGrid& b, uint32_t block) {
    const uint32_t xcd = block & 7u, j = block >> 3;
    const uint32_t group = (j >> 3) * 8u + xcd, within = j & 7u;
    const uint32_t gx = (b.bx + 7u) >> 3;  // groups along x
    const uint32_t row = group / gx, x = (group % gx) * 8u + within;
    if (x >= b.bx || row >= b.by * b.bz) return kNone;
    return x + b.bx * row;
}

// Workgroups of brick_kernel: the groups rounded up to a multiple of 8, so
// that brick_of_block reaches every group.
__host__ __forceinline__ uint32_t brick_blocks(const BrickGrid& b) {
    const uint32_t groups = ((b.bx + 7u) >> 3) * b.by * b.bz;
    return ((groups + 7u) & ~7u) * 8u;
}

// One brick per workgroup: the samples of its bin's pairs that fall inside
// it, accumulated in LDS, then the brick's outputs.  The tangent sums are
// 32-bit: at 2^16 per unit they hold up to 32767 contributions exactly,
// and a voxel of more gets tangent 0 (as from the packed accumulator), so
// the quantized tangents are those of vkhr_volume_finalize.  With the
// overflow flag set the workgroups voxelize every pair into the workspace
// instead.
constexpr uint32_t kBrickBlock = 128;

__global__ void __launch_bounds__(kBrickBlock) brick_kernel(BuildArgs a) {
    constexpr uint32_t kVox = 512;
    __shared__ uint32_t count[kVox];
    __shared__ int sum[3][kVox];
    __shared__ uint32_t red[2][kBrickBlock / kWave];
    if (a.status[1]) {
        // grid-stride over wave-aligned ranges (voxelize_segment is wave-cooperative)
        const float* tangents = a.tangent_i8x4 ? a.tangents : a.xyz;  // sums unused without a tangent volume
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < a.n_pairs;
             base += stride) {
            const uint32_t s = base + (threadIdx.x & uint32_t(kWave - 1));
            const bool has = s < a.n_pairs;
            voxelize_segment(a.g, a.xyz, a.n_vertices, has, has ? a.indices[2u * s] : 0u,
                             has ? a.indices[2u * s + 1u] : 0u, tangents, nullptr, a.acc);
        }
        return;
    }
    const uint32_t brick = brick_of_block(a.b, blockIdx.x);
    if (brick == kNone) return;
    for (uint32_t l = threadIdx.x; l < kVox; l += kBrickBlock) {
        count[l] = 0u;
        sum[0][l] = sum[1][l] = sum[2][l] = 0;
    }
    const uint32_t bxo = brick % a.b.bx, byo = (brick / a.b.bx) % a.b.by, bzo = brick / (a.b.bx * a.b.by);
    __syncthreads();
    const uint32_t end = a.cursor[brick], n = a.bins[brick];
    const bool with_tangents = a.tangent_i8x4 != nullptr;
    for (uint32_t e = end - n + threadIdx.x; e < end; e += kBrickBlock) {
        const uint2 pair = a.entries[e];
        SegmentDda w = segment_dda(a.g, a.xyz, a.n_vertices, pair.x, pair.y);
        int tx = 0, ty = 0, tz = 0;
        if (with_tangents && w.steps > 0.0f) {
            const float3 t = load_xyz(a.tangents, pair.x);
            tx = int(tangent_fixed(t.x));
            ty = int(tangent_fixed(t.y));
            tz = int(tangent_fixed(t.z));
        }
        bool entered = false;
        while (w.steps > 0.0f) {
            w.steps -= 1.0f;
            const uint3 v = voxel_coords(a.g, w.rx, w.ry, w.rz);
            const bool inside = (v.x >> 3) == bxo && (v.y >> 3) == byo && (v.z >> 3) == bzo;
            if (inside) {
                const uint32_t l = (v.x & 7u) + 8u * ((v.y & 7u) + 8u * (v.z & 7u));
                atomicAdd(&count[l], 1u);
                if (with_tangents) {  // wraps only past 32767 contributions, where the tangent is 0
                    atomicAdd(reinterpret_cast<unsigned*>(&sum[0][l]), unsigned(tx));
                    atomicAdd(reinterpret_cast<unsigned*>(&sum[1][l]), unsigned(ty));
                    atomicAdd(reinterpret_cast<unsigned*>(&sum[2][l]), unsigned(tz));
                }
            } else if (entered) {
                break;  // the brick run is over: the pair never comes back
            }
            entered = entered || inside;
            w.advance();
        }
    }
    __syncthreads();
    uint32_t lo = 255u, hi = 0u;
    for (uint32_t l = threadIdx.x; l < kVox; l += kBrickBlock) {
        const uint32_t x = bxo * 8u + (l & 7u), y = byo * 8u + ((l >> 3) & 7u), z = bzo * 8u + (l >> 6);
        if (x >= a.g.nx || y >= a.g.ny || z >= a.g.nz) continue;
        const size_t v = size_t(x) + size_t(a.g.nx) * (size_t(y) + size_t(a.g.ny) * z);
        const uint32_t c = count[l];
        const uint32_t d = c < 255u ? c : 255u;
        a.density[v] = uint8_t(d);
        if (a.counts) a.counts[v] = c;
        if (with_tangents)
            reinterpret_cast<char4*>(a.tangent_i8x4)[v] = quantize_tangent_sums(sum[0][l], sum[1][l], sum[2][l], c);
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t p = __shfl_xor(lo, o, kWave), q = __shfl_xor(hi, o, kWave);
        lo = p < lo ? p : lo;
        hi = q > hi ? q : hi;
    }
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane == 0) {
        red[0][wave] = lo;
        red[1][wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < kBrickBlock / kWave; ++w) {
            lo = red[0][w] < lo ? red[0][w] : lo;
            hi = red[1][w] > hi ? red[1][w] : hi;
        }
        if (lo <= hi)  // the brick holds a voxel of the volume
            fold_minmax(a.minmax + (blockIdx.x % kMinMaxSlots) * kMinMaxStride, lo, hi);
    }
}

// The finalize of the workspace path (returns at once while the flag is
// clear): the outputs of the brick pass from the packed accumulator.
__global__ void __launch_bounds__(256) overflow_finalize_kernel(BuildArgs a, size_t n) {
    if (!a.status[1]) return;
    uint32_t lo = 255u, hi = 0u;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t first = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t i = first; i < n; i += stride) {
        const long long w0 = a.acc[2 * i], w1 = a.acc[2 * i + 1];
        const uint32_t c = uint32_t((unsigned long long)(w1) & 0xffffffffull);
        const uint32_t d = c < 255u ? c : 255u;
        a.density[i] = uint8_t(d);
        if (a.counts) a.counts[i] = c;
        if (a.tangent_i8x4) reinterpret_cast<char4*>(a.tangent_i8x4)[i] = quantize_tangent_words(w0, w1, c);
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t p = __shfl_xor(lo, o, kWave), q = __shfl_xor(hi, o, kWave);
        lo = p < lo ? p : lo;
        hi = q > hi ? q : hi;
    }
    if (threadIdx.x % kWave == 0 && lo <= hi)
        fold_minmax(a.minmax + (blockIdx.x % kMinMaxSlots) * kMinMaxStride, lo, hi);
}

// The corner-packed cells from the uint8 view with the normalize applied as
// the texels load (the map from the min/max slots), writing the normalized
// volume and its float view on the way (cells_block<true>).
__global__ void __launch_bounds__(256)
mapped_cells_kernel(const uint8_t* __restrict__ raw, uint32_t nx, uint32_t ny, uint32_t nz,
                    const uint32_t* __restrict__ minmax, int normalize, unsigned long long* __restrict__ cells,
                    uint32_t* __restrict__ occupancy, uint8_t* __restrict__ out_u8, float* __restrict__ out_f32) {
    __shared__ uint8_t lut[256];
    const DensityMap m = density_map_slots(minmax, normalize != 0);
    // the map of every uint8 value; the identity (no normalize, or a volume
    // spanning 0..255) needs none
    bool identity = true;
    for (uint32_t v = threadIdx.x; v < 256u; v += blockDim.x) {
        const uint32_t d = texel_density(v, m);
        lut[v] = uint8_t(d);
        identity = identity && d == v;
    }
    identity = __syncthreads_and(identity);
    const uint32_t n_blocks = cell_blocks(nx, ny, nz);
    for (uint32_t b = blockIdx.x; b < n_blocks; b += gridDim.x) {
        if (identity)
            cells_block<1>(raw, nx, ny, nz, b, cells, occupancy, nullptr, out_u8, out_f32);
        else
            cells_block<2>(raw, nx, ny, nz, b, cells, occupancy, lut, out_u8, out_f32);
    }
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_volume_build_segments(const float* xyz, uint32_t n_vertices, const uint32_t* indices,
                                          uint32_t n_indices, const float* tangents, const vkhr_aabb* bounds,
                                          uint32_t nx, uint32_t ny, uint32_t nz, int normalize, uint32_t* counts,
                                          uint8_t* density_u8, int8_t* tangent_i8x4, float* density_f32,
                                          uint64_t* cells, uint32_t* occupancy, int64_t* workspace,
                                          vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VoxelGrid g;
    if ((rc = make_grid(bounds, nx, ny, nz, &g))) return rc;
    if ((rc = check_volume(nx, ny, nz, cells != nullptr, "volume build"))) return rc;
    const uint32_t n_pairs = n_indices / 2;  // the reference loops while i < size-1
    VKHR_REQUIRE(n_pairs == 0 || (xyz && indices && n_vertices), VKHR_ERR_INVALID, "xyz/indices is null");
    VKHR_REQUIRE(density_u8 && workspace, VKHR_ERR_INVALID, "density_u8 and workspace are required");
    VKHR_REQUIRE(!tangent_i8x4 || tangents, VKHR_ERR_INVALID, "tangent_i8x4 needs tangents");
    VKHR_REQUIRE(!occupancy || cells, VKHR_ERR_INVALID, "occupancy needs cells");
    VKHR_REQUIRE(reinterpret_cast<uintptr_t>(density_u8) % 4 == 0 && reinterpret_cast<uintptr_t>(density_f32) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(workspace) % 16 == 0,
                 VKHR_ERR_INVALID, "density_u8 must be 4-byte and density_f32, workspace 16-byte aligned");
    hipStream_t s = as_stream(stream);
    const BrickGrid bg = brick_grid(nx, ny, nz);
    const size_t n_bricks = size_t(bg.bx) * bg.by * bg.bz;  // <= 2^24 (make_grid bounds the voxels)
    const size_t n = size_t(nx) * ny * nz;
    const size_t capacity = std::min<size_t>(4 * size_t(n_pairs) + 65536, size_t(1) << 28);
    // with cells the brick pass writes the uint8 view into the scratch and
    // the cells pass writes the normalized volume
    const bool mapped = cells != nullptr;
    const BuildLayout lay = build_layout(n_bricks, n, mapped, capacity);
    void* scr = nullptr;
    if ((rc = stream_scratch(lay.words * 4, s, &scr))) return rc;
    uint32_t* w = static_cast<uint32_t*>(scr);
    BuildArgs a;
    a.g = g;
    a.b = bg;
    a.xyz = xyz;
    a.indices = indices;
    a.tangents = tangents;
    a.n_vertices = n_vertices;
    a.n_pairs = n_pairs;
    a.n_bricks = uint32_t(n_bricks);
    // ranges of 256 .. 1024 pairs (kPairsPerThread per thread), about 2048 workgroups when that fits
    a.pairs_per_block = std::min(kPairsPerThread * kBuildBlock, std::max(kBuildBlock, (n_pairs + 2047) / 2048));
    const uint32_t bin_blocks = std::max(1u, (n_pairs + a.pairs_per_block - 1) / a.pairs_per_block);
    a.bins = w + lay.bins;
    a.cursor = w + lay.cursor;
    a.status = w + lay.status;
    a.minmax = w + lay.minmax;
    a.entries = reinterpret_cast<uint2*>(w + lay.entries);
    a.capacity = uint32_t(capacity);
    a.occupancy = occupancy;
    a.occupancy_words = occupancy ? occupancy_words(nx, ny, nz) : 0;
    a.counts = counts;
    a.density = mapped ? reinterpret_cast<uint8_t*>(w + lay.raw) : density_u8;
    a.tangent_i8x4 = tangent_i8x4;
    a.acc = reinterpret_cast<long long*>(workspace);

    build_init_kernel<<<grid_for(std::max(n_bricks, a.occupancy_words), 256, 1024), 256, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        bin_count_kernel<<<bin_blocks, kBuildBlock, 0, s>>>(a);
        VKHR_CHECK_LAUNCH();
    }
    scan_kernel<<<1, 1024, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        bin_fill_kernel<<<bin_blocks, kBuildBlock, 0, s>>>(a, 4 * n);
        VKHR_CHECK_LAUNCH();
    }
    brick_kernel<<<brick_blocks(bg), kBrickBlock, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    // the workspace path's finalize: every block returns at once unless the bins overflowed
    overflow_finalize_kernel<<<1024, 256, 0, s>>>(a, n);
    VKHR_CHECK_LAUNCH();
    if (!mapped)
        return finish_volume(density_u8, nx, ny, nz, normalize, density_f32, nullptr, nullptr,
                             normalize ? a.minmax : nullptr, s);
    mapped_cells_kernel<<<grid_for(cell_blocks(nx, ny, nz), 1, 1u << 20), 256, 0, s>>>(
        a.density, nx, ny, nz, a.minmax, normalize, reinterpret_cast<unsigned long long*>(cells), occupancy,
        density_u8, density_f32);
    VKHR_CHECK_LAUNCH();
    return occupancy ? launch_distance_field(occupancy, nx, ny, nz, s) : VKHR_OK;
}
