// Device side of the LDS-binned volume build (HairStyle::voxelize_segments,
// src/vkhr/scene_graph/hair_style.cc:296-342, then Volume::normalize,
// :344-357), shared by the one-call build of one style and the batched
// build of many.
//
// Bricks of 8x8x8 voxels.  The index pairs are binned by the bricks their
// samples visit -- a counting sort: count, scan, fill -- then one
// workgroup per brick walks the DDA of every pair in its
// bin (from the run's first sample) and adds the samples
// that fall inside its brick into an LDS accumulator (count and fixed-point
// tangent sums: integer adds, so the order of the adds does not matter).
// The workgroup finalizes its brick straight from LDS: the saturating uint8
// view, the quantized tangent (the sums and the expression of
// vkhr_volume_finalize, so both give the same bits), the counts when asked
// for, and the min/max of the uint8 view.  The corner-packed cells pass
// then applies the normalize as it loads the uint8 view (writing the
// normalized volume and the float view as it goes) and sets the brick
// occupancy.
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
// If the pairs visit more bricks than the entry capacity (long segments), a
// flag set on the device by the scan switches the build to the workspace
// path: the fill pass zeroes the workspace instead, the brick pass
// voxelizes into it with the device-wide atomics of vkhr_voxelize_segments,
// and one more pass (which returns at once while the flag is clear)
// finalizes from it, so the host never waits for the count.
//
// Every body takes its workgroup's index within the style's own launch
// (block) and that launch's size (blocks), so one style's passes run as
// plain kernels and many styles' as one launch over the concatenated
// blocks.  Compiled with -ffp-contract=off: the DDA and the quantization
// must round as the reference's float arithmetic does.
#pragma once

#include "volume.hpp"
#include "voxel.hpp"

namespace vkhr {

constexpr uint32_t kBuildBlock = 256;
constexpr uint32_t kBrickBlock = 128;
constexpr uint32_t kBuildNone = 0xffffffffu;

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

// A bin entry: one pair's run through one brick -- the DDA at the run's
// first sample (position and step; the walk on from there makes the same
// float adds as the walk from the pair's root) and a 64-bit word holding
// the run's sample limit (bits 0-9: the samples left, at most kMaxRun) and
// the fixed-point tangent of the pair's root (three 18-bit fields, each
// tangent_fixed + 65536), so the brick pass gathers nothing.  Runs longer
// than kMaxRun samples (a clamped walk along the grid's edge can stay in
// one brick) are split into entries of kMaxRun, by the count and the fill
// pass alike.
struct BinEntry {
    float rx, ry, rz, dx;
    float dy, dz;
    uint32_t word_lo, word_hi;
};
static_assert(sizeof(BinEntry) == 32, "two 16-byte words per entry");
constexpr uint32_t kMaxRun = 1023;

__device__ __forceinline__ unsigned long long pack_tangent(float3 t) {
    const unsigned long long x = (unsigned long long)(tangent_fixed(t.x) + 65536);
    const unsigned long long y = (unsigned long long)(tangent_fixed(t.y) + 65536);
    const unsigned long long z = (unsigned long long)(tangent_fixed(t.z) + 65536);
    return (x << 10) | (y << 28) | (z << 46);
}

// Where the build keeps its state (32-bit words from one base).
struct BuildLayout {
    size_t bins;        // bin sizes, one per brick
    size_t cursor;      // bin starts, then (after the fill) bin ends
    size_t status;      // [0] total entries, [1] overflow flag
    size_t minmax;      // the min/max slots (kMinMaxBytes)
    size_t brick_mm;    // per brick: the min/max of the uint8 view of its two waves' voxels (4 bytes)
    size_t raw;         // the uint8 view before the normalize (with cells), n voxels
    size_t entries;     // bin entries (BinEntry)
    size_t words;
};

__host__ __forceinline__ BuildLayout build_layout(size_t n_bricks, size_t n_voxels, bool raw, size_t capacity) {
    auto up = [](size_t w) { return (w + 63) & ~size_t(63); };  // 256-byte aligned sections
    BuildLayout l;
    l.bins = 0;
    l.cursor = up(n_bricks);
    l.status = l.cursor + up(n_bricks);
    l.minmax = l.status + 64;
    l.brick_mm = l.minmax + up(kMinMaxBytes / 4);
    l.raw = l.brick_mm + up(n_bricks);
    l.entries = l.raw + (raw ? up((n_voxels + 3) / 4) : 0);
    l.words = l.entries + up(sizeof(BinEntry) / 4 * capacity);
    return l;
}

// Bin entries per style: 4 per pair plus 64 Ki, at most 2^26 (2 GiB).
__host__ __forceinline__ size_t build_capacity(uint32_t n_pairs) {
    const size_t c = 4 * size_t(n_pairs) + 65536;
    return c < (size_t(1) << 26) ? c : (size_t(1) << 26);
}

// Pairs per bin workgroup: ranges of 256 .. 1024 (kPairsPerThread per
// thread), about 2048 workgroups when that fits.
__host__ __forceinline__ uint32_t build_pairs_per_block(uint32_t n_pairs) {
    const uint32_t want = (n_pairs + 2047) / 2048;
    return want < kBuildBlock ? kBuildBlock : (want > 4 * kBuildBlock ? 4 * kBuildBlock : want);
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
    uint32_t* brick_mm;             // bytes lo0, hi0, lo1, hi1 per brick (lo > hi: no voxel)
    BinEntry* entries;
    uint32_t capacity;
    uint32_t* occupancy;            // zeroed by the init pass (nullable)
    size_t occupancy_words;
    uint32_t* counts;               // outputs (counts and tangents nullable)
    uint8_t* density;               // the uint8 view the brick pass writes (raw, or the output)
    int8_t* tangent_i8x4;
    long long* acc;                 // workspace of the overflow path: count + tangent sums, 16 B per voxel,
    uint32_t* acc_counts;           // or (acc null, no tangent volume) the counts alone
};

// The words the overflow path zeroes and accumulates into.
__host__ __device__ __forceinline__ size_t overflow_words(const BuildArgs& a, size_t n_voxels) {
    return a.acc ? 4 * n_voxels : n_voxels;
}

// Zeroes the bin sizes, the status and the occupancy, and empties the
// min/max slots.
__device__ __forceinline__ void build_init_body(const BuildArgs& a, uint32_t block, uint32_t blocks) {
    const size_t stride = size_t(blocks) * blockDim.x;
    const size_t first = size_t(block) * blockDim.x + threadIdx.x;
    for (size_t i = first; i < a.n_bricks; i += stride) a.bins[i] = 0u;
    for (size_t i = first; i < a.occupancy_words; i += stride) a.occupancy[i] = 0u;
    if (block == 0) {
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

// An add to a table counter at workgroup scope (an LDS atomic): written
// apart from the device-wide adds of the table misses, which the compiler
// would otherwise merge with it into one flat atomic on a selected address.
__device__ __forceinline__ uint32_t table_add(uint32_t* counter, uint32_t v) {
    return __hip_atomic_fetch_add(counter, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t table_insert(BinTable& t, uint32_t k) {
    const uint32_t h = table_home(k);
    for (uint32_t p = 0; p < kTableProbes; ++p) {
        const uint32_t slot = (h + p) & (kTableSlots - 1u);
        const uint32_t prev = atomicCAS(&t.key[slot], kBuildNone, k);
        if (prev == kBuildNone || prev == k) return slot;
    }
    return kBuildNone;
}

__device__ __forceinline__ uint32_t table_find(const BinTable& t, uint32_t k) {
    const uint32_t h = table_home(k);
    for (uint32_t p = 0; p < kTableProbes; ++p) {
        const uint32_t slot = (h + p) & (kTableSlots - 1u);
        const uint32_t key = t.key[slot];
        if (key == k) return slot;
        if (key == kBuildNone) return kBuildNone;
    }
    return kBuildNone;
}

__device__ __forceinline__ void table_clear(BinTable& t) {
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x) {
        t.key[i] = kBuildNone;
        t.count[i] = 0u;
    }
}

// The DDAs of one thread's pairs of workgroup block's range (at most
// kPairsPerThread; the index and vertex loads of all of them are issued
// before the first walk), kept so that a pass can walk them twice.
constexpr uint32_t kPairsPerThread = 4;

struct PairWalks {
    SegmentDda w[kPairsPerThread];
    unsigned long long tangent[kPairsPerThread];  // pack_tangent of the pair's root (0 without tangents)
};

__device__ __forceinline__ void load_walks(const BuildArgs& a, uint32_t block, PairWalks& p) {
    const uint32_t r0 = block * a.pairs_per_block;
    const uint32_t r1 = min(a.n_pairs, r0 + a.pairs_per_block);
    uint2 pr[kPairsPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        const uint32_t s = r0 + threadIdx.x + j * blockDim.x;
        pr[j] = s < r1 ? make_uint2(a.indices[2u * s], a.indices[2u * s + 1u]) : make_uint2(kBuildNone, kBuildNone);
    }
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        p.w[j] = segment_dda(a.g, a.xyz, a.n_vertices, pr[j].x, pr[j].y);  // kBuildNone: no samples
        p.tangent[j] = a.tangent_i8x4 && p.w[j].steps > 0.0f ? pack_tangent(load_vertex(a.tangents, pr[j].x)) : 0ull;
    }
}

// Calls run(j, DDA at the sample, steps before the sample, brick) at the
// first sample of each brick run (split every kMaxRun samples) of each of
// the thread's pairs.
template <class Run>
__device__ __forceinline__ void walk_runs(const BuildArgs& a, const PairWalks& p, Run run) {
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        SegmentDda w = p.w[j];
        uint32_t last = kBuildNone, len = 0;
        while (w.steps > 0.0f) {
            const float before = w.steps;
            w.steps -= 1.0f;
            const uint32_t k = brick_index(a.b, voxel_coords(a.g, w.rx, w.ry, w.rz));
            if (k != last || len == kMaxRun) {
                run(j, w, before, k);
                len = 0;
            }
            ++len;
            last = k;
            w.advance();
        }
    }
}

// Exclusive scan of the bin sizes into the bin starts (one workgroup of at
// most kScanThreads, 16 sizes per thread per round), the total, and the
// overflow flag.
constexpr uint32_t kScanThreads = 1024, kMaxScanWaves = kScanThreads / kWave;

struct ScanLds {
    unsigned long long wave_sums[kMaxScanWaves];
    unsigned long long carry;  // 64-bit: the total of absurd inputs may pass 2^32
};

__device__ __forceinline__ void scan_body(const BuildArgs& a, ScanLds& l) {
    constexpr uint32_t kPer = 16;
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const uint32_t n = a.n_bricks;
    if (threadIdx.x == 0) l.carry = 0;
    __syncthreads();
    for (uint32_t chunk = 0; chunk < n; chunk += kPer * blockDim.x) {
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
        if (lane == kWave - 1) l.wave_sums[wave] = inc;
        __syncthreads();
        unsigned long long before = l.carry;
        for (uint32_t w = 0; w < wave; ++w) before += l.wave_sums[w];
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
        if (threadIdx.x == blockDim.x - 1) l.carry = before;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.status[0] = uint32_t(l.carry < 0xffffffffull ? l.carry : 0xffffffffull);
        a.status[1] = l.carry > a.capacity ? 1u : 0u;
    }
}

// The bin sizes: runs counted per brick in the table, then one add per
// table entry.  One walk per pair: each DDA is set up as it is walked
// (the index loads of the thread's pairs issued first), which keeps fewer
// registers live than PairWalks.
__device__ __forceinline__ void bin_count_body(const BuildArgs& a, BinTable& t, uint32_t block) {
    const uint32_t r0 = block * a.pairs_per_block;
    const uint32_t r1 = min(a.n_pairs, r0 + a.pairs_per_block);
    uint2 pr[kPairsPerThread];
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        const uint32_t s = r0 + threadIdx.x + j * blockDim.x;
        pr[j] = s < r1 ? make_uint2(a.indices[2u * s], a.indices[2u * s + 1u]) : make_uint2(kBuildNone, kBuildNone);
    }
    table_clear(t);
    __syncthreads();
    const auto count = [&](uint32_t k) {  // one run (walk_runs' split)
        const uint32_t slot = table_insert(t, k);
        if (slot != kBuildNone)
            table_add(&t.count[slot], 1u);
        else
            atomicAdd(&a.bins[k], 1u);
    };
#pragma unroll
    for (uint32_t j = 0; j < kPairsPerThread; ++j) {
        SegmentDda w = segment_dda(a.g, a.xyz, a.n_vertices, pr[j].x, pr[j].y);  // kBuildNone: no samples
        uint32_t last = kBuildNone, len = 0;
        while (w.steps > 0.0f) {
            w.steps -= 1.0f;
            const uint32_t k = brick_index(a.b, voxel_coords(a.g, w.rx, w.ry, w.rz));
            if (k != last || len == kMaxRun) {
                count(k);
                len = 0;
            }
            ++len;
            last = k;
            w.advance();
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x)
        if (t.key[i] != kBuildNone) atomicAdd(&a.bins[t.key[i]], t.count[i]);
}

// The bin entries: the same counts in the table, one range reserved per
// table entry in its brick's bin, then a second walk of the same DDAs that
// places each run in its range (table misses take a slot of their own):
// the run's first sample and the DDA state there, from which the brick
// pass walks on with the same float adds.  With the overflow flag set the
// workgroups zero the workspace instead.
__device__ __forceinline__ void bin_fill_body(const BuildArgs& a, BinTable& t, uint32_t block, uint32_t blocks,
                                              size_t n_voxels) {
    if (a.status[1]) {
        uint32_t* acc = a.acc ? reinterpret_cast<uint32_t*>(a.acc) : a.acc_counts;
        const size_t words = overflow_words(a, n_voxels);
        const size_t stride = size_t(blocks) * blockDim.x;
        for (size_t i = size_t(block) * blockDim.x + threadIdx.x; i < words; i += stride) acc[i] = 0u;
        return;
    }
    PairWalks p;
    load_walks(a, block, p);
    table_clear(t);
    __syncthreads();
    walk_runs(a, p, [&](uint32_t, const SegmentDda&, float, uint32_t k) {
        const uint32_t slot = table_insert(t, k);
        if (slot != kBuildNone) table_add(&t.count[slot], 1u);
    });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kTableSlots; i += blockDim.x) {
        if (t.key[i] != kBuildNone) t.base[i] = atomicAdd(&a.cursor[t.key[i]], t.count[i]);
        t.count[i] = 0u;
    }
    __syncthreads();
    walk_runs(a, p, [&](uint32_t j, const SegmentDda& w, float before, uint32_t k) {
        const uint32_t slot = table_find(t, k);
        const uint32_t pos =
            slot != kBuildNone ? t.base[slot] + table_add(&t.count[slot], 1u) : atomicAdd(&a.cursor[k], 1u);
        if (pos < a.capacity) {
            // samples left (this one included): the walk's iterations, ceil(before)
            const float left = ceilf(before);
            const unsigned long long word =
                p.tangent[j] | (left < float(kMaxRun) ? (unsigned long long)(left) : (unsigned long long)(kMaxRun));
            float4* e = reinterpret_cast<float4*>(a.entries + pos);
            e[0] = make_float4(w.rx, w.ry, w.rz, w.dx);
            e[1] = make_float4(w.dy, w.dz, __uint_as_float(uint32_t(word)), __uint_as_float(uint32_t(word >> 32)));
        }
    });
}

// Brick of a workgroup: groups of 8 consecutive bricks along x go to the
// same XCD (workgroups are dealt to the 8 XCDs in turn), so the rows of a
// group's output are written through one L2.  Bricks past the grid
// (groups padded to 8 along x) return kBuildNone.
__device__ __forceinline__ uint32_t brick_of_block(const BrickGrid& b, uint32_t block) {
    const uint32_t xcd = block & 7u, j = block >> 3;
    const uint32_t group = (j >> 3) * 8u + xcd, within = j & 7u;
    const uint32_t gx = (b.bx + 7u) >> 3;  // groups along x
    const uint32_t row = group / gx, x = (group % gx) * 8u + within;
    if (x >= b.bx || row >= b.by * b.bz) return kBuildNone;
    return x + b.bx * row;
}

// Workgroups of the brick pass: the groups rounded up to a multiple of 8,
// so that brick_of_block reaches every group (a multiple of 64).
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
// instead.  block is a multiple-of-8 offset from the launch's own block
// index, so block & 7 is the XCD.
struct BrickLds {
    uint32_t count[512];
    int sum[3][512];
};

__device__ __forceinline__ void brick_body(const BuildArgs& a, BrickLds& l, uint32_t block, uint32_t blocks) {
    constexpr uint32_t kVox = 512;
    if (a.status[1]) {
        // grid-stride over wave-aligned ranges (voxelize_segment is wave-cooperative)
        const float* tangents = a.tangent_i8x4 ? a.tangents : a.xyz;  // sums unused without a tangent volume
        const uint32_t stride = blocks * blockDim.x;
        for (uint32_t base = block * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < a.n_pairs;
             base += stride) {
            const uint32_t s = base + (threadIdx.x & uint32_t(kWave - 1));
            const bool has = s < a.n_pairs;
            voxelize_segment(a.g, a.xyz, a.n_vertices, has, has ? a.indices[2u * s] : 0u,
                             has ? a.indices[2u * s + 1u] : 0u, tangents, a.acc ? nullptr : a.acc_counts, a.acc);
        }
        return;
    }
    const uint32_t brick = brick_of_block(a.b, block);
    if (brick == kBuildNone) return;
    const uint32_t bxo = brick % a.b.bx, byo = (brick / a.b.bx) % a.b.by, bzo = brick / (a.b.bx * a.b.by);
    const uint32_t end = a.cursor[brick], n = a.bins[brick];  // workgroup-uniform
    const bool with_tangents = a.tangent_i8x4 != nullptr;
    // With nx % 4 == 0 thread t writes voxels 4 (t & 1) .. +3 of brick row
    // t >> 1 (y = row & 7, z = row >> 3), whole words at a time; else one
    // voxel at a time.
    static_assert(kBrickBlock * 4 == kVox, "one quad of voxels per thread");
    const bool quads = (a.g.nx & 3u) == 0u;
    const uint32_t t = threadIdx.x, row = t >> 1, q0 = 8u * row + 4u * (t & 1u);
    const uint32_t qx = bxo * 8u + 4u * (t & 1u), qy = byo * 8u + (row & 7u), qz = bzo * 8u + (row >> 3);
    const bool quad_in = qx < a.g.nx && qy < a.g.ny && qz < a.g.nz;  // the quad is whole when nx % 4 == 0
    const size_t qv = size_t(qx) + size_t(a.g.nx) * (size_t(qy) + size_t(a.g.ny) * qz);
    uint32_t lo = 255u, hi = 0u;
    if (n == 0u) {
        // an empty bin: every voxel of the brick is 0, and no LDS is needed
        if (quads) {
            if (quad_in) {
                *reinterpret_cast<uint32_t*>(a.density + qv) = 0u;
                if (a.counts) *reinterpret_cast<uint4*>(a.counts + qv) = make_uint4(0u, 0u, 0u, 0u);
                if (with_tangents) *reinterpret_cast<uint4*>(a.tangent_i8x4 + 4 * qv) = make_uint4(0u, 0u, 0u, 0u);
            }
        } else {
            for (uint32_t i = t; i < kVox; i += kBrickBlock) {
                const uint32_t x = bxo * 8u + (i & 7u), y = byo * 8u + ((i >> 3) & 7u), z = bzo * 8u + (i >> 6);
                if (x >= a.g.nx || y >= a.g.ny || z >= a.g.nz) continue;
                const size_t v = size_t(x) + size_t(a.g.nx) * (size_t(y) + size_t(a.g.ny) * z);
                a.density[v] = 0u;
                if (a.counts) a.counts[v] = 0u;
                if (with_tangents) reinterpret_cast<char4*>(a.tangent_i8x4)[v] = make_char4(0, 0, 0, 0);
            }
        }
        if (t == 0) a.brick_mm[brick] = 0u;  // the brick holds a voxel of the volume (brick_of_block)
        return;
    }
    reinterpret_cast<uint4*>(l.count)[t] = make_uint4(0u, 0u, 0u, 0u);
    reinterpret_cast<int4*>(l.sum[0])[t] = make_int4(0, 0, 0, 0);
    reinterpret_cast<int4*>(l.sum[1])[t] = make_int4(0, 0, 0, 0);
    reinterpret_cast<int4*>(l.sum[2])[t] = make_int4(0, 0, 0, 0);
    __syncthreads();
    // kEntriesInFlight entries per thread per round: their loads are issued
    // before the first walk
    constexpr uint32_t kEntriesInFlight = 2;
    for (uint32_t e0 = end - n + t; e0 < end; e0 += kEntriesInFlight * kBrickBlock) {
        float4 p0[kEntriesInFlight], p1[kEntriesInFlight];
#pragma unroll
        for (uint32_t u = 0; u < kEntriesInFlight; ++u) {
            const uint32_t e = e0 + u * kBrickBlock;
            p0[u] = p1[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // a sample limit of 0: no samples
            if (e < end) {
                const float4* en = reinterpret_cast<const float4*>(a.entries + e);
                p0[u] = en[0];
                p1[u] = en[1];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kEntriesInFlight; ++u) {
            SegmentDda w = {p0[u].x, p0[u].y, p0[u].z, p0[u].w, p1[u].x, p1[u].y, 0.0f};
            const unsigned long long word =
                (unsigned long long)__float_as_uint(p1[u].z) | ((unsigned long long)__float_as_uint(p1[u].w) << 32);
            const uint32_t limit = uint32_t(word & 1023ull);
            const int tx = int((word >> 10) & 0x3ffffull) - 65536, ty = int((word >> 28) & 0x3ffffull) - 65536,
                      tz = int((word >> 46) & 0x3ffffull) - 65536;
            // the run starts inside the brick and, once out, never comes back
            for (uint32_t k = 0; k < limit; ++k) {
                const uint3 v = voxel_coords(a.g, w.rx, w.ry, w.rz);
                if ((v.x >> 3) != bxo || (v.y >> 3) != byo || (v.z >> 3) != bzo) break;
                const uint32_t i = (v.x & 7u) + 8u * ((v.y & 7u) + 8u * (v.z & 7u));
                atomicAdd(&l.count[i], 1u);
                if (with_tangents) {  // wraps only past 32767 contributions, where the tangent is 0
                    atomicAdd(reinterpret_cast<unsigned*>(&l.sum[0][i]), unsigned(tx));
                    atomicAdd(reinterpret_cast<unsigned*>(&l.sum[1][i]), unsigned(ty));
                    atomicAdd(reinterpret_cast<unsigned*>(&l.sum[2][i]), unsigned(tz));
                }
                w.advance();
            }
        }
    }
    __syncthreads();
    if (quads) {
        if (quad_in) {
            const uint4 c4 = *reinterpret_cast<const uint4*>(&l.count[q0]);
            const uint32_t c[4] = {c4.x, c4.y, c4.z, c4.w};
            uint32_t dens = 0u, tq[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (uint32_t m = 0; m < 4; ++m) {
                const uint32_t d = c[m] < 255u ? c[m] : 255u;
                dens |= d << (8u * m);
                lo = d < lo ? d : lo;
                hi = d > hi ? d : hi;
                if (with_tangents && c[m]) {
                    const char4 k = quantize_tangent_sums(l.sum[0][q0 + m], l.sum[1][q0 + m], l.sum[2][q0 + m], c[m]);
                    tq[m] = uint32_t(uint8_t(k.x)) | (uint32_t(uint8_t(k.y)) << 8) | (uint32_t(uint8_t(k.z)) << 16) |
                            (uint32_t(uint8_t(k.w)) << 24);
                }
            }
            *reinterpret_cast<uint32_t*>(a.density + qv) = dens;
            if (a.counts) *reinterpret_cast<uint4*>(a.counts + qv) = c4;
            if (with_tangents) *reinterpret_cast<uint4*>(a.tangent_i8x4 + 4 * qv) = make_uint4(tq[0], tq[1], tq[2], tq[3]);
        }
    } else {
        for (uint32_t i = t; i < kVox; i += kBrickBlock) {
            const uint32_t x = bxo * 8u + (i & 7u), y = byo * 8u + ((i >> 3) & 7u), z = bzo * 8u + (i >> 6);
            if (x >= a.g.nx || y >= a.g.ny || z >= a.g.nz) continue;
            const size_t v = size_t(x) + size_t(a.g.nx) * (size_t(y) + size_t(a.g.ny) * z);
            const uint32_t c = l.count[i];
            const uint32_t d = c < 255u ? c : 255u;
            a.density[v] = uint8_t(d);
            if (a.counts) a.counts[v] = c;
            if (with_tangents)
                reinterpret_cast<char4*>(a.tangent_i8x4)[v] =
                    quantize_tangent_sums(l.sum[0][i], l.sum[1][i], l.sum[2][i], c);
            lo = d < lo ? d : lo;
            hi = d > hi ? d : hi;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const uint32_t p = __shfl_xor(lo, o, kWave), q = __shfl_xor(hi, o, kWave);
        lo = p < lo ? p : lo;
        hi = q > hi ? q : hi;
    }
    // each wave's min/max into its half of the brick's word (folded into
    // the slots by the finalize pass, off this pass's critical path)
    if (threadIdx.x % kWave == 0)
        reinterpret_cast<uint16_t*>(a.brick_mm)[2u * brick + threadIdx.x / kWave] = uint16_t(lo | (hi << 8));
}

// The min/max of the brick pass's words folded into the slots (one fold
// per workgroup), or, with the overflow flag set, the finalize of the
// workspace path: the outputs of the brick pass from the workspace.
__device__ __forceinline__ void overflow_finalize_body(const BuildArgs& a, size_t n, uint32_t block, uint32_t blocks,
                                                       uint32_t (&red)[2][256 / kWave]) {
    uint32_t lo = 255u, hi = 0u;
    if (!a.status[1]) {
        if (size_t(block) * blockDim.x >= a.n_bricks) return;  // workgroup-uniform
        for (size_t i = size_t(block) * blockDim.x + threadIdx.x; i < a.n_bricks; i += size_t(blocks) * blockDim.x) {
            const uint32_t w = a.brick_mm[i];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t l0 = (w >> (16 * h)) & 0xffu, h0 = (w >> (16 * h + 8)) & 0xffu;
                lo = l0 < lo ? l0 : lo;
                hi = h0 > hi ? h0 : hi;
            }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const uint32_t p = __shfl_xor(lo, o, kWave), q = __shfl_xor(hi, o, kWave);
            lo = p < lo ? p : lo;
            hi = q > hi ? q : hi;
        }
        const uint32_t wave = threadIdx.x / kWave;
        if (threadIdx.x % kWave == 0) {
            red[0][wave] = lo;
            red[1][wave] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t w = 1; w < blockDim.x / kWave; ++w) {
                lo = red[0][w] < lo ? red[0][w] : lo;
                hi = red[1][w] > hi ? red[1][w] : hi;
            }
            if (lo <= hi) fold_minmax(a.minmax + (block % kMinMaxSlots) * kMinMaxStride, lo, hi);
        }
        return;
    }
    const size_t stride = size_t(blocks) * blockDim.x;
    const size_t first = size_t(block) * blockDim.x + threadIdx.x;
    for (size_t i = first; i < n; i += stride) {
        long long w0 = 0, w1 = 0;
        uint32_t c;
        if (a.acc) {
            w0 = a.acc[2 * i];
            w1 = a.acc[2 * i + 1];
            c = uint32_t((unsigned long long)(w1) & 0xffffffffull);
        } else {
            c = a.acc_counts[i];
        }
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
        fold_minmax(a.minmax + (block % kMinMaxSlots) * kMinMaxStride, lo, hi);
}

// The map of every uint8 value of the normalize (from the min/max slots)
// into lut; returns whether it is the identity (no normalize, or a volume
// spanning 0..255), for which the cells pass needs no lookups.  All
// threads of the workgroup call it.
__device__ __forceinline__ bool build_lut(const uint32_t* minmax, bool normalize, uint8_t* lut) {
    const DensityMap m = density_map_slots(minmax, normalize);
    bool identity = true;
    for (uint32_t v = threadIdx.x; v < 256u; v += blockDim.x) {
        const uint32_t d = texel_density(v, m);
        lut[v] = uint8_t(d);
        identity = identity && d == v;
    }
    return __syncthreads_and(identity);
}

}  // namespace vkhr
