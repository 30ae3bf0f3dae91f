// The load-time volume build in one call: HairStyle::voxelize_segments
// (src/vkhr/scene_graph/hair_style.cc:296-342) and Volume::normalize
// (:344-357), run back to back as the reference's load path does
// (src/vkhr/rasterizer/hair_style.cc:75-77), with every voxel sum kept in
// LDS instead of device-wide atomics.
//
// Bricks of 8x8x8 voxels.  The index pairs are binned by the bricks their
// samples visit -- a counting sort: count (one add per brick a pair visits,
// aggregated over the lanes of a wave), scan, fill -- then one workgroup
// per brick walks the DDA of every pair in its bin and adds the samples
// that fall inside its brick into an LDS accumulator (count and fixed-point
// tangent sums: integer adds, so the order of the adds does not matter).
// The workgroup finalizes its brick straight from LDS: the saturating uint8
// view, the quantized tangent (the sums and the expression of
// vkhr_volume_finalize, so both give the same bits), the counts when asked
// for, and the min/max of the uint8 view for the normalize.  The normalize,
// the float view, the corner-packed cells and the brick distance field
// follow as in vkhr_volume_finalize.
//
// A pair's voxel coordinates are monotone in its sample index (the root is
// advanced by a fixed direction and the clamps are monotone), so the bricks
// it visits form runs that never repeat: a pair lands in a bin at most once.
//
// The bin entries live in the per-stream scratch, 4 per pair plus 64 Ki.
// If the pairs visit more bricks than that (long segments), a flag set on
// the device switches the build to the workspace path: the packed
// accumulator of vkhr_voxelize_segments in the caller's workspace,
// device-wide atomics, then the finalize of vkhr_volume_finalize.  Its
// kernels are always enqueued and return at once while the flag is clear,
// so the host never waits for the count.
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
    size_t entries;     // bin entries: pair indices
    size_t words;
};

__host__ __forceinline__ BuildLayout build_layout(size_t n_bricks, size_t capacity) {
    auto up = [](size_t w) { return (w + 63) & ~size_t(63); };  // 256-byte aligned sections
    BuildLayout l;
    l.bins = 0;
    l.cursor = up(n_bricks);
    l.status = l.cursor + up(n_bricks);
    l.minmax = l.status + 64;
    l.entries = l.minmax + up(kMinMaxBytes / 4);
    l.words = l.entries + up(capacity);
    return l;
}

struct BuildArgs {
    VoxelGrid g;
    BrickGrid b;
    const float* xyz;
    const uint32_t* indices;
    const float* tangents;          // per vertex (null: no tangent volume)
    uint32_t n_vertices, n_pairs;
    uint32_t* bins;
    uint32_t* cursor;
    uint32_t* status;
    uint32_t* minmax;
    uint32_t* entries;
    uint32_t capacity;
    uint32_t* counts;               // outputs (counts and tangents nullable)
    uint8_t* density;
    int8_t* tangent_i8x4;
    long long* acc;                 // workspace of the overflow path
};

// The bricks a wave's lanes enter in this step (kNone for none), one
// atomic per distinct brick.  Every lane of the wave calls it.
__device__ __forceinline__ void count_runs(uint32_t key, uint32_t* bins) {
    const int lane = __lane_id();
    unsigned long long pending = __ballot(key != kNone);
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const uint32_t k = __shfl(key, leader, kWave);
        const unsigned long long same = __ballot(key == k);
        if (lane == leader) atomicAdd(&bins[k], uint32_t(__popcll(same)));
        pending &= ~same;
    }
}

// As count_runs, taking a slot in each entered brick's bin for pair s.
__device__ __forceinline__ void fill_runs(uint32_t key, uint32_t s, uint32_t* cursor, uint32_t* entries,
                                          uint32_t capacity) {
    const int lane = __lane_id();
    unsigned long long pending = __ballot(key != kNone);
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const uint32_t k = __shfl(key, leader, kWave);
        const unsigned long long same = __ballot(key == k);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursor[k], uint32_t(__popcll(same)));
        base = __shfl(base, leader, kWave);
        if (key == k) {
            const uint32_t slot = base + uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
            if (slot < capacity) entries[slot] = s;
        }
        pending &= ~same;
    }
}

// Walks the DDA of every pair (one per lane, waves stepping together) and
// reports each brick run: to the bin sizes (kFill = false) or as a bin
// entry (kFill = true).
template <bool kFill>
__global__ void __launch_bounds__(kBuildBlock) bin_kernel(BuildArgs a) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < a.n_pairs;
         base += stride) {
        const uint32_t s = base + (threadIdx.x & uint32_t(kWave - 1));
        SegmentDda w = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (s < a.n_pairs) w = segment_dda(a.g, a.xyz, a.n_vertices, a.indices[2u * s], a.indices[2u * s + 1u]);
        uint32_t last = kNone;
        while (__ballot(w.steps > 0.0f)) {
            uint32_t key = kNone;
            if (w.steps > 0.0f) {
                w.steps -= 1.0f;
                const uint32_t k = brick_index(a.b, voxel_coords(a.g, w.rx, w.ry, w.rz));
                key = k != last ? k : kNone;
                last = k;
                w.advance();
            }
            if constexpr (kFill)
                fill_runs(key, s, a.cursor, a.entries, a.capacity);
            else
                count_runs(key, a.bins);
        }
    }
}

// Exclusive scan of the bin sizes into the bin starts (one block), the
// total, and the overflow flag.
__global__ void __launch_bounds__(1024) scan_kernel(BuildArgs a, uint32_t n_bricks) {
    constexpr uint32_t kPer = 4, kThreads = 1024;
    __shared__ unsigned long long wave_sums[kThreads / kWave];
    __shared__ unsigned long long carry_lds;  // 64-bit: the total of absurd inputs may pass 2^32
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (threadIdx.x == 0) carry_lds = 0;
    __syncthreads();
    for (uint32_t chunk = 0; chunk < n_bricks; chunk += kPer * kThreads) {
        uint32_t v[kPer];
        unsigned long long sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t i = chunk + kPer * threadIdx.x + j;
            v[j] = i < n_bricks ? a.bins[i] : 0u;
            sum += v[j];
        }
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
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            const uint32_t i = chunk + kPer * threadIdx.x + j;
            // a wrapped start only happens with the overflow flag set, and
            // then no bin is read
            if (i < n_bricks) a.cursor[i] = uint32_t(before);
            before += v[j];
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
// it, accumulated in LDS, then the brick's outputs.
__global__ void __launch_bounds__(kBuildBlock) brick_kernel(BuildArgs a) {
    constexpr uint32_t kVox = 512;
    __shared__ uint32_t count[kVox];
    __shared__ long long sum[3][kVox];
    __shared__ uint32_t red[2][kBuildBlock / kWave];
    const uint32_t brick = brick_of_block(a.b, blockIdx.x);
    if (brick == kNone || a.status[1]) return;  // the overflow path writes the outputs
    for (uint32_t l = threadIdx.x; l < kVox; l += kBuildBlock) {
        count[l] = 0u;
        sum[0][l] = sum[1][l] = sum[2][l] = 0;
    }
    const uint32_t bxo = brick % a.b.bx, byo = (brick / a.b.bx) % a.b.by, bzo = brick / (a.b.bx * a.b.by);
    __syncthreads();
    const uint32_t end = a.cursor[brick], n = a.bins[brick];
    const bool with_tangents = a.tangent_i8x4 != nullptr;
    for (uint32_t e = end - n + threadIdx.x; e < end; e += kBuildBlock) {
        const uint32_t s = a.entries[e];
        const uint32_t ia = a.indices[2u * s], ib = a.indices[2u * s + 1u];
        SegmentDda w = segment_dda(a.g, a.xyz, a.n_vertices, ia, ib);
        long long tx = 0, ty = 0, tz = 0;
        if (with_tangents && w.steps > 0.0f) {
            const float3 t = load_xyz(a.tangents, ia);
            tx = tangent_fixed(t.x);
            ty = tangent_fixed(t.y);
            tz = tangent_fixed(t.z);
        }
        bool entered = false;
        while (w.steps > 0.0f) {
            w.steps -= 1.0f;
            const uint3 v = voxel_coords(a.g, w.rx, w.ry, w.rz);
            const bool inside = (v.x >> 3) == bxo && (v.y >> 3) == byo && (v.z >> 3) == bzo;
            if (inside) {
                const uint32_t l = (v.x & 7u) + 8u * ((v.y & 7u) + 8u * (v.z & 7u));
                atomicAdd(&count[l], 1u);
                if (with_tangents) {
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sum[0][l]), (unsigned long long)tx);
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sum[1][l]), (unsigned long long)ty);
                    atomicAdd(reinterpret_cast<unsigned long long*>(&sum[2][l]), (unsigned long long)tz);
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
    for (uint32_t l = threadIdx.x; l < kVox; l += kBuildBlock) {
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
        for (uint32_t w = 1; w < kBuildBlock / kWave; ++w) {
            lo = red[0][w] < lo ? red[0][w] : lo;
            hi = red[1][w] > hi ? red[1][w] : hi;
        }
        if (lo <= hi) {  // the brick holds a voxel of the volume
            uint32_t* slot = a.minmax + (blockIdx.x % kMinMaxSlots) * kMinMaxStride;
            atomicMin(&slot[0], lo);
            atomicMax(&slot[1], hi);
        }
    }
}

__global__ void __launch_bounds__(64) build_init_kernel(BuildArgs a) {
    if (threadIdx.x < kMinMaxSlots) {
        a.minmax[threadIdx.x * kMinMaxStride] = 255u;
        a.minmax[threadIdx.x * kMinMaxStride + 1] = 0u;
    }
}

// --- the workspace path, taken only when the bins overflowed

__global__ void __launch_bounds__(256) overflow_clear_kernel(BuildArgs a, size_t words) {
    if (!a.status[1]) return;
    uint32_t* acc = reinterpret_cast<uint32_t*>(a.acc);
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) acc[i] = 0u;
}

__global__ void __launch_bounds__(256) overflow_voxelize_kernel(BuildArgs a) {
    if (!a.status[1]) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < a.n_pairs;
         base += stride) {
        const uint32_t s = base + (threadIdx.x & uint32_t(kWave - 1));
        const bool has = s < a.n_pairs;
        // tangent sums only matter for the tangent volume; the packed count
        // is exact either way
        const float* tangents = a.tangent_i8x4 ? a.tangents : nullptr;
        voxelize_segment(a.g, a.xyz, a.n_vertices, has, has ? a.indices[2u * s] : 0u,
                         has ? a.indices[2u * s + 1u] : 0u, tangents ? tangents : a.xyz, nullptr, a.acc);
    }
}

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
    if (threadIdx.x % kWave == 0 && lo <= hi) {
        uint32_t* slot = a.minmax + (blockIdx.x % kMinMaxSlots) * kMinMaxStride;
        atomicMin(&slot[0], lo);
        atomicMax(&slot[1], hi);
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
    const size_t capacity = std::min<size_t>(4 * size_t(n_pairs) + 65536, size_t(1) << 28);
    const BuildLayout lay = build_layout(n_bricks, capacity);
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
    a.bins = w + lay.bins;
    a.cursor = w + lay.cursor;
    a.status = w + lay.status;
    a.minmax = w + lay.minmax;
    a.entries = w + lay.entries;
    a.capacity = uint32_t(capacity);
    a.counts = counts;
    a.density = density_u8;
    a.tangent_i8x4 = tangent_i8x4;
    a.acc = reinterpret_cast<long long*>(workspace);

    if ((rc = fill_u32(a.bins, n_bricks, 0u, s))) return rc;
    build_init_kernel<<<1, 64, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    const unsigned pair_blocks = grid_for(n_pairs, kBuildBlock, 65536);
    if (n_pairs) {
        bin_kernel<false><<<pair_blocks, kBuildBlock, 0, s>>>(a);
        VKHR_CHECK_LAUNCH();
    }
    scan_kernel<<<1, 1024, 0, s>>>(a, uint32_t(n_bricks));
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        bin_kernel<true><<<pair_blocks, kBuildBlock, 0, s>>>(a);
        VKHR_CHECK_LAUNCH();
    }
    brick_kernel<<<brick_blocks(bg), kBuildBlock, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    // the workspace path: every block returns at once unless the bins overflowed
    const size_t n = size_t(nx) * ny * nz;
    overflow_clear_kernel<<<1024, 256, 0, s>>>(a, 4 * n);
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        overflow_voxelize_kernel<<<grid_for(n_pairs, 256, 1024), 256, 0, s>>>(a);
        VKHR_CHECK_LAUNCH();
    }
    overflow_finalize_kernel<<<1024, 256, 0, s>>>(a, n);
    VKHR_CHECK_LAUNCH();
    return finish_volume(density_u8, nx, ny, nz, normalize, density_f32, cells, occupancy,
                         normalize ? a.minmax : nullptr, s);
}
