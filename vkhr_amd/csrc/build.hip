// The load-time volume build of one style in one call: HairStyle::
// voxelize_segments (src/vkhr/scene_graph/hair_style.cc:296-342) and
// Volume::normalize (:344-357), run back to back as the reference's load
// path does (src/vkhr/rasterizer/hair_style.cc:75-77), with every voxel sum
// kept in LDS instead of device-wide atomics (the passes are described
// with their bodies in build.hpp).  The bin entries and the bin state live
// in the per-stream scratch; the brick distance field follows the cells as
// in vkhr_volume_finalize.
//
// Compiled with -ffp-contract=off: the DDA and the quantization must round
// as the reference's float arithmetic does.
#include "build.hpp"

#include <algorithm>

namespace vkhr {
namespace {

__global__ void __launch_bounds__(256) build_init_kernel(BuildArgs a) { build_init_body(a, blockIdx.x, gridDim.x); }

__global__ void __launch_bounds__(kBuildBlock) bin_count_kernel(BuildArgs a) {
    __shared__ BinTable t;
    bin_count_body(a, t, blockIdx.x);
}

__global__ void __launch_bounds__(kScanThreads) scan_kernel(BuildArgs a) {
    __shared__ ScanLds l;
    scan_body(a, l);
}

__global__ void __launch_bounds__(kBuildBlock) bin_fill_kernel(BuildArgs a, size_t n_voxels) {
    __shared__ BinTable t;
    bin_fill_body(a, t, blockIdx.x, gridDim.x, n_voxels);
}

__global__ void __launch_bounds__(kBrickBlock) brick_kernel(BuildArgs a) {
    __shared__ BrickLds l;
    brick_body(a, l, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256) overflow_finalize_kernel(BuildArgs a, size_t n) {
    __shared__ uint32_t red[2][256 / kWave];
    overflow_finalize_body(a, n, blockIdx.x, gridDim.x, red);
}

// The corner-packed cells from the uint8 view with the normalize applied as
// the texels load (the map from the min/max slots), writing the normalized
// volume and its float view on the way (cells_pass modes 1 and 2).
__global__ void __launch_bounds__(256)
mapped_cells_kernel(const uint8_t* __restrict__ raw, uint32_t nx, uint32_t ny, uint32_t nz,
                    const uint32_t* __restrict__ minmax, int normalize, unsigned long long* __restrict__ cells,
                    uint32_t* __restrict__ occupancy, uint8_t* __restrict__ out_u8, float* __restrict__ out_f32) {
    __shared__ uint8_t lut[256];
    const bool identity = build_lut(minmax, normalize != 0, lut);
    const uint32_t n_blocks = cells_pass_blocks(nx, ny, nz);
    for (uint32_t b = blockIdx.x; b < n_blocks; b += gridDim.x) {
        if (identity)
            cells_pass<1>(raw, nx, ny, nz, b, cells, occupancy, nullptr, out_u8, out_f32);
        else
            cells_pass<2>(raw, nx, ny, nz, b, cells, occupancy, lut, out_u8, out_f32);
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
                     reinterpret_cast<uintptr_t>(workspace) % 16 == 0 && reinterpret_cast<uintptr_t>(counts) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(tangent_i8x4) % 16 == 0,
                 VKHR_ERR_INVALID,
                 "density_u8 must be 4-byte and counts, tangent_i8x4, density_f32, workspace 16-byte aligned");
    hipStream_t s = as_stream(stream);
    const BrickGrid bg = brick_grid(nx, ny, nz);
    const size_t n_bricks = size_t(bg.bx) * bg.by * bg.bz;  // <= 2^24 (make_grid bounds the voxels)
    const size_t n = size_t(nx) * ny * nz;
    const size_t capacity = build_capacity(n_pairs);
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
    a.pairs_per_block = build_pairs_per_block(n_pairs);
    const uint32_t bin_blocks = std::max(1u, (n_pairs + a.pairs_per_block - 1) / a.pairs_per_block);
    a.bins = w + lay.bins;
    a.cursor = w + lay.cursor;
    a.status = w + lay.status;
    a.minmax = w + lay.minmax;
    a.brick_mm = w + lay.brick_mm;
    a.entries = reinterpret_cast<BinEntry*>(w + lay.entries);
    a.capacity = uint32_t(capacity);
    a.occupancy = occupancy;
    a.occupancy_words = occupancy ? occupancy_words(nx, ny, nz) : 0;
    a.counts = counts;
    a.density = mapped ? reinterpret_cast<uint8_t*>(w + lay.raw) : density_u8;
    a.tangent_i8x4 = tangent_i8x4;
    a.acc = reinterpret_cast<long long*>(workspace);
    a.acc_counts = nullptr;

    build_init_kernel<<<grid_for(std::max(n_bricks, a.occupancy_words), 256, 1024), 256, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        bin_count_kernel<<<bin_blocks, kBuildBlock, 0, s>>>(a);
        VKHR_CHECK_LAUNCH();
    }
    scan_kernel<<<1, kScanThreads, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    if (n_pairs) {
        bin_fill_kernel<<<bin_blocks, kBuildBlock, 0, s>>>(a, n);
        VKHR_CHECK_LAUNCH();
    }
    brick_kernel<<<brick_blocks(bg), kBrickBlock, 0, s>>>(a);
    VKHR_CHECK_LAUNCH();
    // the min/max of the bricks into the slots, or the workspace path's finalize
    overflow_finalize_kernel<<<1024, 256, 0, s>>>(a, n);
    VKHR_CHECK_LAUNCH();
    if (!mapped)
        return finish_volume(density_u8, nx, ny, nz, normalize, density_f32, nullptr, nullptr,
                             normalize ? a.minmax : nullptr, s);
    mapped_cells_kernel<<<grid_for(cells_pass_blocks(nx, ny, nz), 1, 1u << 20), 256, 0, s>>>(
        a.density, nx, ny, nz, a.minmax, normalize, reinterpret_cast<unsigned long long*>(cells), occupancy,
        density_u8, density_f32);
    VKHR_CHECK_LAUNCH();
    return occupancy ? finish_occupancy(occupancy, nx, ny, nz, s) : VKHR_OK;
}
