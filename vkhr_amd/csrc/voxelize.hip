// Strand voxelization: the reference's CPU voxelizers
// (src/vkhr/scene_graph/hair_style.cc:257-342) as gfx950 kernels.
//
// Counts are integers, so the result is independent of the order in which
// threads add them: one thread walks one segment with exactly the
// reference's float arithmetic (this file is compiled with
// -ffp-contract=off, division and sqrt are IEEE correctly rounded by
// default on HIP) and adds 1 to each voxel it visits.  The reference's
// float linear index x + y*w + z*w*h is formed from integers < 2^24 and is
// therefore exact; it is computed here in integer arithmetic, which is
// identical as long as nx*ny*nz <= 2^24 (enforced).
#include "voxel.hpp"

#include <cmath>

namespace vkhr {
namespace {

// Grid-stride over wave-aligned ranges so that whole waves stay converged
// for the wave-cooperative bodies.
__global__ void __launch_bounds__(256)
voxelize_vertices_kernel(const float* __restrict__ xyz, uint32_t n_vertices,
                         const float* __restrict__ tangents, VoxelGrid g,
                         uint32_t* __restrict__ counts, long long* __restrict__ sums) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < n_vertices;
         base += stride) {
        const uint32_t i = base + (threadIdx.x & uint32_t(kWave - 1));
        voxelize_vertex(g, xyz, i < n_vertices, i, tangents, counts, sums);
    }
}

__global__ void __launch_bounds__(256)
voxelize_segments_kernel(const float* __restrict__ xyz, uint32_t n_vertices,
                         const uint32_t* __restrict__ indices, uint32_t n_pairs, const float* __restrict__ tangents, VoxelGrid g,
                         uint32_t* __restrict__ counts, long long* __restrict__ sums) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~uint32_t(kWave - 1)); base < n_pairs;
         base += stride) {
        const uint32_t s = base + (threadIdx.x & uint32_t(kWave - 1));
        const bool has = s < n_pairs;
        voxelize_segment(g, xyz, n_vertices, has, has ? indices[2u * s] : 0u, has ? indices[2u * s + 1u] : 0u,
                         tangents, counts, sums);
    }
}

// counts[v] = the count held in word 1 of the packed accumulator.
__global__ void __launch_bounds__(256)
unpack_counts_kernel(const long long* __restrict__ acc, uint32_t* __restrict__ counts, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t v = size_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) counts[v] = packed_count(acc, v);
}

// Zeroes what the voxelizer accumulates into: the packed accumulator when
// there is one (the counts are unpacked from it afterwards if asked for),
// else the counts.
int clear_targets(uint32_t* counts, int64_t* acc, size_t n, hipStream_t s) {
    return acc ? fill_u32(reinterpret_cast<uint32_t*>(acc), 4 * n, 0u, s) : fill_u32(counts, n, 0u, s);
}

int unpack_counts(uint32_t* counts, const int64_t* acc, size_t n, hipStream_t s) {
    if (!acc || !counts) return VKHR_OK;
    unpack_counts_kernel<<<grid_for(n, 256, 8192), 256, 0, s>>>(reinterpret_cast<const long long*>(acc), counts, n);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_voxelize_vertices(const float* xyz, uint32_t n_vertices, const float* tangents,
                                      const vkhr_aabb* bounds, uint32_t nx, uint32_t ny, uint32_t nz,
                                      uint32_t* counts, int64_t* tangent_acc, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VoxelGrid g;
    if ((rc = make_grid(bounds, nx, ny, nz, &g))) return rc;
    VKHR_REQUIRE(counts || tangent_acc, VKHR_ERR_INVALID, "neither counts nor tangent_acc");
    VKHR_REQUIRE(n_vertices == 0 || xyz, VKHR_ERR_INVALID, "xyz is null");
    VKHR_REQUIRE(!tangent_acc || tangents, VKHR_ERR_INVALID, "tangent_acc needs tangents");
    hipStream_t s = as_stream(stream);
    size_t n = size_t(nx) * ny * nz;
    if ((rc = clear_targets(counts, tangent_acc, n, s))) return rc;
    if (n_vertices) {
        voxelize_vertices_kernel<<<grid_for(n_vertices, 256, 65536), 256, 0, s>>>(
            xyz, n_vertices, tangents, g, counts, reinterpret_cast<long long*>(tangent_acc));
        VKHR_CHECK_LAUNCH();
    }
    return unpack_counts(counts, tangent_acc, n, s);
}

extern "C" int vkhr_voxelize_segments(const float* xyz, uint32_t n_vertices, const uint32_t* indices,
                                      uint32_t n_indices, const float* tangents, const vkhr_aabb* bounds,
                                      uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* counts,
                                      int64_t* tangent_acc, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VoxelGrid g;
    if ((rc = make_grid(bounds, nx, ny, nz, &g))) return rc;
    VKHR_REQUIRE(counts || tangent_acc, VKHR_ERR_INVALID, "neither counts nor tangent_acc");
    uint32_t n_pairs = n_indices / 2;  // the reference loops while i < size-1
    VKHR_REQUIRE(n_pairs == 0 || (xyz && indices && n_vertices), VKHR_ERR_INVALID, "xyz/indices is null");
    VKHR_REQUIRE(!tangent_acc || tangents, VKHR_ERR_INVALID, "tangent_acc needs tangents");
    hipStream_t s = as_stream(stream);
    size_t n = size_t(nx) * ny * nz;
    if ((rc = clear_targets(counts, tangent_acc, n, s))) return rc;
    if (n_pairs) {
        voxelize_segments_kernel<<<grid_for(n_pairs, 256, 65536), 256, 0, s>>>(
            xyz, n_vertices, indices, n_pairs, tangents, g, counts, reinterpret_cast<long long*>(tangent_acc));
        VKHR_CHECK_LAUNCH();
    }
    return unpack_counts(counts, tangent_acc, n, s);
}
