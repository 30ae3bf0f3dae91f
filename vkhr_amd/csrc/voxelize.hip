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
#include "common.hpp"

#include <cmath>

namespace vkhr {
namespace {

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

__device__ __forceinline__ void add_tangent(float* __restrict__ sums, uint32_t v, float3 t) {
    // Non-returning float atomics resolve in L2.
    unsafeAtomicAdd(&sums[3u * v + 0u], t.x);
    unsafeAtomicAdd(&sums[3u * v + 1u], t.y);
    unsafeAtomicAdd(&sums[3u * v + 2u], t.z);
}

__global__ void __launch_bounds__(256)
voxelize_vertices_kernel(const float* __restrict__ xyz, uint32_t n_vertices,
                         const float* __restrict__ tangents, VoxelGrid g,
                         uint32_t* __restrict__ counts, float* __restrict__ sums) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_vertices; i += stride) {
        float3 p = load_xyz(xyz, i);
        // (vertex - origin) / voxel_size, per component (hair_style.cc:274)
        float px = (p.x - g.ox) / g.vx;
        float py = (p.y - g.oy) / g.vy;
        float pz = (p.z - g.oz) / g.vz;
        uint32_t v = voxel_of(g, px, py, pz);
        atomicAdd(&counts[v], 1u);
        if (sums) add_tangent(sums, v, load_xyz(tangents, i));
    }
}

__global__ void __launch_bounds__(256)
voxelize_segments_kernel(const float* __restrict__ xyz, uint32_t n_vertices,
                         const uint32_t* __restrict__ indices, uint32_t n_pairs, const float* __restrict__ tangents, VoxelGrid g,
                         uint32_t* __restrict__ counts, float* __restrict__ sums) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_pairs; s += stride) {
        uint2 ab = make_uint2(indices[2u * s], indices[2u * s + 1u]);
        if (ab.x >= n_vertices || ab.y >= n_vertices) continue;  // malformed pair: skipped
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
        if (!(steps <= kMaxSteps)) continue;
        dx /= steps;
        dy /= steps;
        dz /= steps;
        float3 t = make_float3(0.0f, 0.0f, 0.0f);
        if (sums) t = load_xyz(tangents, ab.x);
        // while (steps-- > 0.0f): ceil(steps) samples, root accumulated (:319-328)
        while (steps > 0.0f) {
            steps -= 1.0f;
            uint32_t v = voxel_of(g, rx, ry, rz);
            atomicAdd(&counts[v], 1u);
            if (sums) add_tangent(sums, v, t);
            rx += dx;
            ry += dy;
            rz += dz;
        }
    }
}

int make_grid(const vkhr_aabb* bounds, uint32_t nx, uint32_t ny, uint32_t nz, VoxelGrid* g) {
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

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_voxelize_vertices(const float* xyz, uint32_t n_vertices, const float* tangents,
                                      const vkhr_aabb* bounds, uint32_t nx, uint32_t ny, uint32_t nz,
                                      uint32_t* counts, float* tangent_sums, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VoxelGrid g;
    if ((rc = make_grid(bounds, nx, ny, nz, &g))) return rc;
    VKHR_REQUIRE(counts, VKHR_ERR_INVALID, "counts is null");
    VKHR_REQUIRE(n_vertices == 0 || xyz, VKHR_ERR_INVALID, "xyz is null");
    VKHR_REQUIRE(!tangent_sums || tangents, VKHR_ERR_INVALID, "tangent_sums needs tangents");
    hipStream_t s = as_stream(stream);
    size_t n = size_t(nx) * ny * nz;
    if ((rc = fill_u32(counts, n, 0u, s))) return rc;
    if (tangent_sums && (rc = fill_u32(reinterpret_cast<uint32_t*>(tangent_sums), 3 * n, 0u, s))) return rc;
    if (n_vertices == 0) return VKHR_OK;
    voxelize_vertices_kernel<<<grid_for(n_vertices, 256, 65536), 256, 0, s>>>(xyz, n_vertices, tangents, g,
                                                                             counts, tangent_sums);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_voxelize_segments(const float* xyz, uint32_t n_vertices, const uint32_t* indices,
                                      uint32_t n_indices, const float* tangents, const vkhr_aabb* bounds,
                                      uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* counts,
                                      float* tangent_sums, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VoxelGrid g;
    if ((rc = make_grid(bounds, nx, ny, nz, &g))) return rc;
    VKHR_REQUIRE(counts, VKHR_ERR_INVALID, "counts is null");
    uint32_t n_pairs = n_indices / 2;  // the reference loops while i < size-1
    VKHR_REQUIRE(n_pairs == 0 || (xyz && indices && n_vertices), VKHR_ERR_INVALID, "xyz/indices is null");
    VKHR_REQUIRE(!tangent_sums || tangents, VKHR_ERR_INVALID, "tangent_sums needs tangents");
    hipStream_t s = as_stream(stream);
    size_t n = size_t(nx) * ny * nz;
    if ((rc = fill_u32(counts, n, 0u, s))) return rc;
    if (tangent_sums && (rc = fill_u32(reinterpret_cast<uint32_t*>(tangent_sums), 3 * n, 0u, s))) return rc;
    if (n_pairs == 0) return VKHR_OK;
    voxelize_segments_kernel<<<grid_for(n_pairs, 256, 65536), 256, 0, s>>>(xyz, n_vertices, indices, n_pairs,
                                                                           tangents, g, counts, tangent_sums);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
