// Load-time preparation of a hair style: the reference's generators
// HairStyle::generate_bounding_box / generate_indices / generate_tangents
// (src/vkhr/scene_graph/hair_style.cc:171-255) on the device.
// Compiled with -ffp-contract=off: tangents match the reference's float
// arithmetic bit for bit (glm::normalize = v * (1 / sqrt(dot(v, v))), dot
// summed x, y, then z).
#include "common.hpp"

#include <cmath>

namespace vkhr {
namespace {

constexpr int kAabbBlock = 256;
constexpr int kAabbGrid = 1024;

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

// Stage 1: each block reduces a strided share of the vertices to 6 floats
// (min xyz, max xyz), both started at the origin as the reference does.
__global__ void __launch_bounds__(kAabbBlock)
aabb_partial_kernel(const float* __restrict__ xyz, uint32_t n, float* __restrict__ partial) {
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float3 p = load_xyz(xyz, i);
        mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
    __shared__ float red[6][kAabbBlock / kWave];
    int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    for (int c = 0; c < 3; ++c) {
        float a = wave_min(mn[c]), b = wave_max(mx[c]);
        if (lane == 0) { red[c][wid] = a; red[3 + c][wid] = b; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int c = threadIdx.x;
        float r = red[c][0];
        for (int w = 1; w < kAabbBlock / kWave; ++w) r = c < 3 ? fminf(r, red[c][w]) : fmaxf(r, red[c][w]);
        partial[6 * blockIdx.x + c] = r;
    }
}

// Stage 2: one wave folds the per-block results (lanes stride over blocks).
__global__ void __launch_bounds__(kWave)
aabb_final_kernel(const float* __restrict__ partial, int blocks, float* __restrict__ out) {
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    for (int b = threadIdx.x; b < blocks; b += kWave) {
        for (int c = 0; c < 3; ++c) {
            mn[c] = fminf(mn[c], partial[6 * b + c]);
            mx[c] = fmaxf(mx[c], partial[6 * b + 3 + c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        float a = wave_min(mn[c]), b = wave_max(mx[c]);
        if (threadIdx.x == 0) { out[c] = a; out[3 + c] = b; }
    }
}

__global__ void __launch_bounds__(256)
indices_kernel(const uint32_t* __restrict__ offsets, uint32_t n_strands, uint32_t* __restrict__ indices) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_strands; s += stride) {
        uint32_t first = offsets[s], end = offsets[s + 1];
        uint32_t out = 2u * (first - s);  // every earlier strand skipped its last vertex
        for (uint32_t v = first; v + 1u < end; ++v) {
            indices[out++] = v;
            indices[out++] = v + 1u;
        }
    }
}

__device__ __forceinline__ float3 glm_normalize(float3 v) {
    float d = (v.x * v.x + v.y * v.y) + v.z * v.z;
    float r = 1.0f / sqrtf(d);
    return make_float3(v.x * r, v.y * r, v.z * r);
}

__global__ void __launch_bounds__(256)
tangents_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ offsets, uint32_t n_strands,
                float* __restrict__ tangents) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_strands; s += stride) {
        uint32_t first = offsets[s], end = offsets[s + 1];
        if (end <= first) continue;
        float3 t = make_float3(0.0f, 0.0f, 0.0f);
        float3 cur = load_xyz(xyz, first);
        for (uint32_t v = first; v + 1u < end; ++v) {
            float3 nxt = load_xyz(xyz, v + 1u);
            t = glm_normalize(make_float3(nxt.x - cur.x, nxt.y - cur.y, nxt.z - cur.z));
            tangents[3u * v + 0u] = t.x;
            tangents[3u * v + 1u] = t.y;
            tangents[3u * v + 2u] = t.z;
            cur = nxt;
        }
        // the strand's last vertex repeats the previous tangent
        tangents[3u * (end - 1u) + 0u] = t.x;
        tangents[3u * (end - 1u) + 1u] = t.y;
        tangents[3u * (end - 1u) + 2u] = t.z;
    }
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_aabb_compute(const float* xyz, uint32_t n_vertices, vkhr_aabb* out, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(out, VKHR_ERR_INVALID, "out is null");
    VKHR_REQUIRE(n_vertices == 0 || xyz, VKHR_ERR_INVALID, "xyz is null");
    float mm[6] = {0, 0, 0, 0, 0, 0};
    if (n_vertices) {
        void* scr = nullptr;
        if ((rc = stream_scratch(sizeof(float) * (6 * kAabbGrid + 8), as_stream(stream), &scr))) return rc;
        float* partial = static_cast<float*>(scr);
        float* result = partial + 6 * kAabbGrid;
        hipStream_t s = as_stream(stream);
        int blocks = int(grid_for(n_vertices, kAabbBlock, kAabbGrid));
        aabb_partial_kernel<<<blocks, kAabbBlock, 0, s>>>(xyz, n_vertices, partial);
        VKHR_CHECK_LAUNCH();
        aabb_final_kernel<<<1, kWave, 0, s>>>(partial, blocks, result);
        VKHR_CHECK_LAUNCH();
        VKHR_CHECK_HIP(hipMemcpyAsync(mm, result, sizeof(mm), hipMemcpyDeviceToHost, s));
        VKHR_CHECK_HIP(hipStreamSynchronize(s));
    }
    // get_bounding_box (hair_style.cc:236-255), float arithmetic
    for (int c = 0; c < 3; ++c) {
        out->origin[c] = mm[c];
        out->size[c] = mm[3 + c] - mm[c];
    }
    volatile float sx = out->size[0], sy = out->size[1], sz = out->size[2];
    float xx = sx * sx, yy = sy * sy, zz = sz * sz;
    float dot = (xx + yy) + zz;
    out->radius = std::sqrt(dot);
    float xy = sx * sy;
    out->volume = xy * sz;
    return VKHR_OK;
}

extern "C" int vkhr_generate_indices(const uint32_t* strand_offsets, uint32_t n_strands, uint32_t* indices,
                                     vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    if (n_strands == 0) return VKHR_OK;
    VKHR_REQUIRE(strand_offsets && indices, VKHR_ERR_INVALID, "null pointer");
    indices_kernel<<<grid_for(n_strands, 256, 65536), 256, 0, as_stream(stream)>>>(strand_offsets, n_strands,
                                                                                   indices);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_generate_tangents(const float* xyz, const uint32_t* strand_offsets, uint32_t n_strands,
                                      float* tangents, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    if (n_strands == 0) return VKHR_OK;
    VKHR_REQUIRE(xyz && strand_offsets && tangents, VKHR_ERR_INVALID, "null pointer");
    tangents_kernel<<<grid_for(n_strands, 256, 65536), 256, 0, as_stream(stream)>>>(xyz, strand_offsets,
                                                                                    n_strands, tangents);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
