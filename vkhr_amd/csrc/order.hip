// Load-time spatial orders of a style's segments: the index pairs sorted by
// the Morton code of each segment's midpoint (vkhr_sort_segments), or
// strand by strand with the strands sorted by the Morton code of their
// root vertex (vkhr_sort_strands).  No counterpart in the
// reference, whose .hair assets store strands in scalp order; a seeded
// synthetic style roots its strands in random order, and the strand passes
// then gather their shadow, density and depth taps from all over the volume
// within one workgroup.  The output is a permutation of the pairs, so the
// voxel counts, fragments and shadow depths are the same on it.
// Compiled with -ffp-contract=off: the codes are reproducible from float
// arithmetic alone (midpoint, offset, scale, truncation).
#include "common.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

namespace vkhr {
namespace {

constexpr uint32_t kInvalidCode = 1u << 30;  // above every 30-bit code: malformed pairs sort last

// 10 bits of v spread to every third bit.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void __launch_bounds__(256)
segment_codes_kernel(const float* __restrict__ xyz, uint32_t n_vertices, const uint32_t* __restrict__ indices,
                     uint32_t n, vkhr_aabb box, uint32_t* __restrict__ keys, uint32_t* __restrict__ values) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t ia = indices[2ull * k], ib = indices[2ull * k + 1];
    uint32_t code = kInvalidCode;
    if (ia < n_vertices && ib < n_vertices) {
        const float3 a = load_xyz(xyz, ia), b = load_xyz(xyz, ib);
        const float m[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        uint32_t q[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float ext = box.size[c];
            float f = ext > 0.0f ? (m[c] - box.origin[c]) / ext * 1024.0f : 0.0f;
            f = fminf(fmaxf(f, 0.0f), 1023.0f);  // NaN becomes 0
            q[c] = uint32_t(f);
        }
        code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    }
    keys[k] = code;
    values[k] = k;
}

// The first pair of k's strand run: k when the pair does not continue the
// previous one (its first index is not the previous pair's second), to be
// carried forward by an inclusive max-scan.
__global__ void __launch_bounds__(256)
run_heads_kernel(const uint2* __restrict__ pairs, uint32_t n, uint32_t* __restrict__ heads) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    heads[k] = k == 0 || pairs[k - 1].y != pairs[k].x ? k : 0u;
}

// Pair k's key: the Morton code of its strand's root vertex (the first
// index of the run's first pair), as segment_codes_kernel codes a midpoint.
__global__ void __launch_bounds__(256)
strand_codes_kernel(const float* __restrict__ xyz, uint32_t n_vertices, const uint2* __restrict__ pairs, uint32_t n,
                    const uint32_t* __restrict__ first, vkhr_aabb box, uint32_t* __restrict__ keys,
                    uint32_t* __restrict__ values) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t root = pairs[first[k]].x;
    uint32_t code = kInvalidCode;
    if (root < n_vertices) {
        const float3 a = load_xyz(xyz, root);
        const float m[3] = {a.x, a.y, a.z};
        uint32_t q[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float ext = box.size[c];
            float f = ext > 0.0f ? (m[c] - box.origin[c]) / ext * 1024.0f : 0.0f;
            f = fminf(fmaxf(f, 0.0f), 1023.0f);  // NaN becomes 0
            q[c] = uint32_t(f);
        }
        code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    }
    keys[k] = code;
    values[k] = k;
}

__global__ void __launch_bounds__(256)
gather_pairs_kernel(const uint2* __restrict__ pairs, const uint32_t* __restrict__ order, uint32_t n,
                    uint2* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = pairs[order[k]];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Checks shared by the two orders; n_out: the pair count.
int check_order_args(const float* xyz, const uint32_t* indices, uint32_t n_indices, const vkhr_aabb* bounds,
                     const uint32_t* sorted, uint32_t* n_out) {
    VKHR_REQUIRE(bounds && sorted, VKHR_ERR_INVALID, "null argument");
    VKHR_REQUIRE(n_indices % 2 == 0, VKHR_ERR_INVALID, "odd index count %u", n_indices);
    const uint32_t n = n_indices / 2;
    *n_out = n;
    if (n == 0) return VKHR_OK;
    VKHR_REQUIRE(n <= (1u << 28), VKHR_ERR_OVERFLOW, "too many segments (%u)", n);  // one lane per pair
    const auto plausible = [](const void* p) { return reinterpret_cast<uintptr_t>(p) >= 65536u; };
    VKHR_REQUIRE(plausible(xyz) && plausible(indices) && plausible(sorted), VKHR_ERR_INVALID,
                 "null or implausible pointer");
    VKHR_REQUIRE(reinterpret_cast<uintptr_t>(indices) % 8 == 0 && reinterpret_cast<uintptr_t>(sorted) % 8 == 0,
                 VKHR_ERR_INVALID, "index buffers must be 8-byte aligned");
    const uintptr_t in0 = reinterpret_cast<uintptr_t>(indices), out0 = reinterpret_cast<uintptr_t>(sorted);
    const uintptr_t bytes = uintptr_t(n) * 8u;
    VKHR_REQUIRE(in0 + bytes <= out0 || out0 + bytes <= in0, VKHR_ERR_INVALID, "indices and sorted overlap");
    return VKHR_OK;
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_sort_segments(const float* xyz, uint32_t n_vertices, const uint32_t* indices,
                                  uint32_t n_indices, const vkhr_aabb* bounds, uint32_t* sorted,
                                  vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    uint32_t n = 0;
    if ((rc = check_order_args(xyz, indices, n_indices, bounds, sorted, &n)) || n == 0) return rc;
    hipStream_t s = as_stream(stream);
    size_t sort_bytes = 0;
    VKHR_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), size_t(n), 0, 31, s));
    const size_t words = align256(size_t(n) * 4);
    void* tmp = nullptr;
    VKHR_CHECK_HIP(hipMalloc(&tmp, 4 * words + align256(sort_bytes)));
    char* p = static_cast<char*>(tmp);
    uint32_t* keys = reinterpret_cast<uint32_t*>(p);
    uint32_t* values = reinterpret_cast<uint32_t*>(p + words);
    uint32_t* keys_sorted = reinterpret_cast<uint32_t*>(p + 2 * words);
    uint32_t* order = reinterpret_cast<uint32_t*>(p + 3 * words);
    void* sort_tmp = p + 4 * words;
    const uint32_t blocks = grid_for(n, 256);
    segment_codes_kernel<<<blocks, 256, 0, s>>>(xyz, n_vertices, indices, n, *bounds, keys, values);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys_sorted, values, order, size_t(n), 0, 31, s);
    if (e == hipSuccess) {
        gather_pairs_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint2*>(indices), order, n,
                                                   reinterpret_cast<uint2*>(sorted));
        e = hipGetLastError();
    }
    const hipError_t se = hipStreamSynchronize(s);  // the temporaries are freed below
    (void)hipFree(tmp);
    if (e == hipSuccess) e = se;
    VKHR_REQUIRE(e == hipSuccess, VKHR_ERR_HIP, "segment sort failed: %s", hipGetErrorString(e));
    return VKHR_OK;
}

extern "C" int vkhr_sort_strands(const float* xyz, uint32_t n_vertices, const uint32_t* indices, uint32_t n_indices,
                                 const vkhr_aabb* bounds, uint32_t* sorted, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    uint32_t n = 0;
    if ((rc = check_order_args(xyz, indices, n_indices, bounds, sorted, &n)) || n == 0) return rc;
    hipStream_t s = as_stream(stream);
    size_t sort_bytes = 0, scan_bytes = 0;
    VKHR_CHECK_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), size_t(n), 0, 31, s));
    VKHR_CHECK_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, static_cast<uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), size_t(n), rocprim::maximum<uint32_t>(),
                                           s));
    const size_t words = align256(size_t(n) * 4);
    void* tmp = nullptr;
    VKHR_CHECK_HIP(hipMalloc(&tmp, 5 * words + align256(std::max(sort_bytes, scan_bytes))));
    char* p = static_cast<char*>(tmp);
    uint32_t* keys = reinterpret_cast<uint32_t*>(p);
    uint32_t* values = reinterpret_cast<uint32_t*>(p + words);
    uint32_t* keys_sorted = reinterpret_cast<uint32_t*>(p + 2 * words);
    uint32_t* order = reinterpret_cast<uint32_t*>(p + 3 * words);
    uint32_t* first = reinterpret_cast<uint32_t*>(p + 4 * words);
    void* lib_tmp = p + 5 * words;
    const uint2* pairs = reinterpret_cast<const uint2*>(indices);
    const uint32_t blocks = grid_for(n, 256);
    // heads in `order` (free until the sort), their running maximum in `first`
    run_heads_kernel<<<blocks, 256, 0, s>>>(pairs, n, order);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::inclusive_scan(lib_tmp, scan_bytes, order, first, size_t(n), rocprim::maximum<uint32_t>(), s);
    if (e == hipSuccess) {
        strand_codes_kernel<<<blocks, 256, 0, s>>>(xyz, n_vertices, pairs, n, first, *bounds, keys, values);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(lib_tmp, sort_bytes, keys, keys_sorted, values, order, size_t(n), 0, 31, s);
    if (e == hipSuccess) {
        gather_pairs_kernel<<<blocks, 256, 0, s>>>(pairs, order, n, reinterpret_cast<uint2*>(sorted));
        e = hipGetLastError();
    }
    const hipError_t se = hipStreamSynchronize(s);  // the temporaries are freed below
    (void)hipFree(tmp);
    if (e == hipSuccess) e = se;
    VKHR_REQUIRE(e == hipSuccess, VKHR_ERR_HIP, "strand sort failed: %s", hipGetErrorString(e));
    return VKHR_OK;
}
