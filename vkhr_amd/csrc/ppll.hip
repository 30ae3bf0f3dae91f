// Per-pixel linked list: clear and the k-buffer resolve
// (share/shaders/transparency/resolve.comp:15-102), one thread per pixel.
// The K=16 slots live in registers: every slot access is an unrolled
// compare/select, so no private array spills to scratch memory.
#include "render.hpp"

namespace vkhr {
namespace {

constexpr int kK = 16;
constexpr int kMaxFragments = 1024;
constexpr float kMaxDepth = 3.e+38f;
// LDS the resolve reserves per workgroup and does not use: two workgroups
// (8 waves) fit in a gfx950 CU's 160 KB, three do not.  The lists are
// chains of dependent 12-byte loads scattered over the node pool, and the
// nodes of a workgroup's pixels share 128-byte lines; with fewer waves in
// flight per CU those lines stay in L2 between the loads that share them.
// Measured at 136k strands / 1080p (16.7 M nodes): 5 waves per SIMD (the
// register limit) 0.225 ms, 4: 0.197, 3: 0.179, 2: 0.160, 1: 0.237.
constexpr size_t kResolveLdsReserve = 64 * 1024;

__global__ void __launch_bounds__(256)
ppll_clear_kernel(uint32_t* __restrict__ heads, size_t n, uint32_t* __restrict__ counter) {
    size_t stride = size_t(gridDim.x) * blockDim.x;
    size_t i0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i0 == 0) *counter = 0u;
    for (size_t i = i0; i < n; i += stride) heads[i] = kPpllNull;
}

// b / 255 correctly rounded for an integer 0 <= b <= 255 (unpackUnorm4x8):
// the quotient by the rounded reciprocal, corrected once by its residual
// (Markstein's division step; the residual fma is exact).  Equal to the
// IEEE division for all 256 inputs (checked exhaustively in exact rational
// arithmetic), at 3 instructions instead of a full division's 10.
__device__ __forceinline__ float unorm8(uint32_t b) {
    constexpr float kInv255 = 1.0f / 255.0f;
    const float c = float(b);
    const float q = c * kInv255;
    return __builtin_fmaf(__builtin_fmaf(-q, 255.0f, c), kInv255, q);
}

__device__ __forceinline__ float4 unpack_unorm4x8(uint32_t c) {
    return make_float4(unorm8(c & 0xffu), unorm8((c >> 8) & 0xffu), unorm8((c >> 16) & 0xffu), unorm8(c >> 24));
}

// mix(x, y, a) = x * (1 - a) + y * a, on all four channels
__device__ __forceinline__ float4 blend(float4 x, float4 y) {
    float a = y.w;
    return make_float4(x.x * (1.0f - a) + y.x * a, x.y * (1.0f - a) + y.y * a, x.z * (1.0f - a) + y.z * a,
                       x.w * (1.0f - a) + y.w * a);
}

// Node i of the pool: {colour, depth bits, next node} in one dwordx3 load.
__device__ __forceinline__ uint3 load_node(const uint32_t* __restrict__ nodes, uint32_t i) {
    const uint32_t* n = nodes + 3ull * i;
    return make_uint3(n[0], n[1], n[2]);
}

// Furthest slot: strictly greater than the running maximum, which starts at
// 0 with index 0 (resolve.comp:48-57, 81-90).
__device__ __forceinline__ int furthest(const float (&d)[kK], float& best) {
    best = 0.0f;
    int idx = 0;
#pragma unroll
    for (int k = 0; k < kK; ++k)
        if (d[k] > best) { best = d[k]; idx = k; }
    return idx;
}

__global__ void __launch_bounds__(256)
ppll_resolve_kernel(const uint32_t* __restrict__ heads, const uint32_t* __restrict__ nodes, uint32_t node_count,
                    int width, int height, const float4* __restrict__ background, float4* __restrict__ out) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= width || py >= height) return;
    const size_t pix = size_t(py) * width + px;
    float4 resolved = background ? background[pix] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    uint32_t head = heads[pix];
    if (head == kPpllNull) {
        out[pix] = resolved;
        return;
    }
    float depth[kK];
    uint32_t color[kK];
#pragma unroll
    for (int k = 0; k < kK; ++k) {
        depth[k] = kMaxDepth;
        color[k] = 0u;
    }
    // the K newest nodes (a node index past the pool ends the list)
#pragma unroll
    for (int k = 0; k < kK; ++k) {
        if (head != kPpllNull && head < node_count) {
            const uint32_t* n = nodes + 3ull * head;
            color[k] = n[0];
            depth[k] = __uint_as_float(n[1]);
            head = n[2];
        } else {
            head = kPpllNull;
        }
    }
    // tail: swap closer nodes into the k-buffer, blend the evicted ones.  The
    // furthest slot (resolve.comp:48-57) changes only when a node is swapped
    // in, so it is rescanned only then.  The walk runs one node ahead: the
    // load of node f+1 is issued before node f is swapped and blended, so
    // the list's latency and the slot arithmetic overlap.
    float far_depth;
    int far_idx = furthest(depth, far_depth);
    bool have = head != kPpllNull && head < node_count;
    uint3 node = make_uint3(0u, 0u, kPpllNull);
    if (have) node = load_node(nodes, head);
    for (int f = 0; f < kMaxFragments && have; ++f) {
        const uint32_t next = node.z;
        const bool have_next = f + 1 < kMaxFragments && next != kPpllNull && next < node_count;
        // the current node's address stands in when there is no next one
        const uint3 ahead = load_node(nodes, have_next ? next : head);
        uint32_t fc = node.x;
        float fd = __uint_as_float(node.y);
        if (fd < far_depth) {
#pragma unroll
            for (int k = 0; k < kK; ++k) {
                if (k == far_idx) {
                    uint32_t tc = color[k];
                    float td = depth[k];
                    color[k] = fc;
                    depth[k] = fd;
                    fc = tc;
                    fd = td;
                }
            }
            far_idx = furthest(depth, far_depth);
        }
        resolved = blend(resolved, unpack_unorm4x8(fc));
        node = ahead;
        head = next;
        have = have_next;
    }
    // The k-buffer, back to front (resolve.comp:81-99).  Each of the K rounds
    // there blends the slot of greatest depth above 0 -- the lowest slot
    // among equal depths, since the scan's running maximum starts at 0 and
    // takes only strictly greater depths -- and zeroes its depth; once no
    // depth above 0 is left, a round finds slot 0 and blends slot 0's colour
    // again.  The same sequence comes from one descending sort of the keys
    // (depth bits << 4 | K-1-slot) for depths above 0 (positive float bits
    // order as the floats; NaN is not above 0) and 0 for the others: a
    // bitonic network over registers, 80 compare-exchanges instead of K
    // scans of K slots.
    const uint32_t color0 = color[0];
    unsigned long long key[kK];
#pragma unroll
    for (int k = 0; k < kK; ++k)
        key[k] = depth[k] > 0.0f
                     ? (static_cast<unsigned long long>(__float_as_uint(depth[k])) << 4) | unsigned(kK - 1 - k)
                     : 0ull;
#pragma unroll
    for (int size = 2; size <= kK; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < kK; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    // runs of `size` alternate direction; the last run (all K) is descending
                    const bool out_of_order = (i & size) == 0 ? key[i] < key[j] : key[i] > key[j];
                    const unsigned long long ki = key[i], kj = key[j];
                    const uint32_t ci = color[i], cj = color[j];
                    key[i] = out_of_order ? kj : ki;
                    key[j] = out_of_order ? ki : kj;
                    color[i] = out_of_order ? cj : ci;
                    color[j] = out_of_order ? ci : cj;
                }
            }
#pragma unroll
    for (int r = 0; r < kK; ++r) resolved = blend(resolved, unpack_unorm4x8(key[r] ? color[r] : color0));
    out[pix] = resolved;
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_ppll_clear(const vkhr_ppll* ppll, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(ppll && ppll->heads && ppll->counter, VKHR_ERR_INVALID, "ppll is incomplete");
    VKHR_REQUIRE(ppll->width && ppll->height, VKHR_ERR_INVALID, "empty ppll");
    size_t n = size_t(ppll->width) * ppll->height;
    ppll_clear_kernel<<<grid_for(n, 256, 8192), 256, 0, as_stream(stream)>>>(ppll->heads, n, ppll->counter);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_ppll_resolve(const vkhr_ppll* ppll, const float* background, float* out, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(ppll && ppll->heads && ppll->nodes && out, VKHR_ERR_INVALID, "null argument");
    VKHR_REQUIRE(ppll->width && ppll->height, VKHR_ERR_INVALID, "empty ppll");
    dim3 grid((ppll->width + 15) / 16, (ppll->height + 15) / 16);
    ppll_resolve_kernel<<<grid, 256, kResolveLdsReserve, as_stream(stream)>>>(
        ppll->heads, ppll->nodes, ppll->node_count, int(ppll->width), int(ppll->height),
        reinterpret_cast<const float4*>(background), reinterpret_cast<float4*>(out));
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
