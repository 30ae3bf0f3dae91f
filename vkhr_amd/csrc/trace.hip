// Strand ray tracer on the GPU: the reference's Embree path tracer of flat
// linear curves (src/vkhr/ray_tracer.cc:76-169, ray_tracer/hair_style.cc:
// 18-44,50-110) as a ground-truth for the raymarcher's shadow and AO
// approximations.
//
// Acceleration structure: a linear BVH over the segments (Karras, "Maximizing
// Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG
// 2012): 30-bit Morton codes of the segment centroids, sorted with rocPRIM's
// radix sort, a binary radix tree over (code, index) -- which bounds its depth
// by the 62 key bits -- and a bottom-up refit with one arrival counter per
// node.  Each internal node holds its two children's boxes (64 bytes, four
// 16-byte loads).
//
// Traversal is stackless (bit-trail with parent links, after Hapala et al.,
// "Efficient Stack-less BVH Traversal for Ray Tracing", SCCG 2011, and Afra
// and Szirmay-Kalos, "Stackless Multi-BVH Traversal", 2014): a 64-bit word
// records, per level of the current path, whether the far sibling is still to
// be visited, so no per-lane stack lives in LDS or scratch.
//
// Primitive: the ray-facing ribbon of a segment whose radius varies linearly
// (Embree's flat linear curve): the axis is projected on the plane through
// the ray origin perpendicular to the ray; u minimises the projected
// distance to the ray, and the hit needs that distance within the radius at
// u and the axis depth t in [tnear, tfar].  Compiled with -ffp-contract=off:
// the float32 expressions are evaluated in the order the CPU restatement
// evaluates them, so hits agree bit for bit with it.
#include <rocprim/device/device_radix_sort.hpp>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace vkhr {
namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kNone = 0xffffffffu;

struct TraceBvh {
    const float4* nodes;       // 4 per internal node: lo_l|cl, hi_l|cr, lo_r, hi_r
    const uint32_t* parent;    // internal nodes [0, n-1), then leaves [n-1, 2n-1)
    const float4* prim_a;      // p0, r0 (leaf order)
    const float4* prim_b;      // p1, r1
    const uint32_t* prim_seg;  // segment index (index pair) of the leaf
    uint32_t n;                // leaves
};

struct Hit {
    float t, u;
    uint32_t seg, leaf;
};

__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ bool flat_curve(float3 o, float3 d, float dd, float4 a, float4 b, float tnear,
                                           float tfar, float* t_out, float* u_out) {
    const float3 q0 = make_float3(a.x - o.x, a.y - o.y, a.z - o.z);
    const float3 q1 = make_float3(b.x - o.x, b.y - o.y, b.z - o.z);
    const float z0 = dot3(q0, d) / dd, z1 = dot3(q1, d) / dd;
    const float3 x0 = make_float3(q0.x - z0 * d.x, q0.y - z0 * d.y, q0.z - z0 * d.z);
    const float3 v = make_float3((q1.x - z1 * d.x) - x0.x, (q1.y - z1 * d.y) - x0.y, (q1.z - z1 * d.z) - x0.z);
    const float vv = dot3(v, v);
    float u = 0.0f;
    if (vv > 0.0f) {
        u = -dot3(x0, v) / vv;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    }
    const float3 p = make_float3(x0.x + u * v.x, x0.y + u * v.y, x0.z + u * v.z);
    const float z = z0 + u * (z1 - z0);
    const float r = a.w + u * (b.w - a.w);
    if (!(dot3(p, p) <= r * r)) return false;
    if (!(z >= tnear && z <= tfar)) return false;
    *t_out = z;
    *u_out = u;
    return true;
}

// Directions with a zero component get a tiny one of the same sign, so the
// slab reciprocals stay finite (boxes are padded, so this never loses a hit).
__device__ __forceinline__ float3 safe_inverse(float3 d) {
    auto f = [](float x) { return 1.0f / (fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x); };
    return make_float3(f(d.x), f(d.y), f(d.z));
}

__device__ __forceinline__ bool slab(float4 lo, float4 hi, float3 o, float3 inv, float tnear, float tfar,
                                     float* t_enter) {
    const float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    const float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    const float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    const float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tnear));
    const float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tfar));
    *t_enter = t0;
    return t0 <= t1;
}

__device__ __forceinline__ uint32_t parent_of(const TraceBvh& b, uint32_t x) {
    return (x & kLeaf) ? b.parent[b.n - 1u + (x & ~kLeaf)] : b.parent[x];
}

// Closest (kAny = false) or any (kAny = true) hit of the ray in [tnear, tfar];
// closest ties go to the lowest segment index.
// A short stack of the deepest pending siblings of one lane, in LDS (entry
// k of lane x at node[k * stride + x], t[k * stride + x]): the far child
// and its entry distance, pushed when both children are hit.  The trail
// bits stay the record of every pending level; the stack caches the deepest
// ones, most recent on top, so a pop visits exactly the sibling the parent
// walk would find, without the walk and the re-read of the parent's node.
// When it is full the oldest (shallowest) entry is dropped, and that level
// is found by the parent walk once the stack is empty.
constexpr uint32_t kTraceStack = 16;
struct TraceStack {
    uint32_t* node;
    float* t;
    uint32_t stride;
};

template <bool kAny>
__device__ Hit trace(const TraceBvh& b, float3 o, float3 d, float tnear, float tfar, TraceStack stk,
                     uint32_t skip = kNone) {
    Hit best{tfar, 0.0f, kNone, kNone};
    if (b.n == 0u) return best;
    uint32_t top = 0u, count = 0u;  // ring buffer: top entry index, entries held
    const float dd = dot3(d, d);
    const float3 inv = safe_inverse(d);
    auto test_leaf = [&](uint32_t leaf) -> bool {
        float t, u;
        const uint32_t seg = b.prim_seg[leaf];
        if (seg == skip) return false;
        if (!flat_curve(o, d, dd, b.prim_a[leaf], b.prim_b[leaf], tnear, best.t, &t, &u)) return false;
        if (best.seg == kNone || t < best.t || (t == best.t && seg < best.seg)) best = Hit{t, u, seg, leaf};
        return true;
    };
    if (b.n == 1u) {
        test_leaf(0u);
        return best;
    }
    uint32_t cur = 0u;          // the node at the bottom of the path (root = internal node 0)
    uint64_t trail = 0ull;      // bit k: the sibling at level depth-k is still to be visited
    bool expand = true;         // cur is an internal node to open
    // A traversal opens each of the 2n-1 nodes at most once (and backs up
    // through each at most once): the cap only guards against a malformed tree.
    for (uint64_t steps = 4ull * b.n + 256ull; steps != 0ull; --steps) {
        if (expand) {
            const float4* nd = b.nodes + 4ull * cur;
            const float4 l0 = nd[0], l1 = nd[1], r0 = nd[2], r1 = nd[3];
            const uint32_t cl = __float_as_uint(l0.w), cr = __float_as_uint(l1.w);
            float tl, tr;
            const bool hl = slab(l0, l1, o, inv, tnear, best.t, &tl);
            const bool hr = slab(r0, r1, o, inv, tnear, best.t, &tr);
            uint32_t next = kNone;
            if (hl && hr) {
                next = tl <= tr ? cl : cr;
                trail = (trail << 1) | 1ull;
                top = count ? (top + 1u) % kTraceStack : top;
                count = count < kTraceStack ? count + 1u : count;
                stk.node[top * stk.stride] = tl <= tr ? cr : cl;
                stk.t[top * stk.stride] = tl <= tr ? tr : tl;
            } else if (hl || hr) {
                next = hl ? cl : cr;
                trail <<= 1;
            }
            if (next != kNone) {
                cur = next;
                if (!(next & kLeaf)) continue;  // open it
                if (test_leaf(next & ~kLeaf) && kAny) return best;
            }
        }
        if (trail == 0ull) return best;
        if (count) {
            // the deepest pending level is the lowest set trail bit; its sibling is the stack top
            trail >>= __builtin_ctzll(trail);
            trail ^= 1ull;
            const uint32_t sib = stk.node[top * stk.stride];
            const float ts = stk.t[top * stk.stride];
            top = (top + kTraceStack - 1u) % kTraceStack;
            --count;
            cur = sib;
            expand = false;
            if (!(ts <= best.t)) continue;  // culled by the closer hit found meanwhile
            if (sib & kLeaf) {
                if (test_leaf(sib & ~kLeaf) && kAny) return best;
                continue;
            }
            expand = true;
            continue;
        }
        // back up to the deepest level whose sibling is pending
        while ((trail & 1ull) == 0ull) {
            if (trail == 0ull) return best;
            cur = parent_of(b, cur);
            trail >>= 1;
        }
        trail ^= 1ull;
        const uint32_t p = parent_of(b, cur);
        const float4* nd = b.nodes + 4ull * p;
        const float4 l0 = nd[0], l1 = nd[1];
        const bool left = __float_as_uint(l0.w) == cur;  // cur is the left child: the sibling is the right one
        const uint32_t sib = left ? __float_as_uint(l1.w) : __float_as_uint(l0.w);
        const float4 lo = left ? nd[2] : l0, hi = left ? nd[3] : l1;
        float ts;
        cur = sib;
        expand = false;
        if (!slab(lo, hi, o, inv, tnear, best.t, &ts)) continue;  // culled by the closer hit found meanwhile
        if (sib & kLeaf) {
            if (test_leaf(sib & ~kLeaf) && kAny) return best;
            continue;
        }
        expand = true;
    }
    return best;
}

// ------------------------------------------------------------------ build

struct BuildArgs {
    const float* xyz;
    const float* radius;
    const uint32_t* indices;
    uint32_t n_vertices;
    uint32_t n;                 // segments
    float4* box_lo;             // per segment (input order), w unused
    float4* box_hi;
    float* bounds_partial;      // 6 per block: centroid min xyz, max xyz
    float* bounds;              // 6: centroid min, max; then scene min, max (12)
    uint32_t* keys;
    uint32_t* values;
};

__global__ void __launch_bounds__(256) segment_boxes_kernel(BuildArgs a) {
    __shared__ float red[6][256];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (k < a.n) {
        const uint32_t ia = a.indices[2ull * k], ib = a.indices[2ull * k + 1];
        float4 lo = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.0f), hi = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.0f);
        if (ia < a.n_vertices && ib < a.n_vertices) {
            const float3 p = load_xyz(a.xyz, ia), q = load_xyz(a.xyz, ib);
            const float r = fmaxf(a.radius[ia], a.radius[ib]);
            lo = make_float4(fminf(p.x, q.x) - r, fminf(p.y, q.y) - r, fminf(p.z, q.z) - r, 0.0f);
            hi = make_float4(fmaxf(p.x, q.x) + r, fmaxf(p.y, q.y) + r, fmaxf(p.z, q.z) + r, 0.0f);
            const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
            for (int i = 0; i < 3; ++i) cmin[i] = cmax[i] = c[i];
        }
        a.box_lo[k] = lo;
        a.box_hi[k] = hi;
    }
    for (int i = 0; i < 3; ++i) {
        red[i][threadIdx.x] = cmin[i];
        red[3 + i][threadIdx.x] = cmax[i];
    }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int i = 0; i < 3; ++i) {
                red[i][threadIdx.x] = fminf(red[i][threadIdx.x], red[i][threadIdx.x + s]);
                red[3 + i][threadIdx.x] = fmaxf(red[3 + i][threadIdx.x], red[3 + i][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) a.bounds_partial[6ull * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) bounds_final_kernel(BuildArgs a, uint32_t blocks) {
    __shared__ float red[6][256];
    float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = threadIdx.x; i < blocks; i += blockDim.x)
        for (int c = 0; c < 6; ++c)
            v[c] = c < 3 ? fminf(v[c], a.bounds_partial[6ull * i + c]) : fmaxf(v[c], a.bounds_partial[6ull * i + c]);
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 6; ++c)
                red[c][threadIdx.x] = c < 3 ? fminf(red[c][threadIdx.x], red[c][threadIdx.x + s])
                                            : fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) a.bounds[threadIdx.x] = red[threadIdx.x][0];
}

// 10 bits of v spread to every third bit.
__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void __launch_bounds__(256) morton_kernel(BuildArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const float4 lo = a.box_lo[k], hi = a.box_hi[k];
    uint32_t code = 0u;
    if (lo.x <= hi.x) {
        const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
        uint32_t q[3];
        for (int i = 0; i < 3; ++i) {
            const float ext = a.bounds[3 + i] - a.bounds[i];
            float f = ext > 0.0f ? (c[i] - a.bounds[i]) / ext : 0.0f;
            f = fminf(fmaxf(f * 1024.0f, 0.0f), 1023.0f);
            q[i] = uint32_t(f);
        }
        code = (spread_bits(q[0]) << 2) | (spread_bits(q[1]) << 1) | spread_bits(q[2]);
    }
    a.keys[k] = code;
    a.values[k] = k;
}

// Leaves in code order: the padded boxes and the primitive records.
__global__ void __launch_bounds__(256)
leaves_kernel(BuildArgs a, const uint32_t* __restrict__ order, float pad, float4* leaf_lo, float4* leaf_hi,
              float4* prim_a, float4* prim_b, uint32_t* prim_seg) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.n) return;
    const uint32_t k = order[l];
    const uint32_t ia = a.indices[2ull * k], ib = a.indices[2ull * k + 1];
    float4 pa = make_float4(0.0f, 0.0f, 0.0f, NAN), pb = make_float4(0.0f, 0.0f, 0.0f, NAN);  // never hit
    if (ia < a.n_vertices && ib < a.n_vertices) {
        const float3 p = load_xyz(a.xyz, ia), q = load_xyz(a.xyz, ib);
        pa = make_float4(p.x, p.y, p.z, a.radius[ia]);
        pb = make_float4(q.x, q.y, q.z, a.radius[ib]);
    }
    prim_a[l] = pa;
    prim_b[l] = pb;
    prim_seg[l] = k;
    const float4 lo = a.box_lo[k], hi = a.box_hi[k];
    leaf_lo[l] = make_float4(lo.x - pad, lo.y - pad, lo.z - pad, 0.0f);
    leaf_hi[l] = make_float4(hi.x + pad, hi.y + pad, hi.z + pad, 0.0f);
}

// Length of the common prefix of the (code, index) keys of leaves i and j;
// -1 outside [0, n).
__device__ __forceinline__ int prefix(const uint32_t* __restrict__ codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

// Internal node i of the binary radix tree over the sorted keys: its range
// [min(i, j), max(i, j)] and split, its children and their parent links.
__global__ void __launch_bounds__(256)
radix_tree_kernel(const uint32_t* __restrict__ codes, int n, float4* nodes, uint32_t* parent) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = prefix(codes, n, i, i + 1) > prefix(codes, n, i, i - 1) ? 1 : -1;
    const int dmin = prefix(codes, n, i, i - d);
    int lmax = 2;
    while (prefix(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (prefix(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix(codes, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefix(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const uint32_t left = min(i, j) == gamma ? (kLeaf | uint32_t(gamma)) : uint32_t(gamma);
    const uint32_t right = max(i, j) == gamma + 1 ? (kLeaf | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
    nodes[4ull * i + 0].w = __uint_as_float(left);
    nodes[4ull * i + 1].w = __uint_as_float(right);
    parent[(left & kLeaf) ? uint32_t(n - 1) + (left & ~kLeaf) : left] = uint32_t(i);
    parent[(right & kLeaf) ? uint32_t(n - 1) + (right & ~kLeaf) : right] = uint32_t(i);
    if (i == 0) parent[0] = kNone;
}

// Bottom-up boxes: each leaf climbs; at every node the second arrival
// (atomic counter) takes the union of the two child boxes upwards.  The
// child boxes pass between threads through device-scope atomic stores and
// loads (coherent across the XCDs' L2s without flushing them); the store is
// complete before the arrival is counted.  A device-wide fence instead
// writes back and invalidates the whole L2 at each step: 11.6 ms for 1.6 M
// leaves.
__device__ __forceinline__ void store_box(float4* slo, float4* shi, float4 lo, float4 hi) {
    float* a = reinterpret_cast<float*>(slo);
    float* b = reinterpret_cast<float*>(shi);
    __hip_atomic_store(a + 0, lo.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a + 1, lo.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a + 2, lo.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(b + 0, hi.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(b + 1, hi.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(b + 2, hi.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float3 load3(const float4* p) {
    const float* a = reinterpret_cast<const float*>(p);
    return make_float3(__hip_atomic_load(a + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(a + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __hip_atomic_load(a + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__global__ void __launch_bounds__(256)
refit_kernel(int n, const float4* __restrict__ leaf_lo, const float4* __restrict__ leaf_hi, float4* nodes,
             const uint32_t* __restrict__ parent, uint32_t* arrivals) {
    const int l = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (l >= n) return;
    float4 lo = leaf_lo[l], hi = leaf_hi[l];
    uint32_t child = kLeaf | uint32_t(l);
    uint32_t p = parent[uint32_t(n - 1) + uint32_t(l)];
    for (;;) {
        float4* nd = nodes + 4ull * p;
        const bool is_left = __float_as_uint(nd[0].w) == child;  // written by the tree kernel
        store_box(is_left ? &nd[0] : &nd[2], is_left ? &nd[1] : &nd[3], lo, hi);
        __builtin_amdgcn_s_waitcnt(0);  // the box stores are complete before the arrival
        if (__hip_atomic_fetch_add(&arrivals[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            return;  // the sibling's box is not in yet
        const float3 a0 = load3(&nd[0]), a1 = load3(&nd[1]), b0 = load3(&nd[2]), b1 = load3(&nd[3]);
        lo = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
        hi = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
        if (p == 0u) return;
        child = p;
        p = parent[p];
    }
}

// ---------------------------------------------------------------- tracing

__global__ void __launch_bounds__(256)
rays_kernel(TraceBvh b, const float* __restrict__ rays, uint32_t n, int any, float* t_out, float* u_out,
            uint32_t* seg_out) {
    __shared__ uint32_t stk_node[kTraceStack * 256];
    __shared__ float stk_t[kTraceStack * 256];
    const TraceStack stk{stk_node + threadIdx.x, stk_t + threadIdx.x, blockDim.x};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8ull * i;
    const float3 o = make_float3(r[0], r[1], r[2]), d = make_float3(r[4], r[5], r[6]);
    const Hit h = any ? trace<true>(b, o, d, r[3], r[7], stk) : trace<false>(b, o, d, r[3], r[7], stk);
    const bool hit = h.seg != kNone;
    if (t_out) t_out[i] = hit ? h.t : 0.0f;
    if (u_out) u_out[i] = hit ? h.u : 0.0f;
    seg_out[i] = hit ? h.seg : kNone;
}

struct FrameArgs {
    TraceBvh b;
    const float* tangents;
    const uint32_t* indices;
    float plane[12];
    float light[6];
    float diffuse[3];
    float exponent, ao_radius;
    int method, shadows_on, ignore_self;
    uint32_t seed, frame, width, height;
    double* accum;
};

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float sample(uint32_t seed, uint32_t pixel, uint32_t frame, uint32_t k, float lo,
                                        float hi) {
    const uint32_t r = mix32(mix32(mix32(seed ^ 0x9e3779b9u) ^ pixel) ^ (frame * 8u + k));
    return (float(r) / 4294967296.0f) * (hi - lo) + lo;
}

__device__ __forceinline__ float3 normalized(float3 v) {
    const float l = sqrtf(dot3(v, v));
    return make_float3(v.x / l, v.y / l, v.z / l);
}

constexpr float kEpsilon = 0.000001f;  // Ray::Epsilon (ray_tracer/ray.hh:15)

// One sample of Raytracer::draw per pixel; 16x16 pixels per block, 8x8 per wave.
__global__ void __launch_bounds__(256) frame_kernel(FrameArgs a) {
    __shared__ uint32_t stk_node[kTraceStack * 256];
    __shared__ float stk_t[kTraceStack * 256];
    const TraceStack stk{stk_node + threadIdx.x, stk_t + threadIdx.x, blockDim.x};
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 16u + (lane & 7u) + 8u * (wave & 1u);
    const uint32_t j = blockIdx.y * 16u + (lane >> 3) + 8u * (wave >> 1);
    if (i >= a.width || j >= a.height) return;
    const uint32_t pix = i + j * a.width;
    const float3 X = make_float3(a.plane[0], a.plane[1], a.plane[2]);
    const float3 Y = make_float3(a.plane[3], a.plane[4], a.plane[5]);
    const float3 Z = make_float3(a.plane[6], a.plane[7], a.plane[8]);
    const float3 eye = make_float3(a.plane[9], a.plane[10], a.plane[11]);
    const float fx = float(i) + sample(a.seed, pix, a.frame, 0, 0.0f, 1.0f);
    const float fy = float(j) + sample(a.seed, pix, a.frame, 1, 0.0f, 1.0f);
    const float3 d = make_float3((fx * X.x + fy * Y.x) + Z.x, (fx * X.y + fy * Y.y) + Z.y, (fx * X.z + fy * Y.z) + Z.z);
    float color[3] = {1.0f, 1.0f, 1.0f};
    const Hit h = trace<false>(a.b, eye, d, 0.0f, INFINITY, stk);
    if (h.seg != kNone) {
        const float3 P = make_float3(eye.x + d.x * h.t, eye.y + d.y * h.t, eye.z + d.z * h.t);
        const uint32_t skip = a.ignore_self ? h.seg : kNone;  // secondary rays may skip their own segment
        float L[3] = {1.0f, 1.0f, 1.0f};
        if (a.method != 3) {
            // the shadow ray's direction is the jittered light position itself
            const float3 sd = make_float3(a.light[0] + sample(a.seed, pix, a.frame, 2, -16.0f, 16.0f),
                                          a.light[1] + sample(a.seed, pix, a.frame, 3, -16.0f, 16.0f),
                                          a.light[2] + sample(a.seed, pix, a.frame, 4, -16.0f, 16.0f));
            const bool occluded = trace<true>(a.b, P, sd, kEpsilon, INFINITY, stk, skip).seg != kNone;
            if (!occluded || !a.shadows_on) {
                if (a.method == 0) {
                    const uint32_t ia = a.indices[2ull * h.seg], ib = a.indices[2ull * h.seg + 1];
                    const float3 t0 = load_xyz(a.tangents, ia), t1 = load_xyz(a.tangents, ib);
                    const float3 T = make_float3(t0.x + h.u * (t1.x - t0.x), t0.y + h.u * (t1.y - t0.y),
                                                 t0.z + h.u * (t1.z - t0.z));
                    const float3 l = normalized(make_float3(a.light[0] - P.x, a.light[1] - P.y, a.light[2] - P.z));
                    const float3 e = normalized(make_float3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
                    // embree::HairStyle::kajiya_kay (ray_tracer/hair_style.cc:86-107)
                    const float ctl = dot3(l, T), cte = dot3(e, T);
                    const float sa = 1.0f - ctl * ctl, sb = 1.0f - cte * cte;
                    const float stl = sqrtf(sa > 0.0f ? sa : 0.0f), ste = sqrtf(sb > 0.0f ? sb : 0.0f);
                    const float base = ctl * cte + stl * ste;
                    float spec = base > 0.0f ? powf(base, a.exponent) : 0.0f;
                    spec = spec < 0.0f ? 0.0f : (spec > 1.0f ? 1.0f : spec);
                    for (int c = 0; c < 3; ++c) L[c] = a.diffuse[c] * stl + a.light[3 + c] * spec;
                }
            } else {
                L[0] = L[1] = L[2] = 0.0f;
            }
        }
        for (int c = 0; c < 3; ++c) color[c] = L[c];
        if (a.method != 2) {
            const float3 rd = make_float3(sample(a.seed, pix, a.frame, 5, -1.0f, 1.0f),
                                          sample(a.seed, pix, a.frame, 6, -1.0f, 1.0f),
                                          sample(a.seed, pix, a.frame, 7, -1.0f, 1.0f));
            const float len = sqrtf(dot3(rd, rd));
            const float t_max = len > 0.0f ? (a.ao_radius / len) * 1.001f + kEpsilon : INFINITY;
            const Hit o = trace<false>(a.b, P, rd, kEpsilon, t_max, stk, skip);
            float ao = 2.0f;
            if (o.seg != kNone) {
                const float3 q = make_float3((P.x + rd.x * o.t) - P.x, (P.y + rd.y * o.t) - P.y,
                                             (P.z + rd.z * o.t) - P.z);
                if (!(sqrtf(dot3(q, q)) > a.ao_radius)) ao = 0.0f;
            }
            for (int c = 0; c < 3; ++c) color[c] *= ao;
        }
    }
    double* out = a.accum + 3ull * pix;
    for (int c = 0; c < 3; ++c) out[c] += double(color[c]);
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

struct vkhr_bvh {
    int device = 0;
    uint32_t n = 0;
    void* memory = nullptr;     // one allocation holding every array below
    float4* nodes = nullptr;
    uint32_t* parent = nullptr;
    float4* prim_a = nullptr;
    float4* prim_b = nullptr;
    uint32_t* prim_seg = nullptr;
    float bounds[6] = {0, 0, 0, 0, 0, 0};

    TraceBvh view() const { return TraceBvh{nodes, parent, prim_a, prim_b, prim_seg, n}; }
};

namespace {

template <class T>
T* carve(char*& p, size_t count) {
    T* out = reinterpret_cast<T*>(p);
    p += (count * sizeof(T) + 255) & ~size_t(255);
    return out;
}

size_t carve_size(size_t bytes) { return (bytes + 255) & ~size_t(255); }

}  // namespace

extern "C" int vkhr_bvh_build(const float* xyz, const float* radius, uint32_t n_vertices, const uint32_t* indices,
                              uint32_t n_indices, vkhr_stream stream, vkhr_bvh** out) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(out, VKHR_ERR_INVALID, "out is null");
    *out = nullptr;
    const uint32_t n = n_indices / 2;
    // (an argument passed in the wrong position usually shows up as a small integer here)
    const auto plausible = [](const void* p) { return reinterpret_cast<uintptr_t>(p) >= 65536u; };
    VKHR_REQUIRE(n == 0 || (plausible(xyz) && plausible(radius) && plausible(indices)), VKHR_ERR_INVALID,
                 "null or implausible geometry pointer");
    VKHR_REQUIRE(n < (1u << 30), VKHR_ERR_OVERFLOW, "too many segments");  // int range search in the tree build
    hipStream_t s = as_stream(stream);
    int device = 0;
    VKHR_CHECK_HIP(hipGetDevice(&device));
    vkhr_bvh* b = new vkhr_bvh();
    b->device = device;
    b->n = n;
    if (n == 0) {
        *out = b;
        return VKHR_OK;
    }
    const size_t inner = n > 1 ? n - 1 : 0;
    const size_t keep = carve_size(inner * 4 * sizeof(float4)) + carve_size((inner + n) * sizeof(uint32_t)) +
                        2 * carve_size(n * sizeof(float4)) + carve_size(n * sizeof(uint32_t));
    const uint32_t blocks = grid_for(n, 256);
    // temporaries: boxes (input order), leaf boxes, bounds, keys/values x2, arrivals, sort scratch
    size_t sort_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), size_t(n), 0, 30, s) != hipSuccess) {
        delete b;
        set_error("rocprim::radix_sort_pairs size query failed");
        return VKHR_ERR_HIP;
    }
    const size_t temp = 4 * carve_size(n * sizeof(float4)) + carve_size(size_t(blocks) * 6 * sizeof(float)) +
                        carve_size(12 * sizeof(float)) + 4 * carve_size(n * sizeof(uint32_t)) +
                        carve_size(inner * sizeof(uint32_t)) + carve_size(sort_bytes);
    hipError_t e = hipMalloc(&b->memory, keep);
    void* tmp = nullptr;
    if (e == hipSuccess) e = hipMalloc(&tmp, temp);
    if (e != hipSuccess) {
        if (b->memory) (void)hipFree(b->memory);
        delete b;
        set_error("bvh hipMalloc failed: %s", hipGetErrorString(e));
        return VKHR_ERR_NOMEM;
    }
    char* p = static_cast<char*>(b->memory);
    b->nodes = carve<float4>(p, inner * 4);
    b->parent = carve<uint32_t>(p, inner + n);
    b->prim_a = carve<float4>(p, n);
    b->prim_b = carve<float4>(p, n);
    b->prim_seg = carve<uint32_t>(p, n);
    char* q = static_cast<char*>(tmp);
    BuildArgs a;
    a.xyz = xyz;
    a.radius = radius;
    a.indices = indices;
    a.n_vertices = n_vertices;
    a.n = n;
    a.box_lo = carve<float4>(q, n);
    a.box_hi = carve<float4>(q, n);
    float4* leaf_lo = carve<float4>(q, n);
    float4* leaf_hi = carve<float4>(q, n);
    a.bounds_partial = carve<float>(q, size_t(blocks) * 6);
    a.bounds = carve<float>(q, 12);
    a.keys = carve<uint32_t>(q, n);
    a.values = carve<uint32_t>(q, n);
    uint32_t* keys_sorted = carve<uint32_t>(q, n);
    uint32_t* order = carve<uint32_t>(q, n);
    uint32_t* arrivals = carve<uint32_t>(q, inner);
    void* sort_tmp = carve<char>(q, sort_bytes);
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(tmp);
        (void)hipFree(b->memory);
        delete b;
        return code;
    };
    segment_boxes_kernel<<<blocks, 256, 0, s>>>(a);
    bounds_final_kernel<<<1, 256, 0, s>>>(a, blocks);
    // the scene box (padded segment boxes) for the traversal pad: read back
    float host_bounds[6];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host_bounds, a.bounds, sizeof(host_bounds), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error("bvh bounds pass failed");
        return fail(VKHR_ERR_HIP);
    }
    std::memcpy(b->bounds, host_bounds, sizeof(host_bounds));
    float diag2 = 0.0f, mag = 0.0f;
    for (int c = 0; c < 3; ++c) {
        const float ext = host_bounds[3 + c] - host_bounds[c];
        diag2 += ext * ext;
        mag = std::max(mag, std::max(std::fabs(host_bounds[c]), std::fabs(host_bounds[3 + c])));
    }
    // Boxes are widened by far more than the float rounding of the slab test
    // and of the curve test, so that the box of every hit contains it.
    const float pad = 1e-4f * std::sqrt(diag2) + 1e-6f * mag + 1e-30f;
    morton_kernel<<<blocks, 256, 0, s>>>(a);
    if (hipGetLastError() != hipSuccess ||
        rocprim::radix_sort_pairs(sort_tmp, sort_bytes, a.keys, keys_sorted, a.values, order, size_t(n), 0, 30, s) !=
            hipSuccess) {
        set_error("bvh sort failed");
        return fail(VKHR_ERR_HIP);
    }
    leaves_kernel<<<blocks, 256, 0, s>>>(a, order, pad, leaf_lo, leaf_hi, b->prim_a, b->prim_b, b->prim_seg);
    if (n > 1) {
        if ((rc = fill_u32(reinterpret_cast<uint32_t*>(b->nodes), inner * 16, 0u, s)) ||
            (rc = fill_u32(arrivals, inner, 0u, s)))
            return fail(rc);
        radix_tree_kernel<<<grid_for(inner, 256), 256, 0, s>>>(keys_sorted, int(n), b->nodes, b->parent);
        refit_kernel<<<blocks, 256, 0, s>>>(int(n), leaf_lo, leaf_hi, b->nodes, b->parent, arrivals);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        set_error("bvh build kernels failed");
        return fail(VKHR_ERR_HIP);
    }
    (void)hipFree(tmp);
    *out = b;
    return VKHR_OK;
}

extern "C" int vkhr_bvh_destroy(vkhr_bvh* b) {
    clear_error();
    if (!b) return VKHR_OK;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    if (b->memory) (void)hipFree(b->memory);
    (void)hipSetDevice(prev);
    delete b;
    return VKHR_OK;
}

extern "C" int vkhr_bvh_info(const vkhr_bvh* b, uint32_t* n_segments, float* bounds) {
    clear_error();
    VKHR_REQUIRE(b, VKHR_ERR_INVALID, "bvh is null");
    if (n_segments) *n_segments = b->n;
    if (bounds) std::memcpy(bounds, b->bounds, sizeof(b->bounds));
    return VKHR_OK;
}

extern "C" int vkhr_bvh_nodes(const vkhr_bvh* b, float* nodes, uint32_t* parent, uint32_t* leaf_segment,
                              vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(b, VKHR_ERR_INVALID, "bvh is null");
    hipStream_t s = as_stream(stream);
    const size_t inner = b->n > 1 ? b->n - 1 : 0;
    if (nodes && inner) VKHR_CHECK_HIP(hipMemcpyAsync(nodes, b->nodes, inner * 4 * sizeof(float4), hipMemcpyDefault, s));
    if (parent && b->n)
        VKHR_CHECK_HIP(hipMemcpyAsync(parent, b->parent, (inner + b->n) * sizeof(uint32_t), hipMemcpyDefault, s));
    if (leaf_segment && b->n)
        VKHR_CHECK_HIP(hipMemcpyAsync(leaf_segment, b->prim_seg, b->n * sizeof(uint32_t), hipMemcpyDefault, s));
    return VKHR_OK;
}

extern "C" int vkhr_trace_rays(const vkhr_bvh* b, const float* rays, uint32_t n_rays, int any_hit, float* t_out,
                               float* u_out, uint32_t* segment_out, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(b && (n_rays == 0 || (rays && segment_out)), VKHR_ERR_INVALID, "null argument");
    if (n_rays == 0) return VKHR_OK;
    rays_kernel<<<grid_for(n_rays, 256), 256, 0, as_stream(stream)>>>(b->view(), rays, n_rays, any_hit, t_out, u_out,
                                                                      segment_out);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_raytrace(const vkhr_bvh* b, const float* tangents, const uint32_t* indices,
                             const vkhr_viewing_plane* plane, const vkhr_light* light,
                             const vkhr_trace_params* params, uint32_t width, uint32_t height, double* accum,
                             vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(b && plane && light && params && accum, VKHR_ERR_INVALID, "null argument");
    VKHR_REQUIRE(width && height && width <= 16384 && height <= 16384, VKHR_ERR_INVALID, "bad image size %ux%u",
                 width, height);
    VKHR_REQUIRE(params->method <= 3, VKHR_ERR_INVALID, "unknown method %u", params->method);
    VKHR_REQUIRE(b->n == 0 || indices, VKHR_ERR_INVALID, "indices are null");
    VKHR_REQUIRE(params->method != 0 || b->n == 0 || tangents, VKHR_ERR_INVALID, "shading needs tangents");
    FrameArgs a;
    std::memset(&a, 0, sizeof(a));
    a.b = b->view();
    a.tangents = tangents;
    a.indices = indices;
    for (int c = 0; c < 3; ++c) {
        a.plane[c] = plane->x[c];
        a.plane[3 + c] = plane->y[c];
        a.plane[6 + c] = plane->z[c];
        a.plane[9 + c] = plane->point[c];
        a.light[c] = light->origin[c];
        a.light[3 + c] = light->intensity[c];
        a.diffuse[c] = params->diffuse[c];
    }
    a.exponent = params->exponent;
    a.ao_radius = params->ao_radius;
    a.method = int(params->method);
    a.shadows_on = params->shadows_on;
    a.ignore_self = params->ignore_self ? 1 : 0;
    a.seed = params->seed;
    a.frame = params->frame;
    a.width = width;
    a.height = height;
    a.accum = accum;
    frame_kernel<<<dim3((width + 15) / 16, (height + 15) / 16), 256, 0, as_stream(stream)>>>(a);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
