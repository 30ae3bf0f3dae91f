// Strand rasterizer: the reference's line passes (shadow-map depth pass and
// strand colour pass into the PPLL) as gfx950 kernels.
//
// The reference draws LINE_LIST primitives with a dynamic line width and
// lets the Vulkan rasterizer pick the fragments.  Here one lane owns one
// index pair: it clips the segment to 0 <= z <= w, maps it to window
// coordinates, and walks the pixels of its major axis column by column;
// in each column only the rows of the line's strip (plus a rounding margin)
// are tested, with the exact rule
//     0 <= (c-a).(b-a) < |b-a|^2   and   cross(b-a, c-a)^2 <= (w/2)^2 |b-a|^2
// in float64 (declared in include/vkhr_hip.h).
//
// Colour pass: segments give from zero to dozens of fragments, so shading
// one fragment per lane per round would leave most lanes of a wave idle.
// Instead the lanes push their covered fragments into a per-wave LDS queue
// (ballot + rank), and the float64 shading runs only on full waves of 64
// queued fragments, each lane shading a fragment of whichever segment it
// came from (the segments' shading data sit in LDS).  Each block reserves
// its PPLL node ids with one counter atomic after a counting walk (a
// counter atomic per wave and round serialized on the one address), and
// links each node with an exchange on its pixel's head.
// Window depth interpolates linearly in t, varyings perspective-correctly.
// Compiled with -ffp-contract=off so the rule evaluates as the float64
// oracle evaluates it.
#include "render.hpp"

#include <cmath>
#include <cstring>

namespace vkhr {
namespace {

constexpr int kMaxPcfTaps = 81;  // kernel sizes up to 9x9

struct LineSetup {
    double ax, ay, az, bx, by, bz;   // window coordinates of the clipped ends
    double wa, wb;                   // clip w of the clipped ends
    double s0, s1;                   // clip interval along the original segment
    double vx, vy, l2;               // b - a, |b - a|^2
    double lim;                      // (w/2)^2 |b - a|^2
    double h;                        // half-height of the strip across the minor axis
    double slope;                    // d(minor) / d(major)
    double margin;                   // rounding margin of a column's row range (pixels)
    int col0, col_end;               // columns [col0, col_end) of the major axis
    bool xmajor, valid;
};

// Walk state of one lane: the current column and its remaining rows, and
// the pixel box the walk is clipped to (along the major / minor axis).
struct Walk {
    int col, row, row_end;
    int col_end, minor_lo, minor_hi;  // columns [.., col_end), rows [minor_lo, minor_hi]
};
struct RasterArgs {
    double vp[16];                   // view-projection (column-major)
    double lm[16];                   // light->matrix (biased light view-projection)
    double eye[3], light_origin[3], hair_color[3], intensity[3];
    double alpha, shininess, lod, line_width, res_x, res_y;
    double ln_transmit;              // log(1 - alpha)
    double pcf_dx[kMaxPcfTaps], pcf_dy[kMaxPcfTaps], pcf_w[kMaxPcfTaps];
    int pcf_taps;
    double ao_scaling, ao_clamp, ao_exponent;
    double vorigin[3], vsize[3], vres[3];
    DensityVolume vol;
    int shading_model, adsm_on;
    int width, height;               // target
    const float* shadow_map;
    int sw, sh;
    const float* depth_buffer;
    const float* xyz;
    const float* tangents;
    const float* thickness;
    const uint32_t* indices;
    uint32_t n_pairs, n_vertices;
    PpllView ppll;
    CaptureView cap;
    uint32_t* depth_bits;            // shadow pass output (float bits, min)
};

__device__ __forceinline__ void xform4(const double* m, double x, double y, double z, double o[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = m[0 * 4 + r] * x + m[1 * 4 + r] * y + m[2 * 4 + r] * z + m[3 * 4 + r];
}

// Clip, project and plan the candidate walk of one segment.
__device__ LineSetup setup_line(const double* m, const double pa[3], const double pb[3], int width, int height,
                                double line_width) {
    LineSetup L;
    L.valid = false;
    double ca[4], cb[4];
    xform4(m, pa[0], pa[1], pa[2], ca);
    xform4(m, pb[0], pb[1], pb[2], cb);
    double s0 = 0.0, s1 = 1.0;
    // planes z >= 0 and w - z >= 0
    double fa[2] = {ca[2], ca[3] - ca[2]}, fb[2] = {cb[2], cb[3] - cb[2]};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (fa[k] < 0.0 && fb[k] < 0.0) return L;
        if (fa[k] < 0.0) s0 = fmax(s0, fa[k] / (fa[k] - fb[k]));
        if (fb[k] < 0.0) s1 = fmin(s1, fa[k] / (fa[k] - fb[k]));
    }
    if (!(s0 < s1)) return L;
    double c0[4], c1[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        c0[r] = ca[r] + s0 * (cb[r] - ca[r]);
        c1[r] = ca[r] + s1 * (cb[r] - ca[r]);
    }
    if (!(c0[3] > 0.0 && c1[3] > 0.0)) return L;
    L.s0 = s0; L.s1 = s1;
    L.wa = c0[3]; L.wb = c1[3];
    L.ax = (c0[0] / c0[3] + 1.0) * (0.5 * width);
    L.ay = (c0[1] / c0[3] + 1.0) * (0.5 * height);
    L.az = c0[2] / c0[3];
    L.bx = (c1[0] / c1[3] + 1.0) * (0.5 * width);
    L.by = (c1[1] / c1[3] + 1.0) * (0.5 * height);
    L.bz = c1[2] / c1[3];
    L.vx = L.bx - L.ax;
    L.vy = L.by - L.ay;
    L.l2 = L.vx * L.vx + L.vy * L.vy;
    if (!(L.l2 > 0.0) || !(L.l2 < 1e30)) return L;
    L.xmajor = fabs(L.vx) >= fabs(L.vy);
    double hw = 0.5 * line_width;
    L.lim = hw * hw * L.l2;
    double major_a = L.xmajor ? L.ax : L.ay, major_b = L.xmajor ? L.bx : L.by;
    double vmaj = L.xmajor ? L.vx : L.vy;
    int extent = L.xmajor ? width : height;
    double lo = fmin(major_a, major_b) - hw, hi = fmax(major_a, major_b) + hw;
    double c_lo = floor(lo - 0.5), c_hi = ceil(hi - 0.5);
    c_lo = fmax(c_lo, 0.0);
    c_hi = fmin(c_hi, double(extent - 1));
    if (c_lo > c_hi) return L;
    L.col0 = int(c_lo);
    L.col_end = int(c_hi) + 1;
    // The strip |minor - line(major)| <= h holds exactly the pixels with
    // cross^2 <= (w/2)^2 |v|^2 (the cross product is |v_major| times the
    // minor-axis distance).  Both that rule and the row range below round in
    // float64 with errors of a few ulps of the largest window coordinate;
    // the margin covers them many times over.
    L.h = hw * sqrt(L.l2) / fabs(vmaj);
    L.slope = L.xmajor ? L.vy / L.vx : L.vx / L.vy;
    double mag = fmax(fmax(fabs(L.ax), fabs(L.ay)), fmax(fmax(fabs(L.bx), fabs(L.by)), double(width + height)));
    L.margin = 1e-6 + 1e-10 * mag;
    L.valid = true;
    return L;
}

// The walk over the pixels of the box [x0, x1) x [y0, y1) only.
__device__ __forceinline__ Walk walk_start(const LineSetup& L, int x0, int y0, int x1, int y1) {
    Walk w{0, 1, 0, 0, 0, -1};
    if (!L.valid) return w;
    const int mlo = L.xmajor ? x0 : y0, mhi = L.xmajor ? x1 : y1;
    w.col = max(L.col0, mlo) - 1;
    w.col_end = min(L.col_end, mhi);
    w.minor_lo = L.xmajor ? y0 : x0;
    w.minor_hi = (L.xmajor ? y1 : x1) - 1;
    return w;
}

// The line rule at pixel centre (px+0.5, py+0.5); t returned on success.
__device__ __forceinline__ bool covers(const LineSetup& L, int px, int py, double& t) {
    double cx = px + 0.5 - L.ax, cy = py + 0.5 - L.ay;
    double dp = cx * L.vx + cy * L.vy;
    if (!(dp >= 0.0 && dp < L.l2)) return false;
    double cr = L.vx * cy - L.vy * cx;
    if (!(cr * cr <= L.lim)) return false;
    t = dp / L.l2;
    return true;
}

// Advances a lane's walk to its next covered pixel; false once the walk has
// passed its last column (and on every later call).
__device__ __forceinline__ bool next_fragment(const LineSetup& L, Walk& w, int& px, int& py, double& t) {
    if (!L.valid) return false;
    for (;;) {
        if (w.row > w.row_end) {
            if (w.col + 1 >= w.col_end) {
                w.col = w.col_end;
                return false;
            }
            ++w.col;
            const double cm = w.col + 0.5;
            const double line_minor = L.xmajor ? L.ay + (cm - L.ax) * L.slope : L.ax + (cm - L.ay) * L.slope;
            // rows whose centre (row + 0.5) lies within h (+ margin) of the line
            double r0 = ceil(line_minor - L.h - 0.5 - L.margin), r1 = floor(line_minor + L.h - 0.5 + L.margin);
            r0 = fmax(r0, double(w.minor_lo));
            r1 = fmin(r1, double(w.minor_hi));
            if (r0 > r1) continue;
            w.row = int(r0);
            w.row_end = int(r1);
        }
        const int row = w.row++;
        px = L.xmajor ? w.col : row;
        py = L.xmajor ? row : w.col;
        if (covers(L, px, py, t)) return true;
    }
}

// Bilinear, clamp-to-edge sample of the shadow map at uv.  Branch-free, so
// that the taps' loads can all be in flight together: a coordinate beyond
// +-1e9 texels is replaced by 0, whose weights (0) return texel (0, 0) --
// the value returned for it before.
__device__ __forceinline__ double shadow_tex(const RasterArgs& a, double u, double v) {
    double x = u * a.sw - 0.5, y = v * a.sh - 0.5;
    const bool sane = fabs(x) < 1e9 && fabs(y) < 1e9;
    x = sane ? x : 0.0;
    y = sane ? y : 0.0;
    double fx = floor(x), fy = floor(y);
    double wx = x - fx, wy = y - fy;
    int x0 = int(fx), y0 = int(fy);
    int x1 = min(max(x0 + 1, 0), a.sw - 1), y1 = min(max(y0 + 1, 0), a.sh - 1);
    x0 = min(max(x0, 0), a.sw - 1);
    y0 = min(max(y0, 0), a.sh - 1);
    double t00 = a.shadow_map[size_t(y0) * a.sw + x0], t10 = a.shadow_map[size_t(y0) * a.sw + x1];
    double t01 = a.shadow_map[size_t(y1) * a.sw + x0], t11 = a.shadow_map[size_t(y1) * a.sw + x1];
    double r0 = t00 + wx * (t10 - t00), r1 = t01 + wx * (t11 - t01);
    return r0 + wy * (r1 - r0);
}

// One PCF tap of approximate_deep_shadow (approximate_deep_shadows.glsl:
// 12-21): the light-space depth gap -> strand count -> transmittance.
__device__ __forceinline__ double pcf_tap(const RasterArgs& a, double u0, double v0, double light_depth, int q) {
    double sd = shadow_tex(a, u0 + a.pcf_dx[q], v0 + a.pcf_dy[q]);
    double depth = fmax(light_depth - sd, 0.0);
    double cnt = depth * 15000.0;
    if (depth > 1e-5) cnt += 1.0;
    return double(expf(float(cnt * a.ln_transmit))) * a.pcf_w[q];
}

// The reference's 3x3 kernel with the taps' per-axis terms formed once: tap
// q = 3j + i sits at (u0 + dx[i], v0 + dy[3j]), so there are three texel
// columns and three rows, each with its floor, weight and clamped texels.
// Each tap's operations and the summation order are those of pcf_tap, so the
// sum is the same bit for bit.  Returns false (nothing summed) when a
// coordinate is beyond +-1e9 texels; the caller then takes pcf_tap's
// per-tap path, which replaces such taps by texel (0, 0).
__device__ __forceinline__ bool pcf3x3_separable(const RasterArgs& a, double u0, double v0, double light_depth,
                                                 double* vis) {
    double wx[3], wy[3];
    int x0[3], x1[3];
    uint32_t r0[3], r1[3];
    bool sane = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double x = (u0 + a.pcf_dx[i]) * a.sw - 0.5, y = (v0 + a.pcf_dy[3 * i]) * a.sh - 0.5;
        sane = sane && fabs(x) < 1e9 && fabs(y) < 1e9;
        const double fx = floor(x), fy = floor(y);
        wx[i] = x - fx;
        wy[i] = y - fy;
        const int ix = int(fx), iy = int(fy);
        x0[i] = min(max(ix, 0), a.sw - 1);
        x1[i] = min(max(ix + 1, 0), a.sw - 1);
        r0[i] = uint32_t(min(max(iy, 0), a.sh - 1)) * uint32_t(a.sw);
        r1[i] = uint32_t(min(max(iy + 1, 0), a.sh - 1)) * uint32_t(a.sw);
    }
    if (!sane) return false;
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int q = 3 * j + i;
            const double t00 = a.shadow_map[r0[j] + x0[i]], t10 = a.shadow_map[r0[j] + x1[i]];
            const double t01 = a.shadow_map[r1[j] + x0[i]], t11 = a.shadow_map[r1[j] + x1[i]];
            const double s0 = t00 + wx[i] * (t10 - t00), s1 = t01 + wx[i] * (t11 - t01);
            const double sd = s0 + wy[j] * (s1 - s0);
            const double depth = fmax(light_depth - sd, 0.0);
            double cnt = depth * 15000.0;
            if (depth > 1e-5) cnt += 1.0;
            sum += double(expf(float(cnt * a.ln_transmit))) * a.pcf_w[q];
        }
    *vis = sum;
    return true;
}

__device__ __forceinline__ void normalize3f(float v[3]) {
    const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const float s = l > 0.0f ? 1.0f / l : 0.0f;
    v[0] *= s; v[1] *= s; v[2] *= s;
}

// Kajiya-Kay (kajiya-kay.glsl:7-24) in float32 for the colour pass: a
// continuous term, so float32 stays within the fragment tolerance and
// frees the registers and issue slots float64 takes; sqrt of a negative
// rounding residue is 0.  The colours are the parameter blocks' floats.
__device__ __forceinline__ void kajiya_kay_f32(const double diffuse[3], const double specular[3], double p,
                                               const float t[3], const float l[3], const float e[3], double out[3]) {
    const float ctl = t[0] * l[0] + t[1] * l[1] + t[2] * l[2];
    const float cte = t[0] * e[0] + t[1] * e[1] + t[2] * e[2];
    const float stl = sqrtf(fmaxf(1.0f - ctl * ctl, 0.0f)), ste = sqrtf(fmaxf(1.0f - cte * cte, 0.0f));
    const float base = fmaxf(ctl * cte + stl * ste, 0.0f);
    const float spec = powf(base, float(p));
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = double(float(diffuse[c]) * stl + float(specular[c]) * spec);
}

// Per-wave LDS of the colour pass: the shading data of the wave's 64
// segments (structure of arrays, so lanes reading different segments hit
// different banks) and a ring of queued fragments.  Positions, clip w and
// depth stay float64 (they decide coverage and the shadow test); the
// tangents and thickness feed continuous terms only and are float32, which
// brings a workgroup to 37 KB, four to a CU.
constexpr int kQueue = 128;          // >= 2 x 64: a fill round never overflows
enum SegField { kPa = 0, kPb = 3, kWa = 6, kWb, kAz, kBz, kSegFields };
enum SegFieldF32 { kTa = 0, kTb = 3, kTha = 6, kThb, kSegFieldsF32 };
struct WaveShared {
    double seg[kSegFields][kWave];
    float segf[kSegFieldsF32][kWave];
    uint32_t segment[kWave];
    double qt[kQueue];
    uint32_t qpix[kQueue];
    uint32_t qlane[kQueue];
};
constexpr int kRasterBlock = 256;

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The shadow pass keeps its block's depths per texel in an LDS table: the
// block's 256 segments are neighbours in the load-time order and cover the
// same texels many times over, while a device-scope atomicMin on the map
// is a memory-side atomic per fragment.  A fragment whose texel finds no
// slot within kDepthProbes probes takes the global atomic at once; the
// table's minima are flushed with one global atomicMin per texel at the
// end.  The minimum does not depend on the order: the map is the same bit
// for bit.
constexpr uint32_t kDepthSlotBits = 10, kDepthSlots = 1u << kDepthSlotBits, kDepthProbes = 8;
constexpr uint32_t kDepthEmpty = 0xffffffffu;   // no texel index (maps are at most 16384^2)

struct DepthTable {
    uint32_t key[kDepthSlots];
    uint32_t bits[kDepthSlots];
};

// The slot holding texel k (claimed when free), or kDepthEmpty.
__device__ __forceinline__ uint32_t depth_slot(DepthTable& t, uint32_t k) {
    const uint32_t h = (k * 2654435761u) >> (32 - kDepthSlotBits);
    for (uint32_t i = 0; i < kDepthProbes; ++i) {
        const uint32_t slot = (h + i) & (kDepthSlots - 1u);
        const uint32_t prev = atomicCAS(&t.key[slot], kDepthEmpty, k);
        if (prev == kDepthEmpty || prev == k) return slot;
    }
    return kDepthEmpty;
}

__global__ void __launch_bounds__(kRasterBlock)
shadow_raster_kernel(RasterArgs) {
    const RasterArgs& a = kernel_args<RasterArgs>();
    __shared__ DepthTable table;
    for (uint32_t i = threadIdx.x; i < kDepthSlots; i += blockDim.x) {
        table.key[i] = kDepthEmpty;
        table.bits[i] = kDepthEmpty;
    }
    __syncthreads();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    LineSetup L;
    L.valid = false;
    if (s < a.n_pairs) {
        const uint32_t ia = a.indices[2u * s], ib = a.indices[2u * s + 1u];
        if (ia < a.n_vertices && ib < a.n_vertices) {
            float3 fa = load_xyz(a.xyz, ia), fb = load_xyz(a.xyz, ib);
            double pa[3] = {fa.x, fa.y, fa.z}, pb[3] = {fb.x, fb.y, fb.z};
            L = setup_line(a.lm, pa, pb, a.width, a.height, a.line_width);
        }
    }
    Walk w = walk_start(L, 0, 0, a.width, a.height);
    int px, py;
    double t;
    while (next_fragment(L, w, px, py, t)) {
        const double z = L.az + t * (L.bz - L.az);
        const uint32_t texel = uint32_t(py) * uint32_t(a.width) + uint32_t(px), bits = __float_as_uint(float(z));
        const uint32_t slot = depth_slot(table, texel);
        if (slot != kDepthEmpty)
            __hip_atomic_fetch_min(&table.bits[slot], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicMin(&a.depth_bits[texel], bits);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kDepthSlots; i += blockDim.x)
        if (table.key[i] != kDepthEmpty) atomicMin(&a.depth_bits[table.key[i]], table.bits[i]);
}

// The PPLL insert of the colour pass (ppll.glsl:23-48).  Node ids come
// from the block's reservation (one counter atomic per block, made after a
// counting walk; the wanting lanes of a round take consecutive ids through
// an LDS cursor): a device-wide atomic on the one counter per wave and round
// serializes -- 261 k of them measured 3.4 ms at 16.7 M fragments.  Nodes at
// or past the pool end are dropped, as in the shader; the node is linked by
// exchanging the pixel's head.
__device__ __forceinline__ uint32_t block_insert(const PpllView& p, bool want, uint32_t pixel, float4 color,
                                                 float depth, uint32_t block_base, uint32_t* block_next) {
    unsigned long long mask = __ballot(want);
    if (mask == 0ull) return kPpllNull;
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t off = 0;
    if (lane == leader) off = atomicAdd(block_next, uint32_t(__popcll(mask)));  // LDS
    off = __shfl(off, leader, kWave);
    if (!want) return kPpllNull;
    const uint32_t node = block_base + off + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    if (node >= p.node_count) return kPpllNull;  // pool exhausted: fragment dropped
    uint32_t* n = p.nodes + 3ull * node;
    n[0] = pack_unorm4x8(color);
    n[1] = __float_as_uint(depth);
    n[2] = atomicExch(&p.heads[pixel], node);
    return node;
}

// The colour pass over one segment per lane (s, when has).  Every thread
// of the block calls it (it synchronises the block once, for the
// reservation).
__device__ void colour_segments(const RasterArgs& a, WaveShared& W, uint32_t s, bool has, uint32_t* block_sums) {
    const int lane = __lane_id();
    LineSetup L;
    L.valid = false;
    uint32_t ia = 0, ib = 0;
    double pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
    if (has && s < a.n_pairs) {
        ia = a.indices[2u * s];
        ib = a.indices[2u * s + 1u];
        if (ia < a.n_vertices && ib < a.n_vertices) {
            float3 fa = load_xyz(a.xyz, ia), fb = load_xyz(a.xyz, ib);
            pa[0] = fa.x; pa[1] = fa.y; pa[2] = fa.z;
            pb[0] = fb.x; pb[1] = fb.y; pb[2] = fb.z;
            L = setup_line(a.vp, pa, pb, a.width, a.height, a.line_width);
        }
    }
    // per-segment attribute ends (clipped), parked in LDS for whichever lane
    // shades this segment's fragments
    if (L.valid) {
        float3 ua = load_xyz(a.tangents, ia), ub = load_xyz(a.tangents, ib);
        double t0[3] = {ua.x, ua.y, ua.z}, t1[3] = {ub.x, ub.y, ub.z};
        double h0 = a.thickness[ia], h1 = a.thickness[ib];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            W.segf[kTa + c][lane] = float(t0[c] + L.s0 * (t1[c] - t0[c]));
            W.segf[kTb + c][lane] = float(t0[c] + L.s1 * (t1[c] - t0[c]));
            W.seg[kPa + c][lane] = pa[c] + L.s0 * (pb[c] - pa[c]);
            W.seg[kPb + c][lane] = pa[c] + L.s1 * (pb[c] - pa[c]);
        }
        W.segf[kTha][lane] = float(h0 + L.s0 * (h1 - h0));
        W.segf[kThb][lane] = float(h0 + L.s1 * (h1 - h0));
        W.seg[kWa][lane] = L.wa;
        W.seg[kWb][lane] = L.wb;
        W.seg[kAz][lane] = L.az;
        W.seg[kBz][lane] = L.bz;
        W.segment[lane] = s;
    }
    // The block reserves its node ids at once: a counting walk (the walk is
    // repeated below), one LDS atomic per wave, one counter atomic per block.
    // The count covers every covered candidate; nodes reserved for fragments
    // the shading then discards (coverage below 0.001, the depth test) stay
    // unlinked.
    {
        uint32_t mine = 0;
        Walk cw = walk_start(L, 0, 0, a.width, a.height);
        int px, py;
        double t;
        while (next_fragment(L, cw, px, py, t)) ++mine;
        for (int o = kWave / 2; o > 0; o >>= 1) mine += __shfl_xor(mine, o, kWave);
        if (lane == 0) atomicAdd(&block_sums[0], mine);
    }
    __syncthreads();
    if (threadIdx.x == 0) block_sums[1] = block_sums[0] ? atomicAdd(a.ppll.counter, block_sums[0]) : 0u;
    __syncthreads();
    const uint32_t block_base = block_sums[1];
    Walk w = walk_start(L, 0, 0, a.width, a.height);
    int head = 0, count = 0;  // wave-uniform ring state
    for (;;) {
        // fill: every lane with fragments left queues its next one
        while (count < kWave) {
            int px = 0, py = 0;
            double t = 0.0;
            const bool found = next_fragment(L, w, px, py, t);
            const unsigned long long m = __ballot(found);
            if (m == 0ull) break;  // every walk has ended
            if (found) {
                const int slot = (head + count + int(__popcll(m & ((1ull << lane) - 1ull)))) & (kQueue - 1);
                W.qt[slot] = t;
                W.qpix[slot] = uint32_t(py) * uint32_t(a.width) + uint32_t(px);
                W.qlane[slot] = uint32_t(lane);
            }
            count += int(__popcll(m));
        }
        if (count == 0) break;
        wave_lds_fence();
        // shade: lane i takes the i-th queued fragment
        const int n = count < kWave ? count : kWave;
        const int slot = (head + lane) & (kQueue - 1);
        bool frag = lane < n;
        uint32_t pix = 0, src = 0;
        double t = 0.0;
        if (frag) {
            t = W.qt[slot];
            pix = W.qpix[slot];
            src = W.qlane[slot];
        }
        head = (head + n) & (kQueue - 1);
        count -= n;
        double z = 0.0;
        float4 color = make_float4(0, 0, 0, 0);
        if (frag) {
            z = W.seg[kAz][src] + t * (W.seg[kBz][src] - W.seg[kAz][src]);
            // early depth test (LESS against the subpass depth, no writes)
            frag = float(z) < (a.depth_buffer ? a.depth_buffer[pix] : 1.0f);
        }
        if (frag) {
            const int px = int(pix % uint32_t(a.width)), py = int(pix / uint32_t(a.width));
            // perspective-correct varyings
            const double wa = W.seg[kWa][src], wb = W.seg[kWb][src];
            double iw = (1.0 - t) / wa + t / wb;
            double sp = (t / wb) / iw;
            const float spf = float(sp);
            double P[3];
            float T[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double p0 = W.seg[kPa + c][src];
                const float t0 = W.segf[kTa + c][src];
                P[c] = p0 + sp * (W.seg[kPb + c][src] - p0);
                T[c] = t0 + spf * (W.segf[kTb + c][src] - t0);
            }
            const double tha = W.segf[kTha][src];
            double thick = tha + sp * (double(W.segf[kThb][src]) - tha);
            // gpaa (gpaa.glsl:33-51)
            double cl[4];
            xform4(a.vp, P[0], P[1], P[2], cl);
            double sx = (cl[0] / cl[3] + 1.0) * (a.res_x / 2.0), sy = (cl[1] / cl[3] + 1.0) * (a.res_y / 2.0);
            double ddx = sx - (px + 0.5), ddy = sy - (py + 0.5);
            double coverage = 1.0 - sqrt(ddx * ddx + ddy * ddy) / (a.line_width / 2.0);
            coverage *= a.alpha;
            frag = !(coverage < 0.001);
            if (frag) {
                coverage *= 1.0 - a.lod;
                coverage *= thick * (1.0 / 0.042);
                double shading[3] = {1.0, 1.0, 1.0};
                if (a.shading_model == 0) {
                    float e[3] = {float(P[0] - a.eye[0]), float(P[1] - a.eye[1]), float(P[2] - a.eye[2])};
                    float l[3] = {float(a.light_origin[0] - P[0]), float(a.light_origin[1] - P[1]),
                                  float(a.light_origin[2] - P[2])};
                    normalize3f(e);
                    normalize3f(l);
                    kajiya_kay_f32(a.hair_color, a.intensity, a.shininess, T, l, e, shading);
                }
                double occlusion = 1.0;
                if (a.adsm_on == 1 && a.shading_model != 3) {
                    double ls[4];
                    xform4(a.lm, P[0], P[1], P[2], ls);
                    double light_depth = ls[2] / ls[3];
                    double u0 = ls[0] / ls[3], v0 = ls[1] / ls[3];
                    double vis = 0.0, tot = 0.0;
                    if (a.pcf_taps == 9 && pcf3x3_separable(a, u0, v0, light_depth, &vis)) {
#pragma unroll
                        for (int q = 0; q < 9; ++q) tot += a.pcf_w[q];
                    } else {
                        for (int q = 0; q < a.pcf_taps; ++q) {
                            vis += pcf_tap(a, u0, v0, light_depth, q);
                            tot += a.pcf_w[q];
                        }
                    }
                    occlusion *= vis / tot;
                }
                if (a.shading_model != 2) {
                    double st[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) st[c] = (P[c] - a.vorigin[c]) / a.vsize[c] * a.vres[c] - 0.5;
                    occlusion *= a.vol.f32 ? local_ambient_occlusion_f32pow(a.vol, st, a.ao_scaling, a.ao_clamp,
                                                                            a.ao_exponent)
                                           : local_ambient_occlusion_cells_f32pow(a.vol, st, a.ao_scaling,
                                                                                  a.ao_clamp, a.ao_exponent);
                }
                color = make_float4(float(shading[0] * occlusion), float(shading[1] * occlusion),
                                    float(shading[2] * occlusion), float(coverage));
            }
        }
        const uint32_t node = block_insert(a.ppll, frag, pix, color, float(z), block_base, &block_sums[2]);
        if (node != kPpllNull) {
            if (a.cap.segment) a.cap.segment[node] = W.segment[src];
            if (a.cap.pixel) a.cap.pixel[node] = pix;
            if (a.cap.color) a.cap.color[node] = color;
        }
        wave_lds_fence();
    }
}


// The colour pass: one segment per lane, the whole target.
__global__ void __launch_bounds__(kRasterBlock) raster_kernel(RasterArgs) {
    const RasterArgs& a = kernel_args<RasterArgs>();
    __shared__ WaveShared shared[kRasterBlock / kWave];
    __shared__ uint32_t block_sums[3];  // fragments counted, reservation base, LDS allocation cursor
    if (threadIdx.x < 3) block_sums[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    colour_segments(a, shared[threadIdx.x / kWave], s, s < a.n_pairs, block_sums);
}

void load_cm(const float* m, double* out) {
    for (int i = 0; i < 16; ++i) out[i] = m[i];
}

void matmul_cm(const float* x, const float* y, double* out) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += double(x[k * 4 + r]) * double(y[c * 4 + k]);
            out[c * 4 + r] = acc;
        }
}

__global__ void fill_depth_kernel(uint32_t* bits, size_t n, uint32_t v) {
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) bits[i] = v;
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_shadow_depth(const float* xyz, uint32_t n_vertices, const uint32_t* indices,
                                 uint32_t n_indices, const float* light_vp, float line_width, uint32_t size_w,
                                 uint32_t size_h, float* depth_map, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(light_vp && depth_map, VKHR_ERR_INVALID, "null argument");
    VKHR_REQUIRE(size_w && size_h && size_w <= 16384 && size_h <= 16384, VKHR_ERR_INVALID, "bad map size");
    VKHR_REQUIRE(line_width > 0.0f && line_width < 64.0f, VKHR_ERR_INVALID, "line width %g", line_width);
    uint32_t n_pairs = n_indices / 2;
    VKHR_REQUIRE(n_pairs == 0 || (xyz && indices), VKHR_ERR_INVALID, "null geometry");
    hipStream_t st = as_stream(stream);
    size_t n = size_t(size_w) * size_h;
    fill_depth_kernel<<<grid_for(n, 256, 8192), 256, 0, st>>>(reinterpret_cast<uint32_t*>(depth_map), n,
                                                             __builtin_bit_cast(uint32_t, 1.0f));
    VKHR_CHECK_LAUNCH();
    if (n_pairs == 0) return VKHR_OK;
    RasterArgs* a = new RasterArgs();
    std::memset(a, 0, sizeof(RasterArgs));
    load_cm(light_vp, a->lm);
    a->line_width = line_width;
    a->width = int(size_w);
    a->height = int(size_h);
    a->xyz = xyz;
    a->indices = indices;
    a->n_pairs = n_pairs;
    a->n_vertices = n_vertices;
    a->depth_bits = reinterpret_cast<uint32_t*>(depth_map);
    shadow_raster_kernel<<<(n_pairs + kRasterBlock - 1) / kRasterBlock, kRasterBlock, 0, st>>>(*a);
    delete a;
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_raster_strands(const float* xyz, const float* tangents, const float* thickness,
                                   uint32_t n_vertices, const uint32_t* indices, uint32_t n_indices,
                                   const vkhr_camera* cam, const vkhr_light* light, const vkhr_strand_params* sp,
                                   const vkhr_render_params* rp, const vkhr_volume* vol, const float* shadow_map,
                                   uint32_t shadow_w, uint32_t shadow_h, const float* depth_buffer,
                                   const vkhr_ppll* ppll, const vkhr_frag_capture* capture, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(cam && light && sp && rp && ppll, VKHR_ERR_INVALID, "null parameter struct");
    VKHR_REQUIRE(ppll->heads && ppll->nodes && ppll->counter, VKHR_ERR_INVALID, "ppll is incomplete");
    uint32_t n_pairs = n_indices / 2;
    VKHR_REQUIRE(n_pairs == 0 || (xyz && indices && tangents && thickness), VKHR_ERR_INVALID, "null geometry");
    const bool adsm = rp->adsm_on == 1 && rp->shading_model != 3;
    const bool lao = rp->shading_model != 2;
    VKHR_REQUIRE(!adsm || (shadow_map && shadow_w && shadow_h), VKHR_ERR_INVALID, "ADSM needs a shadow map");
    VKHR_REQUIRE(!lao || (vol && (vol->cells || vol->density_f32)), VKHR_ERR_INVALID,
                 "LAO needs the density volume");
    // the LAO taps index cells with 32-bit sums, the PCF taps texels with 32-bit row offsets
    if (lao && (rc = check_volume(vol->nx, vol->ny, vol->nz, true, "LAO volume"))) return rc;
    VKHR_REQUIRE(!adsm || (shadow_w <= 16384 && shadow_h <= 16384), VKHR_ERR_INVALID, "shadow map %ux%u",
                 shadow_w, shadow_h);
    VKHR_REQUIRE(sp->strand_radius > 0.0f && sp->strand_radius < 64.0f, VKHR_ERR_INVALID, "strand_radius %g",
                 sp->strand_radius);
    if (n_pairs == 0) return VKHR_OK;
    RasterArgs* a = new RasterArgs();
    std::memset(a, 0, sizeof(RasterArgs));
    matmul_cm(cam->projection, cam->view, a->vp);
    load_cm(light->matrix, a->lm);
    for (int c = 0; c < 3; ++c) {
        a->eye[c] = cam->position[c];
        a->light_origin[c] = light->origin[c];
        a->hair_color[c] = sp->hair_color[c];
        a->intensity[c] = light->intensity[c];
        a->vorigin[c] = sp->volume_bounds.origin[c];
        a->vsize[c] = sp->volume_bounds.size[c];
    }
    a->alpha = sp->hair_opacity;
    a->shininess = sp->hair_shininess;
    a->ln_transmit = std::log(1.0 - double(sp->hair_opacity));
    a->lod = rp->renderer == 0 ? 0.0 : rp->renderer == 2 ? 1.0 : [&] {
        double t = (double(cam->look_at_distance) - rp->lod_magnified_distance) /
                   (double(rp->lod_minified_distance) - rp->lod_magnified_distance);
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        return t * t * (3.0 - 2.0 * t);
    }();
    a->line_width = sp->strand_radius;
    a->res_x = cam->resolution[0];
    a->res_y = cam->resolution[1];
    // approximate_deep_shadows (approximate_deep_shadows.glsl:41-73): taps at
    // (x, y) / (size / stride) with Gaussian weights; the shader's exponent
    // reads -(x^2 + y^2) / 2 * sigma^2
    if (adsm) {
        double kw = rp->adsm_kernel_size, smoothing = rp->adsm_stride_size;
        VKHR_REQUIRE(kw >= 1 && kw <= 9 && smoothing > 0, VKHR_ERR_INVALID, "ADSM kernel %g stride %g", kw,
                     smoothing);
        float kr = float((kw - 1.0) / 2.0);
        double sigma = (kw / 2.0) / 2.4, sigma2 = sigma * sigma;
        int q = 0;
        for (float y = -kr; y <= kr; y += 1.0f)
            for (float x = -kr; x <= kr; x += 1.0f) {
                double exponent = -1.0 * (double(x) * x + double(y) * y) / 2.0 * sigma2;
                a->pcf_w[q] = 1.0 / (2.0 * M_PI * sigma2) * std::pow(M_E, exponent);
                a->pcf_dx[q] = double(x) / (double(shadow_w) / smoothing);
                a->pcf_dy[q] = double(y) / (double(shadow_h) / smoothing);
                ++q;
            }
        a->pcf_taps = q;
    }
    a->ao_scaling = double(rp->occlusion_radius) / 0.5;
    a->ao_clamp = rp->ao_clamp;
    a->ao_exponent = rp->ao_exponent;
    if (vol) {
        a->vol.cells = reinterpret_cast<const unsigned long long*>(vol->cells);
        a->vol.f32 = vol->density_f32;
        a->vol.nx = int(vol->nx); a->vol.ny = int(vol->ny); a->vol.nz = int(vol->nz);
        a->vres[0] = vol->nx; a->vres[1] = vol->ny; a->vres[2] = vol->nz;
    }
    a->shading_model = rp->shading_model;
    a->adsm_on = rp->adsm_on;
    a->width = int(ppll->width);
    a->height = int(ppll->height);
    a->shadow_map = shadow_map;
    a->sw = int(shadow_w);
    a->sh = int(shadow_h);
    a->depth_buffer = depth_buffer;
    a->xyz = xyz;
    a->tangents = tangents;
    a->thickness = thickness;
    a->indices = indices;
    a->n_pairs = n_pairs;
    a->n_vertices = n_vertices;
    a->ppll = PpllView{ppll->heads, ppll->nodes, ppll->counter, ppll->node_count};
    if (capture) a->cap = CaptureView{capture->segment, capture->pixel, reinterpret_cast<float4*>(capture->color)};
    raster_kernel<<<(n_pairs + kRasterBlock - 1) / kRasterBlock, kRasterBlock, 0, as_stream(stream)>>>(*a);
    delete a;
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
