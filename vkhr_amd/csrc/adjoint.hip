// Adjoint of the raymarch pass with respect to the float density volume and
// the absorption / ADSM-thickness / Kajiya-Kay-exponent parameters.
//
// The reference's march accumulates plain sums (volume_rendering.glsl:36-37,
// approximate_deep_shadows.glsl:28-33): no transmittance product runs along a
// ray, so every sample of one ray carries the same gradient weight -- for the
// primary ray dL/d(acc) (through the coverage a = lod * alpha * acc / iso),
// for the shadow ray dL/d(strands) (through (1-alpha)^(k * sum)).  The
// isosurface and surface choices are piecewise constant and carry none; the
// clamp to [0,1] passes gradient only strictly inside.
//
// Rays are re-walked in chunks of 32 steps: the march state (the float
// parameter t of the shader's loop) is carried from chunk to chunk and the
// positions are recomputed from the ray set-up (nothing per sample is
// stored).  The scatter of weight * trilinear weight goes to a corner-packed
// accumulator of 8 floats per cell (the cell layout of the density: cell
// (i,j,k) for i,j,k in [-1, n-1] holds texels i..i+1 x j..j+1 x k..k+1), and
// the lanes of a wave are transposed (through LDS) so that 8 lanes add the
// 8 corners of one sample: one 32-byte line per sample instead of 8
// scattered words (device atomics are bound by line requests), and
// neighbouring rays of a pixel row whose samples share a cell add once.  A gather pass then folds the cells
// into the texel gradient.  Parameter gradients are reduced per block, then
// one atomic each.
#include "march.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace vkhr {
namespace {

constexpr int kChunk = 32;
constexpr uint32_t kNoCell = 0xffffffffu;

struct AdjointArgs {
    RaymarchArgs f;
    const MarchRecord* record;
    const float4* d_rgba;
    float* cell_grad;           // 8 floats per cell, (nx+1)(ny+1)(nz+1) cells
    float* d_params;            // absorption, ADSM thickness, shininess
    const float4* target;       // with a target: dL/d(rgba) = 2 (rgba - target), L = sum of squares
    float* loss;                // L is added here (with a target; nullable)
    int n_steps;                // iterations of the shader's float loop
    double ln_transmit;         // log(1 - alpha)
    double light_texel[3];
};

__device__ __forceinline__ bool in_range(const RaymarchArgs& f, const double x[3]) {
    return x[0] >= -1.0 && x[1] >= -1.0 && x[2] >= -1.0 && x[0] < double(f.nx) && x[1] < double(f.ny) &&
           x[2] < double(f.nz);
}

// The cell of a texel coordinate inside [-1, n)^3 and its trilinear
// fractions (the weights of density_at).
__device__ __forceinline__ uint32_t cell_of(const RaymarchArgs& f, const double x[3], float w[3]) {
    const double fx = floor(x[0]), fy = floor(x[1]), fz = floor(x[2]);
    w[0] = float(x[0] - fx);
    w[1] = float(x[1] - fy);
    w[2] = float(x[2] - fz);
    const uint32_t cx = uint32_t(f.nx) + 1u, cy = uint32_t(f.ny) + 1u;
    return uint32_t(int(fx) + 1) + cx * (uint32_t(int(fy) + 1) + cy * uint32_t(int(fz) + 1));
}

__device__ __forceinline__ int wave_max(int v) {
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// A wave's staging area of the scatter: every lane's sample (its cell and
// the 8 corner weights times the ray's gradient weight).
struct ScatterLds {
    uint32_t cell[kWave];
    float w[8][kWave];
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Adds c * (trilinear corner weights) of every lane's sample to the cell
// accumulator; lanes without a sample pass cell == kNoCell.  The samples
// are staged in LDS and transposed: the 8 lanes of each group of 8 (one
// row of the wave's 8x8 pixel tile) walk the group's 8 samples in order,
// lane k adding corner k, so one sample is one 32-byte line request; a run
// of samples of the same cell along the row (neighbouring rays in one cell)
// is summed first and added once.  Every lane of the wave must call it.
__device__ __forceinline__ void scatter_wave(float* __restrict__ cg, ScatterLds& L, uint32_t cell, float wx,
                                             float wy, float wz, float c) {
    if (!__ballot(cell != kNoCell)) return;
    const int lane = __lane_id();
    const int k = lane & 7, group = lane & 56;
    L.cell[lane] = cell;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        L.w[q][lane] = c * (((q & 1) ? wx : 1.0f - wx) * ((q & 2) ? wy : 1.0f - wy) * ((q & 4) ? wz : 1.0f - wz));
    wave_lds_sync();
    uint32_t run = kNoCell;
    float acc = 0.0f;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint32_t sc = L.cell[group | r];
        const float v = L.w[k][group | r];
        if (sc != run) {
            if (run != kNoCell) unsafeAtomicAdd(&cg[size_t(run) * 8u + size_t(k)], acc);
            run = sc;
            acc = 0.0f;
        }
        acc += v;  // (kNoCell runs are never added)
    }
    if (run != kNoCell) unsafeAtomicAdd(&cg[size_t(run) * 8u + size_t(k)], acc);
    wave_lds_sync();  // the area is rewritten by the next call
}

// One family of rays (primary or shadow): x(t) = x0 + t * dx for the
// sample parameters t of the shader's float loop, samples [0, n) of each
// active lane's ray, weight c per sample.  The loop runs to the wave's
// longest ray so that every lane takes part in the scatter.
__device__ void scatter_rays(const AdjointArgs& a, ScatterLds& L, bool active, const double x0[3], const double dx[3],
                             int n, float c) {
    const RaymarchArgs& f = a.f;
    const float step = f.step_size;
    const int n_lane = active ? n : 0;
    const int n_wave = wave_max(n_lane);
    float t = 0.0f;  // the march state carried across chunks
    for (int i0 = 0; i0 < n_wave; i0 += kChunk) {
        const int i1 = min(i0 + kChunk, n_wave);
        float tt = t;
        for (int i = i0; i < i1; ++i, tt += step) {
            const double x[3] = {fma(double(tt), dx[0], x0[0]), fma(double(tt), dx[1], x0[1]),
                                 fma(double(tt), dx[2], x0[2])};
            float w[3] = {0.0f, 0.0f, 0.0f};
            uint32_t cell = kNoCell;
            if (i < n_lane && in_range(f, x)) cell = cell_of(f, x, w);
            scatter_wave(a.cell_grad, L, cell, w[0], w[1], w[2], c);
        }
        // advance the carried state past this chunk
        for (int i = 0; i < kChunk; ++i) t += step;
    }
}

// Folds the corner-packed accumulator into the texel gradient (added).
__global__ void __launch_bounds__(256)
gather_cells_kernel(const float* __restrict__ cg, float* __restrict__ grad, int nx, int ny, int nz) {
    const size_t n = size_t(nx) * ny * nz;
    const size_t cx = size_t(nx) + 1, cy = size_t(ny) + 1;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const size_t x = i % size_t(nx), yz = i / size_t(nx);
        const size_t y = yz % size_t(ny), z = yz / size_t(ny);
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            // texel (x,y,z) is corner q of the cell whose base is (x,y,z) - corner
            const size_t c = (x + 1 - (q & 1)) + cx * ((y + 1 - ((q >> 1) & 1)) + cy * (z + 1 - (q >> 2)));
            s += cg[c * 8u + size_t(q)];
        }
        grad[i] += s;
    }
}

// Which walks adjoint_tile takes: the primary rays with the LAO taps and
// the parameter gradients (and the loss), the shadow rays, or both.  The
// per-style kernel gives each family a workgroup of its own, so a tile's
// critical path is the longer walk rather than the sum of the two.
enum AdjointWalks { kWalkPrimary = 1, kWalkShadow = 2, kWalkBoth = 3 };

// The adjoint of one pixel (volume.frag:31-102 differentiated).  Called by
// all 256 threads of a block: 2x2 waves of 8x8 pixels from (x0, y0); red is
// the block's reduction buffer for the parameter gradients.
__device__ __forceinline__ void adjoint_tile(const AdjointArgs& a, int x0, int y0, float (*red)[256 / kWave],
                                             ScatterLds* scatter, int walks = kWalkBoth) {
    // red: 4 x (waves per block) floats
    const RaymarchArgs& f = a.f;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    ScatterLds& L = scatter[wave];
    const int px = x0 + (wave & 1) * 8 + (lane & 7);
    const int py = y0 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = px < f.width && py < f.height;
    const size_t pix = inside ? size_t(py) * f.width + px : 0;

    bool frag = false;
    MarchRecord r;
    float4 g = make_float4(0, 0, 0, 0);
    float sq = 0.0f;
    if (inside) {
        r = a.record[pix];
        frag = r.state == 1;
        if (a.target) {  // L2 loss against the target image
            const float4 o = f.rgba[pix], t = a.target[pix];
            const float4 d = make_float4(o.x - t.x, o.y - t.y, o.z - t.z, o.w - t.w);
            g = make_float4(2.0f * d.x, 2.0f * d.y, 2.0f * d.z, 2.0f * d.w);
            sq = d.x * d.x + d.y * d.y + d.z * d.z + d.w * d.w;
        } else {
            g = a.d_rgba[pix];
        }
    }
    double c_primary = 0.0, c_shadow = 0.0, G = 0.0;
    double d_alpha = 0.0, d_k = 0.0, d_p = 0.0;
    const bool adsm_on = f.adsm_on == 1 && f.shading_model != 3;
    const bool lao_on = f.shading_model != 2;
    if (frag) {
        const double gv[4] = {g.x, g.y, g.z, g.w};
        const double cov = f.lod * f.alpha * double(r.acc) / f.iso;
        const bool ua = cov > 0.0 && cov < 1.0;
        double rgb[3];
        for (int c = 0; c < 3; ++c) {
            rgb[c] = double(r.shade[c]) * double(r.adsm) * double(r.lao);
            const bool uc = rgb[c] > 0.0 && rgb[c] < 1.0;
            if (uc) G += gv[c] * rgb[c];
            if (uc && f.shading_model == 0 && r.spec_base > 0.0f) {
                double spec = pow(double(r.spec_base), f.shininess);
                d_p += gv[c] * double(r.adsm) * double(r.lao) * f.intensity[c] * spec * log(double(r.spec_base));
            }
        }
        if (ua) {
            c_primary = gv[3] * f.lod * f.alpha / f.iso;
            d_alpha += gv[3] * f.lod * double(r.acc) / f.iso;
        }
        if (adsm_on) {
            c_shadow = G * a.ln_transmit * f.adsm_k;
            d_k = G * a.ln_transmit * double(r.shadow_sum);
            d_alpha += G * (f.adsm_k * double(r.shadow_sum)) * (-1.0 / (1.0 - f.alpha));
        }
    }
    const bool primary = (walks & kWalkPrimary) != 0, shadow = (walks & kWalkShadow) != 0;
    if (primary) {  // workgroup-uniform
        // parameter gradients (and the loss): wave, then block, then one atomic each
        const float pv[4] = {float(d_alpha), float(d_k), float(d_p), sq};
        for (int k = 0; k < 4; ++k) {
            float v = pv[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
            if (lane == 0) red[k][wave] = v;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            float v = 0.0f;
            for (int w = 0; w < int(blockDim.x / kWave); ++w) v += red[threadIdx.x][w];
            float* dst = threadIdx.x < 3 ? &a.d_params[threadIdx.x] : a.loss;
            if (v != 0.0f && dst) atomicAdd(dst, v);
        }
        __syncthreads();  // red is reused by the next tile of a persistent block

        // primary rays
        const Ray ray = setup_ray(f, px, py, inside && frag && c_primary != 0.0);
        scatter_rays(a, L, ray.hit && c_primary != 0.0, ray.xs, ray.dx, frag ? r.n_primary : 0, float(c_primary));
    }
    // shadow rays: from the surface point to the light origin, every sample
    double s[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    if (frag) {
        s[0] = r.sx; s[1] = r.sy; s[2] = r.sz;
        for (int k = 0; k < 3; ++k) sl[k] = a.light_texel[k] - s[k];
    }
    // samples past the volume's support carry no gradient: the walk stops
    // there (an upper bound on their count is enough, in_range does the rest)
    int n_shadow = 0;
    if (frag && shadow) {
        double lo, hi;
        support_interval(f, s, sl, &lo, &hi);
        if (hi >= 0.0 && lo <= hi) n_shadow = min(a.n_steps, int(fmin(hi, 2.0) / double(f.step_size) * 1.0001) + 2);
    }
    if (shadow) scatter_rays(a, L, frag && adsm_on && c_shadow != 0.0, s, sl, n_shadow, float(c_shadow));
    if (!primary) return;
    // LAO taps: d ln(lao)/dD = -e / (8 - D), through min(v, clamp)
    const bool lao = frag && lao_on && G != 0.0;
    double v[8], D = 0.0;
    double tap[8][3];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        tap[n][0] = s[0] + ((n & 1) ? 0.5 : -0.5) * f.ao_scaling;
        tap[n][1] = s[1] + (((n >> 1) & 1) ? 0.5 : -0.5) * f.ao_scaling;
        tap[n][2] = s[2] + ((n >> 2) ? 0.5 : -0.5) * f.ao_scaling;
        v[n] = lao ? double(density_at(f.vol, tap[n][0], tap[n][1], tap[n][2])) : 0.0;
        D += fmin(v[n], f.ao_clamp);
    }
    const float cl = lao ? float(G * (-f.ao_exponent) / (8.0 - D)) : 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        float w[3] = {0.0f, 0.0f, 0.0f};
        uint32_t cell = kNoCell;
        if (lao && v[n] < f.ao_clamp && in_range(f, tap[n])) cell = cell_of(f, tap[n], w);
        scatter_wave(a.cell_grad, L, cell, w[0], w[1], w[2], cl);
    }
}

// blockIdx.z: 0 the primary rays, LAO and parameters, 1 the shadow rays.
__global__ void __launch_bounds__(256)
adjoint_kernel(AdjointArgs) {
    const AdjointArgs& a = kernel_args<AdjointArgs>();
    __shared__ float red[4][256 / kWave];
    __shared__ ScatterLds scatter[256 / kWave];
    adjoint_tile(a, blockIdx.x * 16, blockIdx.y * 16, red, scatter, blockIdx.z == 0 ? kWalkPrimary : kWalkShadow);
}

// ---------------------------------------------------------- batched form

struct AdjointBatchStyle {
    float* cell_grad;           // 8 floats per cell, owned by the batch
    float* d_density;
    float* d_params;
    uint32_t nx, ny, nz;
    uint32_t reserved;
};

// A contiguous range [begin, end) of one style's cell-accumulator floats
// (clear) or voxels (fold).
struct AdjointChunk {
    uint32_t style;
    uint32_t reserved;
    uint64_t begin, end;
};

struct AdjointBatchDevice {
    AdjointBatchStyle* styles;
    AdjointChunk* clear_chunks;
    AdjointChunk* fold_chunks;
    float** losses;             // per render, nullable entries
    uint32_t n_styles, n_renders;
    uint32_t n_clear_chunks, n_fold_chunks;
    uint32_t tiles_x, tiles_per_render;
};

__global__ void __launch_bounds__(256)
adjoint_batch_clear_kernel(AdjointBatchDevice d) {
    if (blockIdx.x == 0) {
        for (uint32_t s = threadIdx.x; s < 3 * d.n_styles; s += blockDim.x) d.styles[s / 3].d_params[s % 3] = 0.0f;
        for (uint32_t r = threadIdx.x; r < d.n_renders; r += blockDim.x)
            if (d.losses[r]) *d.losses[r] = 0.0f;
    }
    for (uint32_t c = blockIdx.x; c < d.n_clear_chunks; c += gridDim.x) {
        const AdjointChunk ch = d.clear_chunks[c];
        float* cg = d.styles[ch.style].cell_grad;
        for (uint64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) cg[i] = 0.0f;
    }
}

// One block per (render, 16x16 tile): the recording forward.
__global__ void __launch_bounds__(256)
adjoint_batch_forward_kernel(AdjointBatchDevice d, const AdjointArgs* __restrict__ args) {
    __shared__ uint32_t occ_lds[1];  // float volumes carry no occupancy
    const uint32_t item = uint32_t(__builtin_amdgcn_readfirstlane(int(blockIdx.x)));
    const uint32_t r = item / d.tiles_per_render, tile = item % d.tiles_per_render;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = int(tile % d.tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = int(tile / d.tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    raymarch_pixel<true>(args[r].f, px, py, OccupancyView{occ_lds, false});
}

// One block per (render, 16x16 tile): the adjoint with the L2 loss gradient.
__global__ void __launch_bounds__(256)
adjoint_batch_backward_kernel(AdjointBatchDevice d, const AdjointArgs* __restrict__ args) {
    __shared__ float red[4][256 / kWave];
    __shared__ ScatterLds scatter[256 / kWave];
    const uint32_t item = uint32_t(__builtin_amdgcn_readfirstlane(int(blockIdx.x)));
    const uint32_t r = item / d.tiles_per_render, tile = item % d.tiles_per_render;
    adjoint_tile(args[r], int(tile % d.tiles_x) * 16, int(tile / d.tiles_x) * 16, red, scatter);
}

// d_density = the fold of each style's cells (overwritten).
__global__ void __launch_bounds__(256)
adjoint_batch_fold_kernel(AdjointBatchDevice d) {
    for (uint32_t c = blockIdx.x; c < d.n_fold_chunks; c += gridDim.x) {
        const AdjointChunk ch = d.fold_chunks[c];
        const AdjointBatchStyle& st = d.styles[ch.style];
        const size_t cx = size_t(st.nx) + 1, cy = size_t(st.ny) + 1;
        for (uint64_t i = ch.begin + threadIdx.x; i < ch.end; i += blockDim.x) {
            const size_t x = i % st.nx, yz = i / st.nx, y = yz % st.ny, z = yz / st.ny;
            float sum = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const size_t cell = (x + 1 - (q & 1)) + cx * ((y + 1 - ((q >> 1) & 1)) + cy * (z + 1 - (q >> 2)));
                sum += st.cell_grad[cell * 8u + size_t(q)];
            }
            st.d_density[i] = sum;
        }
    }
}

}  // namespace
}  // namespace vkhr

using namespace vkhr;

extern "C" size_t vkhr_march_record_bytes(uint32_t width, uint32_t height) {
    return size_t(width) * height * sizeof(MarchRecord);
}

static int diff_args(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                     const vkhr_strand_params* strand, const vkhr_render_params* params, const vkhr_diff_params* diff,
                     uint32_t width, uint32_t height, RaymarchArgs* a) {
    int rc = make_raymarch_args(volume, camera, light, strand, params, width, height, a);
    if (rc) return rc;
    if (diff) {
        VKHR_REQUIRE(diff->absorption >= 0.0f && diff->absorption < 1.0f, VKHR_ERR_INVALID, "absorption %g",
                     diff->absorption);
        a->alpha = diff->absorption;
        a->adsm_k = diff->adsm_thickness;
        a->shininess = diff->shininess;
    }
    return VKHR_OK;
}

// The per-launch constants of the adjoint derived from the forward's.
static void finish_adjoint_args(AdjointArgs* a) {
    int n = 0;
    for (float t = 0.0f; t < 1.0f; t += a->f.step_size) ++n;
    a->n_steps = n;
    a->ln_transmit = std::log(1.0 - a->f.alpha);
    for (int c = 0; c < 3; ++c)
        a->light_texel[c] = (a->f.light_origin[c] - a->f.origin[c]) / a->f.size[c] * a->f.nres[c] - 0.5;
}

extern "C" int vkhr_raymarch_record(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                                    const vkhr_strand_params* strand, const vkhr_render_params* params,
                                    const vkhr_diff_params* diff, uint32_t width, uint32_t height,
                                    const float* depth_buffer, float* rgba, float* depth, void* record,
                                    vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(rgba && record, VKHR_ERR_INVALID, "rgba/record is null");
    RaymarchArgs a;
    if ((rc = diff_args(volume, camera, light, strand, params, diff, width, height, &a))) return rc;
    a.depth_buffer = depth_buffer;
    a.rgba = reinterpret_cast<float4*>(rgba);
    a.depth = depth;
    a.record = static_cast<MarchRecord*>(record);
    launch_raymarch(a, as_stream(stream));
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_raymarch_adjoint(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                                     const vkhr_strand_params* strand, const vkhr_render_params* params,
                                     const vkhr_diff_params* diff, uint32_t width, uint32_t height,
                                     const void* record, const float* d_rgba, float* d_density, float* d_params,
                                     vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(record && d_rgba && d_density && d_params, VKHR_ERR_INVALID, "null argument");
    VKHR_REQUIRE(volume && volume->density_f32, VKHR_ERR_INVALID, "the adjoint needs the float density volume");
    AdjointArgs* a = new AdjointArgs();
    std::memset(a, 0, sizeof(AdjointArgs));
    if ((rc = diff_args(volume, camera, light, strand, params, diff, width, height, &a->f))) {
        delete a;
        return rc;
    }
    a->record = static_cast<const MarchRecord*>(record);
    a->d_rgba = reinterpret_cast<const float4*>(d_rgba);
    a->d_params = d_params;
    finish_adjoint_args(a);
    const size_t n_cells = (size_t(volume->nx) + 1) * (size_t(volume->ny) + 1) * (size_t(volume->nz) + 1);
    void* cg = nullptr;
    if ((rc = stream_scratch(n_cells * 8 * sizeof(float), as_stream(stream), &cg))) {
        delete a;
        return rc;
    }
    a->cell_grad = static_cast<float*>(cg);
    hipStream_t st = as_stream(stream);
    if ((rc = fill_u32(static_cast<uint32_t*>(cg), n_cells * 8, 0u, st))) {
        delete a;
        return rc;
    }
    // each walk family in workgroups of its own: z = 0 the primary rays
    // (with LAO and parameters), z = 1 the shadow rays
    dim3 grid((width + 15) / 16, (height + 15) / 16, 2);
    adjoint_kernel<<<grid, 256, 0, st>>>(*a);
    const size_t n_vox = size_t(volume->nx) * volume->ny * volume->nz;
    gather_cells_kernel<<<grid_for(n_vox, 256, 8192), 256, 0, st>>>(a->cell_grad, d_density, int(volume->nx),
                                                                      int(volume->ny), int(volume->nz));
    delete a;
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

// ------------------------------------------------------------ batched form

struct vkhr_adjoint_batch {
    AdjointBatchDevice d;
    AdjointArgs* args = nullptr;      // device, per render
    std::vector<void*> allocations;
    uint32_t n_items = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void release_adjoint_batch(vkhr_adjoint_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void* p : b->allocations) (void)hipFree(p);
    b->allocations.clear();
}

template <class T>
int upload_table(vkhr_adjoint_batch* b, const std::vector<T>& host, T** out) {
    void* p = nullptr;
    VKHR_CHECK_HIP(hipMalloc(&p, std::max<size_t>(host.size() * sizeof(T), 16)));
    b->allocations.push_back(p);
    if (!host.empty()) VKHR_CHECK_HIP(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<T*>(p);
    return VKHR_OK;
}

int build_adjoint_batch(vkhr_adjoint_batch* b, const vkhr_adjoint_style* styles, uint32_t n_styles,
                        const vkhr_adjoint_render* renders, uint32_t n_renders, uint32_t width, uint32_t height) {
    int rc = VKHR_OK;
    std::vector<AdjointBatchStyle> st(n_styles);
    std::vector<AdjointChunk> clear, fold;
    const uint64_t kChunk = 1u << 16;
    for (uint32_t s = 0; s < n_styles; ++s) {
        const vkhr_adjoint_style& in = styles[s];
        VKHR_REQUIRE(in.volume.density_f32 && in.volume.tangents && in.d_density && in.d_params, VKHR_ERR_INVALID,
                     "style %u: the float density, tangents, d_density and d_params are required", s);
        if ((rc = check_volume(in.volume.nx, in.volume.ny, in.volume.nz, true, "adjoint style volume"))) return rc;
        const uint64_t cells = (uint64_t(in.volume.nx) + 1) * (uint64_t(in.volume.ny) + 1) * (uint64_t(in.volume.nz) + 1);
        void* cg = nullptr;
        VKHR_CHECK_HIP(hipMalloc(&cg, cells * 8 * sizeof(float)));
        b->allocations.push_back(cg);
        st[s] = {static_cast<float*>(cg), in.d_density, in.d_params, in.volume.nx, in.volume.ny, in.volume.nz, 0};
        for (uint64_t a = 0; a < cells * 8; a += kChunk) clear.push_back({s, 0, a, std::min(cells * 8, a + kChunk)});
        const uint64_t vox = uint64_t(in.volume.nx) * in.volume.ny * in.volume.nz;
        for (uint64_t a = 0; a < vox; a += kChunk) fold.push_back({s, 0, a, std::min(vox, a + kChunk)});
    }
    std::vector<AdjointArgs> args(n_renders);
    std::vector<float*> losses(n_renders, nullptr);
    for (uint32_t r = 0; r < n_renders; ++r) {
        const vkhr_adjoint_render& in = renders[r];
        VKHR_REQUIRE(in.style < n_styles, VKHR_ERR_INVALID, "render %u: style %u of %u", r, in.style, n_styles);
        VKHR_REQUIRE(in.target && in.rgba && in.record, VKHR_ERR_INVALID, "render %u: target/rgba/record is null", r);
        AdjointArgs& a = args[r];
        std::memset(&a, 0, sizeof(a));
        if ((rc = diff_args(&styles[in.style].volume, &in.camera, &in.light, &in.strand, &in.params, &in.diff, width,
                            height, &a.f)))
            return rc;
        a.f.rgba = reinterpret_cast<float4*>(in.rgba);
        a.f.record = static_cast<MarchRecord*>(in.record);
        a.record = static_cast<const MarchRecord*>(in.record);
        a.target = reinterpret_cast<const float4*>(in.target);
        a.loss = in.loss;
        a.d_params = styles[in.style].d_params;
        a.cell_grad = st[in.style].cell_grad;
        finish_adjoint_args(&a);
        losses[r] = in.loss;
    }
    AdjointBatchDevice& d = b->d;
    if ((rc = upload_table(b, st, &d.styles)) || (rc = upload_table(b, clear, &d.clear_chunks)) ||
        (rc = upload_table(b, fold, &d.fold_chunks)) || (rc = upload_table(b, losses, &d.losses)) ||
        (rc = upload_table(b, args, &b->args)))
        return rc;
    d.n_styles = n_styles;
    d.n_renders = n_renders;
    d.n_clear_chunks = uint32_t(clear.size());
    d.n_fold_chunks = uint32_t(fold.size());
    d.tiles_x = (width + 15) / 16;
    d.tiles_per_render = d.tiles_x * ((height + 15) / 16);
    b->n_items = d.tiles_per_render * n_renders;
    return VKHR_OK;
}

}  // namespace

extern "C" int vkhr_adjoint_batch_create(const vkhr_adjoint_style* styles, uint32_t n_styles,
                                         const vkhr_adjoint_render* renders, uint32_t n_renders, uint32_t width,
                                         uint32_t height, vkhr_adjoint_batch** out) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(styles && n_styles && renders && n_renders && out, VKHR_ERR_INVALID, "no styles or renders");
    VKHR_REQUIRE(width && height, VKHR_ERR_INVALID, "empty image");
    VKHR_REQUIRE(uint64_t(n_renders) * ((width + 15) / 16) * ((height + 15) / 16) < (uint64_t(1) << 31),
                 VKHR_ERR_INVALID, "too many tiles");
    vkhr_adjoint_batch* b = new vkhr_adjoint_batch();
    std::memset(&b->d, 0, sizeof(b->d));
    if ((rc = build_adjoint_batch(b, styles, n_styles, renders, n_renders, width, height))) {
        release_adjoint_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return VKHR_OK;
}

extern "C" int vkhr_adjoint_batch_launch(vkhr_adjoint_batch* b, vkhr_stream stream) {
    clear_error();
    VKHR_REQUIRE(b, VKHR_ERR_INVALID, "null batch");
    VKHR_REQUIRE(stream, VKHR_ERR_INVALID, "the batched launch is graph-captured: pass a created stream, not NULL");
    hipStream_t s = as_stream(stream);
    if (!b->exec) {
        VKHR_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        adjoint_batch_clear_kernel<<<4096, 256, 0, s>>>(b->d);
        adjoint_batch_forward_kernel<<<b->n_items, 256, 0, s>>>(b->d, b->args);
        adjoint_batch_backward_kernel<<<b->n_items, 256, 0, s>>>(b->d, b->args);
        adjoint_batch_fold_kernel<<<4096, 256, 0, s>>>(b->d);
        hipError_t e = hipStreamEndCapture(s, &b->graph);
        VKHR_CHECK_HIP(e);
        VKHR_CHECK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    VKHR_CHECK_HIP(hipGraphLaunch(b->exec, s));
    return VKHR_OK;
}

extern "C" int vkhr_adjoint_batch_destroy(vkhr_adjoint_batch* b) {
    clear_error();
    if (!b) return VKHR_OK;
    release_adjoint_batch(b);
    delete b;
    return VKHR_OK;
}
