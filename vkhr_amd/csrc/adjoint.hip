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
// the lanes of a wave are transposed so that 8 lanes add the 8 corners of one
// sample: one 32-byte line per sample instead of 8 scattered words (device
// atomics are bound by line requests).  A gather pass then folds the cells
// into the texel gradient.  Parameter gradients are reduced per block, then
// one atomic each.
#include "march.hpp"

#include <cmath>
#include <cstring>

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

// Adds c * (trilinear corner weights) of every lane's sample to the cell
// accumulator; lanes without a sample pass cell == kNoCell.  In round r the
// 8 lanes of each group of 8 add the 8 corners of the sample of lane r of
// the group.  Every lane of the wave must call it (shuffles).
__device__ __forceinline__ void scatter_wave(float* __restrict__ cg, uint32_t cell, float wx, float wy, float wz,
                                             float c) {
    const unsigned long long live = __ballot(cell != kNoCell);
    if (!live) return;
    const int lane = __lane_id();
    const int k = lane & 7, group = lane & 56;
    for (int r = 0; r < 8; ++r) {
        if (!((live >> r) & 0x0101010101010101ull)) continue;
        const int src = group | r;
        const uint32_t sc = __shfl(cell, src, kWave);
        const float sx = __shfl(wx, src, kWave), sy = __shfl(wy, src, kWave), sz = __shfl(wz, src, kWave);
        const float s_c = __shfl(c, src, kWave);
        if (sc != kNoCell) {
            const float wt = ((k & 1) ? sx : 1.0f - sx) * ((k & 2) ? sy : 1.0f - sy) * ((k & 4) ? sz : 1.0f - sz);
            unsafeAtomicAdd(&cg[size_t(sc) * 8u + size_t(k)], s_c * wt);
        }
    }
}

// One family of rays (primary or shadow): x(t) = x0 + t * dx for the
// sample parameters t of the shader's float loop, samples [0, n) of each
// active lane's ray, weight c per sample.  The loop runs to the wave's
// longest ray so that every lane takes part in the scatter.
__device__ void scatter_rays(const AdjointArgs& a, bool active, const double x0[3], const double dx[3], int n,
                             float c) {
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
            scatter_wave(a.cell_grad, cell, w[0], w[1], w[2], c);
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

// The adjoint of one pixel (volume.frag:31-102 differentiated).  Called by
// all 256 threads of a block: 2x2 waves of 8x8 pixels from (x0, y0); red is
// the block's reduction buffer for the parameter gradients.
__device__ __forceinline__ void adjoint_tile(const AdjointArgs& a, int x0, int y0, float (*red)[256 / kWave]) {
    const RaymarchArgs& f = a.f;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = x0 + (wave & 1) * 8 + (lane & 7);
    const int py = y0 + (wave >> 1) * 8 + (lane >> 3);
    const bool inside = px < f.width && py < f.height;
    const size_t pix = inside ? size_t(py) * f.width + px : 0;

    bool frag = false;
    MarchRecord r;
    float4 g = make_float4(0, 0, 0, 0);
    if (inside) {
        r = a.record[pix];
        frag = r.state == 1;
        g = a.d_rgba[pix];
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
    // parameter gradients: wave, then block, then one atomic each
    double pv[3] = {d_alpha, d_k, d_p};
    for (int k = 0; k < 3; ++k) {
        float v = float(pv[k]);
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        float v = 0.0f;
        for (int w = 0; w < int(blockDim.x / kWave); ++w) v += red[threadIdx.x][w];
        if (v != 0.0f) atomicAdd(&a.d_params[threadIdx.x], v);
    }
    __syncthreads();  // red is reused by the next tile of a persistent block

    // primary rays
    const Ray ray = setup_ray(f, px, py, inside && frag && c_primary != 0.0);
    scatter_rays(a, ray.hit && c_primary != 0.0, ray.xs, ray.dx, frag ? r.n_primary : 0, float(c_primary));
    // shadow rays: from the surface point to the light origin, every sample
    double s[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    if (frag) {
        s[0] = r.sx; s[1] = r.sy; s[2] = r.sz;
        for (int k = 0; k < 3; ++k) sl[k] = a.light_texel[k] - s[k];
    }
    // samples past the volume's support carry no gradient: the walk stops
    // there (an upper bound on their count is enough, in_range does the rest)
    int n_shadow = 0;
    if (frag) {
        double lo, hi;
        support_interval(f, s, sl, &lo, &hi);
        if (hi >= 0.0 && lo <= hi) n_shadow = min(a.n_steps, int(fmin(hi, 2.0) / double(f.step_size) * 1.0001) + 2);
    }
    scatter_rays(a, frag && adsm_on && c_shadow != 0.0, s, sl, n_shadow, float(c_shadow));
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
        scatter_wave(a.cell_grad, cell, w[0], w[1], w[2], cl);
    }
}

__global__ void __launch_bounds__(256)
adjoint_kernel(AdjointArgs a) {
    __shared__ float red[3][256 / kWave];
    adjoint_tile(a, blockIdx.x * 16, blockIdx.y * 16, red);
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
    int n = 0;
    for (float t = 0.0f; t < 1.0f; t += a->f.step_size) ++n;
    a->n_steps = n;
    a->ln_transmit = std::log(1.0 - a->f.alpha);
    for (int c = 0; c < 3; ++c)
        a->light_texel[c] = (a->f.light_origin[c] - a->f.origin[c]) / a->f.size[c] * a->f.nres[c] - 0.5;
    const size_t n_cells = size_t(volume->nx + 1) * (volume->ny + 1) * (volume->nz + 1);
    if (n_cells >= (size_t(1) << 32)) {
        delete a;
        set_error("volume %ux%ux%u too large for the adjoint", volume->nx, volume->ny, volume->nz);
        return VKHR_ERR_INVALID;
    }
    void* cg = nullptr;
    if ((rc = scratch(n_cells * 8 * sizeof(float), &cg))) {
        delete a;
        return rc;
    }
    a->cell_grad = static_cast<float*>(cg);
    hipStream_t st = as_stream(stream);
    if ((rc = fill_u32(static_cast<uint32_t*>(cg), n_cells * 8, 0u, st))) {
        delete a;
        return rc;
    }
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    adjoint_kernel<<<grid, 256, 0, st>>>(*a);
    const size_t n_vox = size_t(volume->nx) * volume->ny * volume->nz;
    gather_cells_kernel<<<grid_for(n_vox, 256, 8192), 256, 0, st>>>(a->cell_grad, d_density, int(volume->nx),
                                                                      int(volume->ny), int(volume->nz));
    delete a;
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
