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
// One block owns a 16x16-pixel tile.  It re-walks its rays in chunks of 32
// steps: the march state (t, sample index) is carried from chunk to chunk and
// the positions are recomputed from the ray set-up (nothing per sample is
// stored).  Per chunk the block reduces the bounding box of its samples'
// trilinear footprints; if it fits in the LDS window, the scatter of
// weight * trilinear weight goes to LDS atomics and the window is flushed with
// one global atomic per touched texel, otherwise straight to global atomics.
// LAO taps (8 per fragment) scatter directly.  Parameter gradients are
// reduced per block, then one atomic each.
#include "march.hpp"

#include <cmath>
#include <cstring>

namespace vkhr {
namespace {

constexpr int kChunk = 32;
constexpr int kWindow = 8192;  // floats of LDS per block (32 KiB)

struct AdjointArgs {
    RaymarchArgs f;
    const MarchRecord* record;
    const float4* d_rgba;
    float* d_density;
    float* d_params;            // absorption, ADSM thickness, shininess
    int n_steps;                // iterations of the shader's float loop
    double ln_transmit;         // log(1 - alpha)
    double light_texel[3];
};

__device__ __forceinline__ size_t vidx(const RaymarchArgs& f, int x, int y, int z) {
    return size_t(x) + size_t(f.nx) * (size_t(y) + size_t(f.ny) * size_t(z));
}

// Scatter c * (trilinear weights at x) to global memory.
__device__ __forceinline__ void scatter_global(const AdjointArgs& a, const double x[3], float c) {
    const RaymarchArgs& f = a.f;
    double fx = floor(x[0]), fy = floor(x[1]), fz = floor(x[2]);
    float w[3] = {float(x[0] - fx), float(x[1] - fy), float(x[2] - fz)};
    int ix = int(fx), iy = int(fy), iz = int(fz);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        int tx = ix + (n & 1), ty = iy + ((n >> 1) & 1), tz = iz + (n >> 2);
        if (tx < 0 || ty < 0 || tz < 0 || tx >= f.nx || ty >= f.ny || tz >= f.nz) continue;
        float wt = ((n & 1) ? w[0] : 1.0f - w[0]) * (((n >> 1) & 1) ? w[1] : 1.0f - w[1]) *
                   ((n >> 2) ? w[2] : 1.0f - w[2]);
        unsafeAtomicAdd(&a.d_density[vidx(f, tx, ty, tz)], c * wt);
    }
}

__device__ __forceinline__ bool in_range(const RaymarchArgs& f, const double x[3]) {
    return x[0] >= -1.0 && x[1] >= -1.0 && x[2] >= -1.0 && x[0] < double(f.nx) && x[1] < double(f.ny) &&
           x[2] < double(f.nz);
}

// One family of rays (primary or shadow): x(t) = x0 + t * dx for the
// sample parameters t of the shader's float loop, samples [0, n) of each ray,
// weight c per sample.  All threads of the block call this (barriers).
__device__ void scatter_rays(const AdjointArgs& a, float* win, int* bb, bool active, const double x0[3],
                             const double dx[3], int n, float c) {
    const RaymarchArgs& f = a.f;
    const float step = f.step_size;
    float t = 0.0f;  // the march state carried across chunks
    const int n_chunks = (a.n_steps + kChunk - 1) / kChunk;
    for (int m = 0; m < n_chunks; ++m) {
        const int i0 = m * kChunk;
        const int i1 = min(i0 + kChunk, n);
        const bool live = active && i0 < i1;
        // footprint of this chunk: the segment between its first and last samples
        int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
        if (live) {
            float tl = t;
            for (int i = i0 + 1; i < i1; ++i) tl += step;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double p = fma(double(t), dx[k], x0[k]), q = fma(double(tl), dx[k], x0[k]);
                double mn = fmin(p, q), mx = fmax(p, q);
                int l = int(floor(fmax(mn, -1.0))), h = int(floor(fmin(mx, 1e9))) + 1;
                int nk = k == 0 ? f.nx : (k == 1 ? f.ny : f.nz);
                lo[k] = max(l, 0);
                hi[k] = min(h, nk - 1);
            }
            if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) lo[0] = 0x7fffffff;
        }
        if (threadIdx.x == 0) {
            bb[0] = bb[1] = bb[2] = 0x7fffffff;
            bb[3] = bb[4] = bb[5] = -1;
        }
        __syncthreads();
        if (live && lo[0] != 0x7fffffff) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(&bb[k], lo[k]);
                atomicMax(&bb[3 + k], hi[k]);
            }
        }
        __syncthreads();
        const int wx0 = bb[0], wy0 = bb[1], wz0 = bb[2];
        const int dimx = bb[3] - wx0 + 1, dimy = bb[4] - wy0 + 1, dimz = bb[5] - wz0 + 1;
        const long long vol = (bb[3] < 0) ? 0 : (long long)dimx * dimy * dimz;
        const bool use_lds = vol > 0 && vol <= kWindow;
        if (use_lds) {
            for (int j = threadIdx.x; j < vol; j += blockDim.x) win[j] = 0.0f;
            __syncthreads();
        }
        if (live) {
            float tt = t;
            for (int i = i0; i < i1; ++i, tt += step) {
                double x[3] = {fma(double(tt), dx[0], x0[0]), fma(double(tt), dx[1], x0[1]),
                               fma(double(tt), dx[2], x0[2])};
                if (!in_range(f, x)) continue;
                if (!use_lds) {
                    scatter_global(a, x, c);
                    continue;
                }
                double fx = floor(x[0]), fy = floor(x[1]), fz = floor(x[2]);
                float w[3] = {float(x[0] - fx), float(x[1] - fy), float(x[2] - fz)};
                int ix = int(fx), iy = int(fy), iz = int(fz);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    int tx = ix + (q & 1), ty = iy + ((q >> 1) & 1), tz = iz + (q >> 2);
                    if (tx < 0 || ty < 0 || tz < 0 || tx >= f.nx || ty >= f.ny || tz >= f.nz) continue;
                    float wt = ((q & 1) ? w[0] : 1.0f - w[0]) * (((q >> 1) & 1) ? w[1] : 1.0f - w[1]) *
                               ((q >> 2) ? w[2] : 1.0f - w[2]);
                    atomicAdd(&win[(tx - wx0) + dimx * ((ty - wy0) + dimy * (tz - wz0))], c * wt);
                }
            }
        }
        if (use_lds) {
            __syncthreads();
            for (int j = threadIdx.x; j < vol; j += blockDim.x) {
                float v = win[j];
                if (v != 0.0f) {
                    int lx = j % dimx, ly = (j / dimx) % dimy, lz = j / (dimx * dimy);
                    unsafeAtomicAdd(&a.d_density[vidx(f, wx0 + lx, wy0 + ly, wz0 + lz)], v);
                }
            }
        }
        __syncthreads();
        // advance the carried state past this chunk
        for (int i = 0; i < kChunk; ++i) t += step;
    }
}

__global__ void __launch_bounds__(256)
adjoint_kernel(AdjointArgs a) {
    __shared__ float win[kWindow];
    __shared__ int bb[6];
    __shared__ float red[3][256 / kWave];
    const RaymarchArgs& f = a.f;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
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

    // primary rays
    const Ray ray = setup_ray(f, px, py, inside && frag && c_primary != 0.0);
    scatter_rays(a, win, bb, ray.hit && c_primary != 0.0, ray.xs, ray.dx, frag ? r.n_primary : 0,
                 float(c_primary));
    // shadow rays: from the surface point to the light origin, every sample
    double s[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    if (frag) {
        s[0] = r.sx; s[1] = r.sy; s[2] = r.sz;
        for (int k = 0; k < 3; ++k) sl[k] = a.light_texel[k] - s[k];
    }
    scatter_rays(a, win, bb, frag && adsm_on && c_shadow != 0.0, s, sl, a.n_steps, float(c_shadow));
    // LAO taps: d ln(lao)/dD = -e / (8 - D), through min(v, clamp)
    if (frag && lao_on && G != 0.0) {
        double v[8], D = 0.0;
        double tap[8][3];
        for (int n = 0; n < 8; ++n) {
            tap[n][0] = s[0] + ((n & 1) ? 0.5 : -0.5) * f.ao_scaling;
            tap[n][1] = s[1] + (((n >> 1) & 1) ? 0.5 : -0.5) * f.ao_scaling;
            tap[n][2] = s[2] + ((n >> 2) ? 0.5 : -0.5) * f.ao_scaling;
            v[n] = double(density_at(f.vol, tap[n][0], tap[n][1], tap[n][2]));
            D += fmin(v[n], f.ao_clamp);
        }
        double cl = G * (-f.ao_exponent) / (8.0 - D);
        for (int n = 0; n < 8; ++n)
            if (v[n] < f.ao_clamp && in_range(f, tap[n])) scatter_global(a, tap[n], float(cl));
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
    a->d_density = d_density;
    a->d_params = d_params;
    int n = 0;
    for (float t = 0.0f; t < 1.0f; t += a->f.step_size) ++n;
    a->n_steps = n;
    a->ln_transmit = std::log(1.0 - a->f.alpha);
    for (int c = 0; c < 3; ++c)
        a->light_texel[c] = (a->f.light_origin[c] - a->f.origin[c]) / a->f.size[c] * a->f.nres[c] - 0.5;
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    adjoint_kernel<<<grid, 256, 0, as_stream(stream)>>>(*a);
    delete a;
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
