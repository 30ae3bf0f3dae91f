// Volume raymarcher: the reference's raymarch pass (volume.frag and the GLSL
// it includes) as one gfx950 kernel, one thread per pixel.
//
// Precision plan (the result is checked against a float64 transcription):
//   * ray set-up, the surface point and all shading run in float64 (once per
//     ray);
//   * sample positions are affine in the march parameter t, so each ray keeps
//     its start/end in texel units (float64) and forms x = xs + t*(xe - xs)
//     per step in float64 -- on CDNA the FP64 FMA issues at the rate of
//     scalar FP32, and this keeps positions exact to ~1e-13 texel;
//   * the trilinear interpolation runs in float32 on the 8 texels of one
//     cell, fetched with a single 8-byte gather from the corner-packed
//     layout (vkhr_volume_finalize);
//   * the march sums (accumulated density, ADSM strand count) are float64.
// Loop counts follow the shader's float loop counters exactly.
//
// Launch: 16x16-pixel blocks of 256 threads, each wave covering an 8x8 tile
// (neighbouring rays march neighbouring texels, so gathers hit the same
// cache lines).
#include "march.hpp"

#include <cmath>
#include <cstring>

namespace vkhr {
namespace {

template <bool kF32>
__device__ __forceinline__ void raymarch_block(const RaymarchArgs& a) {
    extern __shared__ uint32_t dist_lds[];   // distance_lds_bytes(a.vol)
    const OccupancyView occ = {dist_lds, stage_distance(a.vol, dist_lds)};
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    raymarch_pixel<kF32>(a, px, py, occ);
}

// The cells volume.  Five waves per SIMD (at most 96 VGPRs): the march
// waits on its gathers, and the fifth wave hides more of that latency than
// the few spills of per-ray state it costs (0.38 -> 0.34 ms at 1080p, 512
// steps, 256^3).
template <bool kF32>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
raymarch_kernel(RaymarchArgs) {
    raymarch_block<kF32>(kernel_args<RaymarchArgs>());
}

// The float density volume (the differentiable forward), at the register
// budget the compiler picks.
__global__ void __launch_bounds__(256)
raymarch_f32_kernel(RaymarchArgs) {
    raymarch_block<true>(kernel_args<RaymarchArgs>());
}

void matmul_cm(const float* x, const float* y, double* out) {  // x * y, column-major
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += double(x[k * 4 + r]) * double(y[c * 4 + k]);
            out[c * 4 + r] = s;
        }
}

bool invert_cm(const double* m, double* inv) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][4 + c] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < 8; ++k) std::swap(a[p][k], a[c][k]);
        double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][4 + c];
    return true;
}

double smoothstep(double e0, double e1, double x) {
    double t = (x - e0) / (e1 - e0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return t * t * (3.0 - 2.0 * t);
}

}  // namespace

void launch_raymarch(const RaymarchArgs& a, hipStream_t stream) {
    dim3 grid((a.width + 15) / 16, (a.height + 15) / 16);
    const size_t lds = distance_lds_bytes(a.vol);
    if (a.vol.f32)
        raymarch_f32_kernel<<<grid, 256, lds, stream>>>(a);
    else
        raymarch_kernel<false><<<grid, 256, lds, stream>>>(a);
}

// Shared with the batched and adjoint paths: fills the per-launch arguments.
int make_raymarch_args(const vkhr_volume* vol, const vkhr_camera* cam, const vkhr_light* light,
                       const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width,
                       uint32_t height, RaymarchArgs* a) {
    VKHR_REQUIRE(vol && cam && light && sp && rp, VKHR_ERR_INVALID, "null parameter struct");
    VKHR_REQUIRE(width && height && width <= 16384 && height <= 16384, VKHR_ERR_INVALID, "bad image size %ux%u",
                 width, height);
    VKHR_REQUIRE(vol->cells || vol->density_f32, VKHR_ERR_INVALID, "volume needs cells or density_f32");
    // the march indexes cells with 24-bit multiplies (and 32-bit sums)
    int rc = check_volume(vol->nx, vol->ny, vol->nz, true, "raymarch volume");
    if (rc) return rc;
    VKHR_REQUIRE(vol->tangents || rp->shading_model != 0, VKHR_ERR_INVALID, "Kajiya-Kay needs tangents");
    VKHR_REQUIRE(rp->raycast_steps >= 1.0f && rp->raycast_steps <= 65536.0f, VKHR_ERR_INVALID,
                 "raycast_steps %g out of range", rp->raycast_steps);
    for (int c = 0; c < 3; ++c)
        VKHR_REQUIRE(sp->volume_bounds.size[c] > 0.0f, VKHR_ERR_INVALID, "flat volume bounds");
    std::memset(a, 0, sizeof(*a));
    double vp[16];
    matmul_cm(cam->projection, cam->view, vp);
    VKHR_REQUIRE(invert_cm(vp, a->ivp), VKHR_ERR_INVALID, "singular view-projection");
    for (int k = 0; k < 4; ++k) {
        a->vp_z[k] = vp[k * 4 + 2];
        a->vp_w[k] = vp[k * 4 + 3];
    }
    for (int c = 0; c < 3; ++c) {
        a->eye[c] = cam->position[c];
        a->origin[c] = sp->volume_bounds.origin[c];
        a->size[c] = sp->volume_bounds.size[c];
        a->bmin[c] = a->origin[c];
        a->bmax[c] = a->origin[c] + a->size[c];
        a->light_origin[c] = light->origin[c];
        a->hair_color[c] = sp->hair_color[c];
        a->intensity[c] = light->intensity[c];
    }
    a->nres[0] = vol->nx; a->nres[1] = vol->ny; a->nres[2] = vol->nz;
    a->nx = int(vol->nx); a->ny = int(vol->ny); a->nz = int(vol->nz);
    a->radius = sp->volume_bounds.radius;
    a->alpha = sp->hair_opacity;
    a->shininess = sp->hair_shininess;
    a->iso = rp->isosurface;
    a->lod = rp->renderer == 0 ? 0.0
             : rp->renderer == 2 ? 1.0
                                 : smoothstep(rp->lod_magnified_distance, rp->lod_minified_distance,
                                              cam->look_at_distance);
    // local_ambient_occlusion(..., kernel_size 2, radius, ...): taps at
    // +-0.5 * radius / ((2 - 1) / 2) voxels
    a->adsm_k = 11.0;  // volume.frag:78
    a->ao_scaling = double(rp->occlusion_radius) / 0.5;
    a->ao_clamp = rp->ao_clamp;
    a->ao_exponent = rp->ao_exponent;
    a->step_size = 1.0f / rp->raycast_steps;
    {
        // a power of two up to 2^20: the float loop parameter is then i / steps exactly
        const float s = rp->raycast_steps;
        int e = 0;
        const float m = std::frexp(s, &e);
        a->steps_per_unit = (m == 0.5f && s >= 1.0f && s <= 1048576.0f) ? s : 0.0f;
    }
    a->shading_model = rp->shading_model;
    a->adsm_on = rp->adsm_on;
    a->width = int(width);
    a->height = int(height);
    a->vol.cells = reinterpret_cast<const unsigned long long*>(vol->cells);
    a->vol.f32 = vol->density_f32;
    a->vol.nx = a->nx; a->vol.ny = a->ny; a->vol.nz = a->nz;
    // the distance field of the occupancy buffer, when its passes ran
    a->vol.distance = vol->occupancy && distance_field_fits(vol->nx, vol->ny, vol->nz)
                          ? vol->occupancy + occupancy_bit_words(vol->nx, vol->ny, vol->nz)
                          : nullptr;
    a->vol.bx = bricks_along(vol->nx);
    a->vol.by = bricks_along(vol->ny);
    a->vol.dist_words = uint32_t(distance_words(vol->nx, vol->ny, vol->nz));
    a->tangents = reinterpret_cast<const char4*>(vol->tangents);
    return VKHR_OK;
}

}  // namespace vkhr

using namespace vkhr;

extern "C" int vkhr_raymarch(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                             const vkhr_strand_params* strand, const vkhr_render_params* params, uint32_t width,
                             uint32_t height, const float* depth_buffer, float* rgba, float* depth,
                             int32_t* surface_info, vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(rgba || surface_info, VKHR_ERR_INVALID, "rgba is null");
    RaymarchArgs a;
    if ((rc = make_raymarch_args(volume, camera, light, strand, params, width, height, &a))) return rc;
    a.depth_buffer = depth_buffer;
    a.rgba = reinterpret_cast<float4*>(rgba);
    a.depth = depth;
    a.info = reinterpret_cast<int4*>(surface_info);
    launch_raymarch(a, as_stream(stream));
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}

extern "C" int vkhr_raymarch_ppll(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                                  const vkhr_strand_params* strand, const vkhr_render_params* params,
                                  uint32_t width, uint32_t height, const float* depth_buffer, const vkhr_ppll* ppll,
                                  vkhr_stream stream) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(ppll && ppll->heads && ppll->nodes && ppll->counter, VKHR_ERR_INVALID, "ppll is incomplete");
    VKHR_REQUIRE(ppll->width == width && ppll->height == height, VKHR_ERR_INVALID,
                 "ppll is %ux%u, target %ux%u", ppll->width, ppll->height, width, height);
    RaymarchArgs a;
    if ((rc = make_raymarch_args(volume, camera, light, strand, params, width, height, &a))) return rc;
    a.depth_buffer = depth_buffer;
    a.ppll = PpllView{ppll->heads, ppll->nodes, ppll->counter, ppll->node_count};
    launch_raymarch(a, as_stream(stream));
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
