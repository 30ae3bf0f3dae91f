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

// Trilinear SNORM tangent (xyz) at a texel coordinate.
__device__ __forceinline__ void tangent_at(const RaymarchArgs& a, double x, double y, double z, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (!(x >= -1.0 && y >= -1.0 && z >= -1.0 && x < double(a.nx) && y < double(a.ny) && z < double(a.nz)))
        return;
    double fx = floor(x), fy = floor(y), fz = floor(z);
    double w[3] = {x - fx, y - fy, z - fz};
    int ix = int(fx), iy = int(fy), iz = int(fz);
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        int tx = ix + (n & 1), ty = iy + ((n >> 1) & 1), tz = iz + (n >> 2);
        if (tx < 0 || ty < 0 || tz < 0 || tx >= a.nx || ty >= a.ny || tz >= a.nz) continue;
        char4 t = a.tangents[size_t(tx) + size_t(a.nx) * (size_t(ty) + size_t(a.ny) * size_t(tz))];
        double wt = ((n & 1) ? w[0] : 1.0 - w[0]) * (((n >> 1) & 1) ? w[1] : 1.0 - w[1]) *
                    ((n >> 2) ? w[2] : 1.0 - w[2]);
        out[0] += wt * fmax(double(t.x) / 127.0, -1.0);
        out[1] += wt * fmax(double(t.y) / 127.0, -1.0);
        out[2] += wt * fmax(double(t.z) / 127.0, -1.0);
    }
}

__global__ void __launch_bounds__(256)
raymarch_kernel(RaymarchArgs a) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    // Out-of-image lanes stay converged (the PPLL insert ballots the wave).
    const bool inside = px < a.width && py < a.height;
    const size_t pix = inside ? size_t(py) * a.width + px : 0;

    float4 color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float out_depth = 1.0f;
    int4 info = make_int4(0, -1, -1, 0);

    const Ray ray = setup_ray(a, px, py, inside);
    if (ray.hit) {
        const double* xs = ray.xs;
        const double* dx = ray.dx;
        const double zs = ray.zs, ws = ray.ws, dz = ray.dz, dw = ray.dw;
        const double* start = ray.start;
        const double* end = ray.end;
        const double db = a.depth_buffer ? double(a.depth_buffer[pix]) : 1.0;

        // --- volume_surface: isosurface march, no early exit
        double acc = 0.0;
        float t_surf = 0.0f, t_entry = 0.0f;
        int surf = -1, entry = -1;
        bool found = false;
        int i = 0;
        for (float t = 0.0f; t < 1.0f; t += a.step_size, ++i) {
            const double td = double(t);
            // depth test: db < z/w  <=>  db*w < z  (w > 0 in front of the eye)
            if (db * fma(td, dw, ws) < fma(td, dz, zs)) break;
            float d = density_at(a.vol, fma(td, dx[0], xs[0]), fma(td, dx[1], xs[1]), fma(td, dx[2], xs[2]));
            acc += double(d);
            if (d != 0.0f) {
                if (acc <= a.iso) { surf = i; t_surf = t; }
                if (acc >= a.iso) found = true;
                if (entry < 0) { entry = i; t_entry = t; }
            }
        }
        info = make_int4(2, surf, entry, found ? 1 : 0);  // marched; 1 once a fragment is written
        if (a.record) {
            a.record[pix].acc = float(acc);
            a.record[pix].n_primary = i;
        }
        double sp[3] = {0.0, 0.0, 0.0};  // the shader's vec3(0) surface
        if (found ? (surf >= 0) : (entry >= 0)) {
            double ts = double(found ? t_surf : t_entry);
#pragma unroll
            for (int c = 0; c < 3; ++c) sp[c] = start[c] * (1.0 - ts) + end[c] * ts;
        }
        const double a_surf = acc / a.iso;
        if (a_surf != 0.0) {
            info.x = 1;
            const double coverage = a.lod * a_surf * a.alpha;
            double sx[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) sx[c] = to_texel(sp[c], a.origin[c], a.size[c], a.nres[c]);

            double shading[3] = {1.0, 1.0, 1.0};
            double spec_base = 0.0, strands = 0.0, adsm = 1.0, lao = 1.0;
            double l[3] = {a.light_origin[0] - sp[0], a.light_origin[1] - sp[1], a.light_origin[2] - sp[2]};
            double e[3] = {sp[0] - a.eye[0], sp[1] - a.eye[1], sp[2] - a.eye[2]};
            normalize3(l);
            normalize3(e);
            if (a.shading_model == 0) {
                double tg[3];
                tangent_at(a, sx[0], sx[1], sx[2], tg);
                normalize3(tg);
                kajiya_kay(a.hair_color, a.intensity, a.shininess, tg, l, e, shading, &spec_base);
            }
            double occlusion = 1.0;
            if (a.adsm_on == 1 && a.shading_model != 3) {
                double xl[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) xl[c] = to_texel(a.light_origin[c], a.origin[c], a.size[c], a.nres[c]) - sx[c];
                for (float t = 0.0f; t < 1.0f; t += a.step_size) {
                    const double td = double(t);
                    strands += double(density_at(a.vol, fma(td, xl[0], sx[0]), fma(td, xl[1], sx[1]),
                                                 fma(td, xl[2], sx[2])));
                }
                adsm = pow(1.0 - a.alpha, strands * a.adsm_k);
                occlusion *= adsm;
            }
            if (a.shading_model != 2) {
                lao = local_ambient_occlusion(a.vol, sx, a.ao_scaling, a.ao_clamp, a.ao_exponent);
                occlusion *= lao;
            }
            if (a.record) {
                MarchRecord& r = a.record[pix];
                r.sx = sx[0]; r.sy = sx[1]; r.sz = sx[2];
                r.shadow_sum = float(strands);
                r.adsm = float(adsm);
                r.lao = float(lao);
                r.shade[0] = float(shading[0]); r.shade[1] = float(shading[1]); r.shade[2] = float(shading[2]);
                r.spec_base = float(spec_base);
            }
            double zz = a.vp_z[0] * sp[0] + a.vp_z[1] * sp[1] + a.vp_z[2] * sp[2] + a.vp_z[3];
            double ww = a.vp_w[0] * sp[0] + a.vp_w[1] * sp[1] + a.vp_w[2] * sp[2] + a.vp_w[3];
            out_depth = float(zz / ww);
            color = make_float4(float(fmin(fmax(shading[0] * occlusion, 0.0), 1.0)),
                                float(fmin(fmax(shading[1] * occlusion, 0.0), 1.0)),
                                float(fmin(fmax(shading[2] * occlusion, 0.0), 1.0)),
                                float(fmin(fmax(coverage, 0.0), 1.0)));
        }
    }
    if (a.ppll.heads) ppll_insert(a.ppll, info.x == 1, uint32_t(pix), color, out_depth);
    if (!inside) return;
    if (a.record) a.record[pix].state = info.x;
    if (a.rgba) a.rgba[pix] = color;
    if (a.depth) a.depth[pix] = out_depth;
    if (a.info) a.info[pix] = info;
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
    raymarch_kernel<<<grid, 256, 0, stream>>>(a);
}

// Shared with the batched and adjoint paths: fills the per-launch arguments.
int make_raymarch_args(const vkhr_volume* vol, const vkhr_camera* cam, const vkhr_light* light,
                       const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width,
                       uint32_t height, RaymarchArgs* a) {
    VKHR_REQUIRE(vol && cam && light && sp && rp, VKHR_ERR_INVALID, "null parameter struct");
    VKHR_REQUIRE(width && height && width <= 16384 && height <= 16384, VKHR_ERR_INVALID, "bad image size %ux%u",
                 width, height);
    VKHR_REQUIRE(vol->nx && vol->ny && vol->nz, VKHR_ERR_INVALID, "zero volume resolution");
    VKHR_REQUIRE(vol->cells || vol->density_f32, VKHR_ERR_INVALID, "volume needs cells or density_f32");
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
    a->shading_model = rp->shading_model;
    a->adsm_on = rp->adsm_on;
    a->width = int(width);
    a->height = int(height);
    a->vol.cells = reinterpret_cast<const unsigned long long*>(vol->cells);
    a->vol.f32 = vol->density_f32;
    a->vol.nx = a->nx; a->vol.ny = a->ny; a->vol.nz = a->nz;
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
