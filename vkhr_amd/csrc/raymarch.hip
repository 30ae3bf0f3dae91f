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
#include "render.hpp"

#include <cmath>
#include <cstring>

namespace vkhr {
namespace {

struct RaymarchArgs {
    double ivp[16];            // inverse(projection * view), column-major
    double vp_z[4], vp_w[4];   // rows 2 and 3 of projection * view
    double eye[3];
    double bmin[3], bmax[3];
    double origin[3], size[3];
    double nres[3];            // volume resolution
    double radius;
    double light_origin[3];
    double hair_color[3], intensity[3];
    double alpha, shininess, iso, lod;
    double ao_scaling;         // LAO tap offset in texels per kernel unit
    double ao_clamp, ao_exponent;
    float step_size;           // 1 / raycast_steps, in float as the shader
    int shading_model, adsm_on;
    int nx, ny, nz;
    int width, height;
    DensityVolume vol;
    const char4* tangents;
    const float* depth_buffer;
    float4* rgba;
    float* depth;
    int4* info;
    PpllView ppll;             // node insertion when ppll.heads is set
};

// Texel-space coordinate of a world point along one axis: uvw * n - 0.5.
__device__ __forceinline__ double to_texel(double p, double o, double s, double n) {
    return (p - o) / s * n - 0.5;
}

struct Texel3 {
    double x, y, z;
};

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

    // --- the pixel-centre ray and its AABB entry (the proxy cube fragment)
    double nxd = 2.0 * (px + 0.5) / a.width - 1.0, nyd = 2.0 * (py + 0.5) / a.height - 1.0;
    double far4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        far4[r] = a.ivp[0 * 4 + r] * nxd + a.ivp[1 * 4 + r] * nyd + a.ivp[2 * 4 + r] + a.ivp[3 * 4 + r];
    double dir[3] = {far4[0] / far4[3] - a.eye[0], far4[1] / far4[3] - a.eye[1], far4[2] / far4[3] - a.eye[2]};
    normalize3(dir);
    double t0 = -1e300, t1 = 1e300;
    bool hit = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (dir[c] == 0.0) {
            if (a.eye[c] < a.bmin[c] || a.eye[c] > a.bmax[c]) hit = false;
            continue;
        }
        double u = (a.bmin[c] - a.eye[c]) / dir[c], v = (a.bmax[c] - a.eye[c]) / dir[c];
        double lo = fmin(u, v), hi = fmax(u, v);
        t0 = fmax(t0, lo);
        t1 = fmin(t1, hi);
    }
    hit = inside && hit && t0 <= t1 && t0 > 0.0;
    double start[3], end[3];
    double zs = 0, ws = 1, dz = 0, dw = 0;
    if (hit) {
#pragma unroll
        for (int c = 0; c < 3; ++c) start[c] = a.eye[c] + dir[c] * t0;
        zs = a.vp_z[0] * start[0] + a.vp_z[1] * start[1] + a.vp_z[2] * start[2] + a.vp_z[3];
        ws = a.vp_w[0] * start[0] + a.vp_w[1] * start[1] + a.vp_w[2] * start[2] + a.vp_w[3];
        double zn = zs / ws;
        hit = zn >= 0.0 && zn <= 1.0;
    }
    if (hit) {
        // end = start + normalize(start - eye) * radius (volume.frag:32-34)
        double sd[3] = {start[0] - a.eye[0], start[1] - a.eye[1], start[2] - a.eye[2]};
        normalize3(sd);
#pragma unroll
        for (int c = 0; c < 3; ++c) end[c] = start[c] + sd[c] * a.radius;
        double ze = a.vp_z[0] * end[0] + a.vp_z[1] * end[1] + a.vp_z[2] * end[2] + a.vp_z[3];
        double we = a.vp_w[0] * end[0] + a.vp_w[1] * end[1] + a.vp_w[2] * end[2] + a.vp_w[3];
        dz = ze - zs;
        dw = we - ws;
        const double db = a.depth_buffer ? double(a.depth_buffer[pix]) : 1.0;

        double xs[3], dx[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xs[c] = to_texel(start[c], a.origin[c], a.size[c], a.nres[c]);
            dx[c] = to_texel(end[c], a.origin[c], a.size[c], a.nres[c]) - xs[c];
        }

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
            double l[3] = {a.light_origin[0] - sp[0], a.light_origin[1] - sp[1], a.light_origin[2] - sp[2]};
            double e[3] = {sp[0] - a.eye[0], sp[1] - a.eye[1], sp[2] - a.eye[2]};
            normalize3(l);
            normalize3(e);
            if (a.shading_model == 0) {
                double tg[3];
                tangent_at(a, sx[0], sx[1], sx[2], tg);
                normalize3(tg);
                kajiya_kay(a.hair_color, a.intensity, a.shininess, tg, l, e, shading);
            }
            double occlusion = 1.0;
            if (a.adsm_on == 1 && a.shading_model != 3) {
                double xl[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) xl[c] = to_texel(a.light_origin[c], a.origin[c], a.size[c], a.nres[c]) - sx[c];
                double strands = 0.0;
                for (float t = 0.0f; t < 1.0f; t += a.step_size) {
                    const double td = double(t);
                    strands += double(density_at(a.vol, fma(td, xl[0], sx[0]), fma(td, xl[1], sx[1]),
                                                 fma(td, xl[2], sx[2])));
                }
                occlusion *= pow(1.0 - a.alpha, strands * 11.0);
            }
            if (a.shading_model != 2)
                occlusion *= local_ambient_occlusion(a.vol, sx, a.ao_scaling, a.ao_clamp, a.ao_exponent);
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
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    raymarch_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
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
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    raymarch_kernel<<<grid, 256, 0, as_stream(stream)>>>(a);
    VKHR_CHECK_LAUNCH();
    return VKHR_OK;
}
