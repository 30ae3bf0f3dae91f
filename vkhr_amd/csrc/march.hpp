// The raymarch pass's per-launch arguments, its ray set-up and the per-pixel
// record the differentiable forward leaves for the adjoint.  Shared by the
// forward (raymarch.hip), the adjoint (adjoint.hip) and the batched launch.
#pragma once

#include "render.hpp"

namespace vkhr {

// What the adjoint needs of one pixel's forward pass (64 bytes).  Sample
// positions are not stored: the adjoint recomputes them from the ray set-up
// at every 32-step chunk.
struct MarchRecord {
    double sx, sy, sz;     // surface point, texel coordinates
    float acc;             // accumulated density of the primary march
    float shadow_sum;      // sum of densities along the shadow ray (before x thickness)
    float adsm, lao;       // the two occlusion factors
    float shade[3];        // shading (Kajiya-Kay or 1)
    float spec_base;       // Kajiya-Kay specular base max(cTL cTE + sTL sTE, 0)
    int n_primary;         // samples taken by the primary march (depth-buffer break)
    int state;             // 0 miss, 1 fragment, 2 marched and discarded
};
static_assert(sizeof(MarchRecord) == 64, "record layout");

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
    double adsm_k;             // ADSM strand thickness per unit density (11 in the shader)
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
    MarchRecord* record;
    PpllView ppll;             // node insertion when ppll.heads is set
};

// Texel-space coordinate of a world point along one axis: uvw * n - 0.5.
__device__ __forceinline__ double to_texel(double p, double o, double s, double n) {
    return (p - o) / s * n - 0.5;
}

struct Ray {
    double start[3], end[3];   // world
    double xs[3], dx[3];       // texel coordinates: x(t) = xs + t * dx
    double zs, ws, dz, dw;     // clip z and w along the ray (affine in t)
    bool hit;
};

// The pixel-centre ray and its AABB entry (the proxy cube fragment):
// volume.frag:32-34 with the proxy of volume.cc:60-203.
__device__ __forceinline__ Ray setup_ray(const RaymarchArgs& a, int px, int py, bool inside) {
    Ray r;
    double nxd = 2.0 * (px + 0.5) / a.width - 1.0, nyd = 2.0 * (py + 0.5) / a.height - 1.0;
    double far4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        far4[k] = a.ivp[0 * 4 + k] * nxd + a.ivp[1 * 4 + k] * nyd + a.ivp[2 * 4 + k] + a.ivp[3 * 4 + k];
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
        t0 = fmax(t0, fmin(u, v));
        t1 = fmin(t1, fmax(u, v));
    }
    r.hit = inside && hit && t0 <= t1 && t0 > 0.0;
    r.zs = 0; r.ws = 1; r.dz = 0; r.dw = 0;
    if (r.hit) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r.start[c] = a.eye[c] + dir[c] * t0;
        r.zs = a.vp_z[0] * r.start[0] + a.vp_z[1] * r.start[1] + a.vp_z[2] * r.start[2] + a.vp_z[3];
        r.ws = a.vp_w[0] * r.start[0] + a.vp_w[1] * r.start[1] + a.vp_w[2] * r.start[2] + a.vp_w[3];
        double zn = r.zs / r.ws;
        r.hit = zn >= 0.0 && zn <= 1.0;
    }
    if (r.hit) {
        double sd[3] = {r.start[0] - a.eye[0], r.start[1] - a.eye[1], r.start[2] - a.eye[2]};
        normalize3(sd);
#pragma unroll
        for (int c = 0; c < 3; ++c) r.end[c] = r.start[c] + sd[c] * a.radius;
        double ze = a.vp_z[0] * r.end[0] + a.vp_z[1] * r.end[1] + a.vp_z[2] * r.end[2] + a.vp_z[3];
        double we = a.vp_w[0] * r.end[0] + a.vp_w[1] * r.end[1] + a.vp_w[2] * r.end[2] + a.vp_w[3];
        r.dz = ze - r.zs;
        r.dw = we - r.ws;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r.xs[c] = to_texel(r.start[c], a.origin[c], a.size[c], a.nres[c]);
            r.dx[c] = to_texel(r.end[c], a.origin[c], a.size[c], a.nres[c]) - r.xs[c];
        }
    }
    return r;
}

void launch_raymarch(const RaymarchArgs& a, hipStream_t stream);

int make_raymarch_args(const vkhr_volume* vol, const vkhr_camera* cam, const vkhr_light* light,
                       const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width, uint32_t height,
                       RaymarchArgs* a);

}  // namespace vkhr
