// ORACLE -- test infrastructure only.  Float64 transcription of the
// reference's volume raymarch pass, used by tests/ and bench.py's CPU
// baseline as the checker; never linked into the product library.
//
// Transcribed (paths under share/shaders/):
//   volumes/volume.frag:31-102              main: ray set-up, shading, occlusion, depth
//   volumes/volume_rendering.glsl:19-52     volume_surface (isosurface march, no early exit)
//   volumes/sample_volume.glsl:7-9          sample_volume
//   self-shadowing/approximate_deep_shadows.glsl:24-36   volume ADSM
//   volumes/local_ambient_occlusion.glsl:9-30            LAO (kernel 2 -> 8 taps)
//   shading/kajiya-kay.glsl:7-24            Kajiya-Kay
//   level_of_detail/scheme.glsl:7-11        lod()
// Fixed-function parts made explicit:
//   * the proxy cube (src/vkhr/rasterizer/volume.cc:60-203, CW front faces,
//     back faces culled) yields one fragment per pixel whose centre ray
//     enters the AABB in front of the eye; its interpolated world position
//     is the ray's entry point (eye inside the box: no fragment);
//   * texture(): Vulkan trilinear filtering with normalized coordinates,
//     texel centres at (i + 0.5) / n, CLAMP_TO_BORDER with a zero border
//     (rasterizer/hair_style.cc:79-85, 112-118; vkpp/sampler.cc:51), exact
//     weights; R8_UNORM texels read c / 255, R8G8B8A8_SNORM texels read
//     max(c / 127, -1) (vkpp/image.cc:453-577);
//   * GLSL float loop counters (t += 1/steps) are kept in float, so the
//     number of samples is the shader's; everything else is float64.
// Defined where the shader is not: normalize(0) = 0, and sqrt of a negative
// rounding residue in Kajiya-Kay is 0 (the GPU build does the same).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../include/vkhr_hip.h"

namespace {

struct d3 {
    double x, y, z;
};
inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline d3 operator*(d3 a, d3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline d3 operator/(d3 a, d3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 normalize(d3 v) {
    double l = std::sqrt(dot(v, v));
    return l > 0.0 ? v * (1.0 / l) : d3{0.0, 0.0, 0.0};
}
inline d3 mix(d3 a, d3 b, double t) { return a * (1.0 - t) + b * t; }
inline d3 F3(const float* p) { return {p[0], p[1], p[2]}; }

struct Volume {
    const uint8_t* density;
    const int8_t* tangent;
    const float* density_f32;
    int nx, ny, nz;
    d3 origin, size;
};

// Column-major 4x4 (m[col*4 + row]) times (p, 1).
inline void xform(const double* m, d3 p, double out[4]) {
    for (int r = 0; r < 4; ++r) out[r] = m[0 * 4 + r] * p.x + m[1 * 4 + r] * p.y + m[2 * 4 + r] * p.z + m[3 * 4 + r];
}

void matmul(const float* a, const float* b, double* out) {  // a * b, column-major
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += double(a[k * 4 + r]) * double(b[c * 4 + k]);
            out[c * 4 + r] = s;
        }
}

bool invert(const double* m, double* inv) {  // Gauss-Jordan, column-major in/out
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

// texture(sampler3D, uvw).r for the density (UNORM or float volume).
double texel_density(const Volume& v, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= v.nx || y >= v.ny || z >= v.nz) return 0.0;
    size_t i = size_t(x) + size_t(v.nx) * (size_t(y) + size_t(v.ny) * size_t(z));
    return v.density_f32 ? double(v.density_f32[i]) : double(v.density[i]) / 255.0;
}

double texel_tangent(const Volume& v, int x, int y, int z, int c) {
    if (x < 0 || y < 0 || z < 0 || x >= v.nx || y >= v.ny || z >= v.nz) return 0.0;
    size_t i = size_t(x) + size_t(v.nx) * (size_t(y) + size_t(v.ny) * size_t(z));
    double t = double(v.tangent[4 * i + c]) / 127.0;
    return t < -1.0 ? -1.0 : t;
}

template <class F>
double trilinear(const Volume& v, d3 uvw, F fetch) {
    double x = uvw.x * v.nx - 0.5, y = uvw.y * v.ny - 0.5, z = uvw.z * v.nz - 0.5;
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int i = int(fx), j = int(fy), k = int(fz);
    double a = x - fx, b = y - fy, c = z - fz;
    double r = 0.0;
    for (int n = 0; n < 8; ++n) {
        int dx = n & 1, dy = (n >> 1) & 1, dz = n >> 2;
        double w = (dx ? a : 1.0 - a) * (dy ? b : 1.0 - b) * (dz ? c : 1.0 - c);
        if (w != 0.0) r += w * fetch(i + dx, j + dy, k + dz);
    }
    return r;
}

// sample_volume(volume, P, origin, size): texture((P - origin) / size)
double sample_density(const Volume& v, d3 p) {
    d3 uvw = (p - v.origin) / v.size;
    if (!(std::fabs(uvw.x) < 1e6 && std::fabs(uvw.y) < 1e6 && std::fabs(uvw.z) < 1e6)) return 0.0;
    return trilinear(v, uvw, [&](int x, int y, int z) { return texel_density(v, x, y, z); });
}

d3 sample_tangent(const Volume& v, d3 p) {
    d3 uvw = (p - v.origin) / v.size;
    d3 r;
    r.x = trilinear(v, uvw, [&](int x, int y, int z) { return texel_tangent(v, x, y, z, 0); });
    r.y = trilinear(v, uvw, [&](int x, int y, int z) { return texel_tangent(v, x, y, z, 1); });
    r.z = trilinear(v, uvw, [&](int x, int y, int z) { return texel_tangent(v, x, y, z, 2); });
    return r;
}

double smoothstep(double e0, double e1, double x) {
    double t = (x - e0) / (e1 - e0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return t * t * (3.0 - 2.0 * t);
}

double lod(const vkhr_render_params& p, double distance) {
    if (p.renderer == 0) return 0.0;
    if (p.renderer == 2) return 1.0;
    return smoothstep(p.lod_magnified_distance, p.lod_minified_distance, distance);
}

d3 kajiya_kay(d3 diffuse, d3 specular, double p, d3 tangent, d3 light, d3 eye) {
    double cos_tl = dot(tangent, light), cos_te = dot(tangent, eye);
    double s1 = 1.0 - cos_tl * cos_tl, s2 = 1.0 - cos_te * cos_te;
    double sin_tl = std::sqrt(s1 > 0.0 ? s1 : 0.0), sin_te = std::sqrt(s2 > 0.0 ? s2 : 0.0);
    double spec = cos_tl * cos_te + sin_tl * sin_te;
    double s = std::pow(spec > 0.0 ? spec : 0.0, p);
    return diffuse * sin_tl + specular * s;
}

}  // namespace

extern "C" {

// Building blocks, exported so that tests can pin each against closed forms.
double oracle_sample_density(const uint8_t* density, const float* density_f32, uint32_t nx, uint32_t ny,
                             uint32_t nz, const float* origin, const float* size, const double* p) {
    Volume v{density, nullptr, density_f32, int(nx), int(ny), int(nz), F3(origin), F3(size)};
    return sample_density(v, d3{p[0], p[1], p[2]});
}

void oracle_kajiya_kay(const double* diffuse, const double* specular, double p, const double* tangent,
                       const double* light, const double* eye, double* out) {
    d3 r = kajiya_kay(d3{diffuse[0], diffuse[1], diffuse[2]}, d3{specular[0], specular[1], specular[2]}, p,
                      d3{tangent[0], tangent[1], tangent[2]}, d3{light[0], light[1], light[2]},
                      d3{eye[0], eye[1], eye[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// volume_approximated_deep_shadows from p toward light_origin (thickness 11).
double oracle_volume_adsm(const uint8_t* density, uint32_t nx, uint32_t ny, uint32_t nz, const float* origin,
                          const float* size, const double* p, const double* light_origin, float steps,
                          double alpha, double* strands_out) {
    Volume v{density, nullptr, nullptr, int(nx), int(ny), int(nz), F3(origin), F3(size)};
    const float step_size = 1.0f / steps;
    double strands = 0.0;
    d3 a{p[0], p[1], p[2]}, b{light_origin[0], light_origin[1], light_origin[2]};
    for (float t = 0.0f; t < 1.0f; t += step_size) strands += sample_density(v, mix(a, b, double(t))) * 11.0;
    if (strands_out) *strands_out = strands;
    return std::pow(1.0 - alpha, strands);
}

// local_ambient_occlusion with kernel size 2 around p.
double oracle_lao(const uint8_t* density, uint32_t nx, uint32_t ny, uint32_t nz, const float* origin,
                  const float* size, const double* p, double radius, double exponent, double clamp) {
    Volume v{density, nullptr, nullptr, int(nx), int(ny), int(nz), F3(origin), F3(size)};
    d3 voxel_space = v.size / d3{double(nx), double(ny), double(nz)};
    d3 scale = voxel_space * (radius / 0.5);
    double dens = 0.0;
    for (float z = -0.5f; z <= 0.5f; z += 1.0f)
        for (float y = -0.5f; y <= 0.5f; y += 1.0f)
            for (float x = -0.5f; x <= 0.5f; x += 1.0f) {
                double s = sample_density(v, d3{p[0], p[1], p[2]} + d3{x, y, z} * scale);
                dens += s < clamp ? s : clamp;
            }
    return std::pow(1.0 - dens / 8.0, exponent);
}

// Per pixel outputs: rgba (clamped to [0,1], as packUnorm4x8 consumes it;
// zero where no fragment), depth (1 where none), info[4] = {fragment (0/1),
// surface step (-1: the vec3(0) surface of the shader), entry step (-1 none),
// found (0/1)}, diag[4] = {accumulated density, ADSM strands, min |acc - iso|
// over the nonzero samples, |cos(tangent, light)|}.
// pixels: list of linear pixel indices to evaluate (n_pixels of them).
// forced: nullable; per evaluated pixel, {surface step, entry step, found}
// decisions to use instead of the oracle's own (the shading is then
// evaluated in float64 at those samples).
void oracle_raymarch(const uint8_t* density, const int8_t* tangent, const float* density_f32, uint32_t nx,
                     uint32_t ny, uint32_t nz, const vkhr_camera* cam, const vkhr_light* light,
                     const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width, uint32_t height,
                     const float* depth_buffer, const int64_t* pixels, uint64_t n_pixels, const int32_t* forced,
                     double* rgba, double* depth, int32_t* info, double* diag) {
    Volume v{density, tangent, density_f32, int(nx), int(ny), int(nz), F3(sp->volume_bounds.origin),
             F3(sp->volume_bounds.size)};
    double vp[16], ivp[16];
    matmul(cam->projection, cam->view, vp);
    invert(vp, ivp);
    d3 eye = F3(cam->position);
    d3 bmin = v.origin, bmax = v.origin + v.size;
    d3 light_origin = F3(light->origin);
    d3 hair_color = F3(sp->hair_color), intensity = F3(light->intensity);
    const double iso = rp->isosurface, alpha = sp->hair_opacity;
    const float steps_f = rp->raycast_steps;
    const float step_size = 1.0f / steps_f;

    for (uint64_t q = 0; q < n_pixels; ++q) {
        int64_t pix = pixels[q];
        int px = int(pix % width), py = int(pix / width);
        double* out = rgba + 4 * q;
        out[0] = out[1] = out[2] = out[3] = 0.0;
        depth[q] = 1.0;
        int32_t* inf = info + 4 * q;
        inf[0] = 0; inf[1] = -1; inf[2] = -1; inf[3] = 0;
        double* dg = diag + 4 * q;
        dg[0] = dg[1] = 0.0; dg[2] = 1e300; dg[3] = 0.0;

        // pixel-centre ray through the eye
        double ndc_x = 2.0 * (px + 0.5) / width - 1.0, ndc_y = 2.0 * (py + 0.5) / height - 1.0;
        double pf[4], h[4] = {ndc_x, ndc_y, 1.0, 1.0};
        for (int r = 0; r < 4; ++r)
            pf[r] = ivp[0 * 4 + r] * h[0] + ivp[1 * 4 + r] * h[1] + ivp[2 * 4 + r] * h[2] + ivp[3 * 4 + r] * h[3];
        d3 far_pt = {pf[0] / pf[3], pf[1] / pf[3], pf[2] / pf[3]};
        d3 dir = normalize(far_pt - eye);
        double t0 = -1e300, t1 = 1e300;
        const double e[3] = {eye.x, eye.y, eye.z}, d[3] = {dir.x, dir.y, dir.z};
        const double lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
        bool miss = false;
        for (int c = 0; c < 3; ++c) {
            if (d[c] == 0.0) {
                if (e[c] < lo[c] || e[c] > hi[c]) miss = true;
                continue;
            }
            double a = (lo[c] - e[c]) / d[c], b = (hi[c] - e[c]) / d[c];
            if (a > b) std::swap(a, b);
            t0 = a > t0 ? a : t0;
            t1 = b < t1 ? b : t1;
        }
        if (miss || t0 > t1 || t0 <= 0.0) continue;
        d3 start = eye + dir * t0;
        {
            double c4[4];
            xform(vp, start, c4);
            double z = c4[2] / c4[3];
            if (!(z >= 0.0 && z <= 1.0)) continue;  // proxy face clipped by near/far
        }
        double db = depth_buffer ? double(depth_buffer[pix]) : 1.0;
        // volume.frag:32-34
        d3 end = start + normalize(start - eye) * double(sp->volume_bounds.radius);

        // volume_surface (volume_rendering.glsl:19-52)
        double acc = 0.0;
        int surf = -1, entry = -1;
        bool found = false;
        int i = 0;
        for (float t = 0.0f; t < 1.0f; t += step_size, ++i) {
            d3 P = mix(start, end, double(t));
            double c4[4];
            xform(vp, P, c4);
            if (db < c4[2] / c4[3]) break;
            double dens = sample_density(v, P);
            acc += dens;
            if (dens != 0.0) {
                double m = std::fabs(acc - iso);
                if (m < dg[2]) dg[2] = m;
                if (acc <= iso) surf = i;
                if (acc >= iso) found = true;
                if (entry < 0) entry = i;
            }
        }
        dg[0] = acc;
        if (forced) {
            surf = forced[3 * q + 0];
            entry = forced[3 * q + 1];
            found = forced[3 * q + 2] != 0;
        }
        inf[0] = 2;  // the ray marched; 1 once a fragment is written
        inf[1] = surf; inf[2] = entry; inf[3] = found ? 1 : 0;
        // sample index -> position (the shader keeps P itself)
        auto pos_of = [&](int k) {
            float t = 0.0f;
            for (int n = 0; n < k; ++n) t += step_size;
            return mix(start, end, double(t));
        };
        d3 surface = {0.0, 0.0, 0.0};
        if (found) {
            if (surf >= 0) surface = pos_of(surf);
        } else if (entry >= 0) {
            surface = pos_of(entry);
        }
        double a_surf = acc / iso;
        if (a_surf == 0.0) continue;  // discard (volume.frag:45-46)
        inf[0] = 1;

        double coverage = lod(*rp, cam->look_at_distance) * a_surf * alpha;
        d3 shading = {1.0, 1.0, 1.0};
        d3 ldir = normalize(light_origin - surface);
        d3 edir = normalize(surface - eye);
        d3 tan = normalize(sample_tangent(v, surface));
        dg[3] = std::fabs(dot(tan, ldir));
        if (rp->shading_model == 0) shading = kajiya_kay(hair_color, intensity, sp->hair_shininess, tan, ldir, edir);

        double occlusion = 1.0;
        if (rp->adsm_on == 1 && rp->shading_model != 3) {
            // volume_approximated_deep_shadows(..., thickness 11)
            double strands = 0.0;
            for (float t = 0.0f; t < 1.0f; t += step_size) {
                d3 P = mix(surface, light_origin, double(t));
                strands += sample_density(v, P) * 11.0;
            }
            dg[1] = strands;
            occlusion *= std::pow(1.0 - alpha, strands);
        }
        if (rp->shading_model != 2) {
            // local_ambient_occlusion(volume, P, origin, size, 2, radius, exponent, clamp)
            const double kernel_size = 2.0;
            double kernel_radius = (kernel_size - 1.0) / 2.0;
            d3 voxel_space = v.size / d3{double(nx), double(ny), double(nz)};
            double voxel_scaling = rp->occlusion_radius / kernel_radius;
            d3 scale = voxel_space * voxel_scaling;
            double dens = 0.0;
            for (float z = -float(kernel_radius); z <= float(kernel_radius); z += 1.0f)
                for (float y = -float(kernel_radius); y <= float(kernel_radius); y += 1.0f)
                    for (float x = -float(kernel_radius); x <= float(kernel_radius); x += 1.0f) {
                        d3 sp3 = surface + d3{x, y, z} * scale;
                        double s = sample_density(v, sp3);
                        dens += s < rp->ao_clamp ? s : double(rp->ao_clamp);
                    }
            occlusion *= std::pow(1.0 - dens / (kernel_size * kernel_size * kernel_size), double(rp->ao_exponent));
        }
        double c4[4];
        xform(vp, surface, c4);
        depth[q] = c4[2] / c4[3];
        double col[4] = {shading.x * occlusion, shading.y * occlusion, shading.z * occlusion, coverage};
        for (int c = 0; c < 4; ++c) out[c] = col[c] < 0.0 ? 0.0 : (col[c] > 1.0 ? 1.0 : col[c]);
    }
}

}  // extern "C"
