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

#include <omp.h>

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

}  // extern "C"

namespace {

struct Ctx {
    Volume v;
    double vp[16], ivp[16];
    d3 eye, bmin, bmax, light_origin, hair_color, intensity;
    double iso, alpha, adsm_k, shininess, lod;
    float step_size;
    const vkhr_camera* cam;
    const vkhr_strand_params* sp;
    const vkhr_render_params* rp;
    uint32_t width, height;
    const float* depth_buffer;
};

// Everything one pixel's forward produces (outputs and per-ray quantities).
struct PixelOut {
    double rgba[4], depth;
    int32_t info[4];
    double diag[4];
    d3 start, end, surface;
    int n_primary;
    double acc, shadow_sum, adsm, lao, shade[3], spec_base, coverage_raw;
};

Ctx make_ctx(const uint8_t* density, const int8_t* tangent, const float* density_f32, uint32_t nx, uint32_t ny,
             uint32_t nz, const vkhr_camera* cam, const vkhr_light* light, const vkhr_strand_params* sp,
             const vkhr_render_params* rp, uint32_t width, uint32_t height, const float* depth_buffer) {
    Ctx c;
    c.v = Volume{density, tangent, density_f32, int(nx), int(ny), int(nz), F3(sp->volume_bounds.origin),
                 F3(sp->volume_bounds.size)};
    matmul(cam->projection, cam->view, c.vp);
    invert(c.vp, c.ivp);
    c.eye = F3(cam->position);
    c.bmin = c.v.origin;
    c.bmax = c.v.origin + c.v.size;
    c.light_origin = F3(light->origin);
    c.hair_color = F3(sp->hair_color);
    c.intensity = F3(light->intensity);
    c.iso = rp->isosurface;
    c.alpha = sp->hair_opacity;
    c.adsm_k = 11.0;  // volume.frag:78
    c.shininess = sp->hair_shininess;
    c.lod = lod(*rp, cam->look_at_distance);
    c.step_size = 1.0f / rp->raycast_steps;
    c.cam = cam; c.sp = sp; c.rp = rp;
    c.width = width; c.height = height;
    c.depth_buffer = depth_buffer;
    return c;
}

void march_pixel(const Ctx& C, int64_t pix, const int32_t* forced, PixelOut& o) {
    const Volume& v = C.v;
    const int px = int(pix % C.width), py = int(pix / C.width);
    o.rgba[0] = o.rgba[1] = o.rgba[2] = o.rgba[3] = 0.0;
    o.depth = 1.0;
    o.info[0] = 0; o.info[1] = -1; o.info[2] = -1; o.info[3] = 0;
    o.diag[0] = o.diag[1] = 0.0; o.diag[2] = 1e300; o.diag[3] = 0.0;
    o.n_primary = 0;
    o.acc = o.shadow_sum = 0.0; o.adsm = o.lao = 1.0; o.spec_base = 0.0; o.coverage_raw = 0.0;
    o.shade[0] = o.shade[1] = o.shade[2] = 1.0;

    // pixel-centre ray through the eye
    double ndc_x = 2.0 * (px + 0.5) / C.width - 1.0, ndc_y = 2.0 * (py + 0.5) / C.height - 1.0;
    double pf[4], h[4] = {ndc_x, ndc_y, 1.0, 1.0};
    for (int r = 0; r < 4; ++r)
        pf[r] = C.ivp[0 * 4 + r] * h[0] + C.ivp[1 * 4 + r] * h[1] + C.ivp[2 * 4 + r] * h[2] + C.ivp[3 * 4 + r] * h[3];
    d3 far_pt = {pf[0] / pf[3], pf[1] / pf[3], pf[2] / pf[3]};
    d3 dir = normalize(far_pt - C.eye);
    double t0 = -1e300, t1 = 1e300;
    const double e[3] = {C.eye.x, C.eye.y, C.eye.z}, d[3] = {dir.x, dir.y, dir.z};
    const double lo[3] = {C.bmin.x, C.bmin.y, C.bmin.z}, hi[3] = {C.bmax.x, C.bmax.y, C.bmax.z};
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
    if (miss || t0 > t1 || t0 <= 0.0) return;
    d3 start = C.eye + dir * t0;
    {
        double c4[4];
        xform(C.vp, start, c4);
        double z = c4[2] / c4[3];
        if (!(z >= 0.0 && z <= 1.0)) return;  // proxy face clipped by near/far
    }
    double db = C.depth_buffer ? double(C.depth_buffer[pix]) : 1.0;
    // volume.frag:32-34
    d3 end = start + normalize(start - C.eye) * double(C.sp->volume_bounds.radius);
    o.start = start;
    o.end = end;

    // volume_surface (volume_rendering.glsl:19-52)
    double acc = 0.0;
    int surf = -1, entry = -1;
    bool found = false;
    int i = 0;
    for (float t = 0.0f; t < 1.0f; t += C.step_size, ++i) {
        d3 P = mix(start, end, double(t));
        double c4[4];
        xform(C.vp, P, c4);
        if (db < c4[2] / c4[3]) break;
        double dens = sample_density(v, P);
        acc += dens;
        if (dens != 0.0) {
            double m = std::fabs(acc - C.iso);
            if (m < o.diag[2]) o.diag[2] = m;
            if (acc <= C.iso) surf = i;
            if (acc >= C.iso) found = true;
            if (entry < 0) entry = i;
        }
    }
    o.n_primary = i;
    o.acc = acc;
    o.diag[0] = acc;
    int forced_state = -1;
    if (forced) {
        forced_state = forced[0];
        surf = forced[1];
        entry = forced[2];
        found = forced[3] != 0;
    }
    o.info[0] = 2;  // the ray marched; 1 once a fragment is written
    o.info[1] = surf; o.info[2] = entry; o.info[3] = found ? 1 : 0;
    // sample index -> position (the shader keeps P itself)
    auto pos_of = [&](int k) {
        float t = 0.0f;
        for (int n = 0; n < k; ++n) t += C.step_size;
        return mix(start, end, double(t));
    };
    d3 surface = {0.0, 0.0, 0.0};
    if (found) {
        if (surf >= 0) surface = pos_of(surf);
    } else if (entry >= 0) {
        surface = pos_of(entry);
    }
    o.surface = surface;
    double a_surf = acc / C.iso;
    // discard (volume.frag:45-46); a forced state overrides it
    if (forced_state >= 0 ? forced_state != 1 : a_surf == 0.0) return;
    o.info[0] = 1;

    double coverage = C.lod * a_surf * C.alpha;
    o.coverage_raw = coverage;
    d3 shading = {1.0, 1.0, 1.0};
    d3 ldir = normalize(C.light_origin - surface);
    d3 edir = normalize(surface - C.eye);
    d3 tan = normalize(sample_tangent(v, surface));
    o.diag[3] = std::fabs(dot(tan, ldir));
    if (C.rp->shading_model == 0) {
        shading = kajiya_kay(C.hair_color, C.intensity, C.shininess, tan, ldir, edir);
        double cos_tl = dot(tan, ldir), cos_te = dot(tan, edir);
        double s1 = 1.0 - cos_tl * cos_tl, s2 = 1.0 - cos_te * cos_te;
        double b = cos_tl * cos_te + std::sqrt(s1 > 0 ? s1 : 0.0) * std::sqrt(s2 > 0 ? s2 : 0.0);
        o.spec_base = b > 0.0 ? b : 0.0;
    }
    o.shade[0] = shading.x; o.shade[1] = shading.y; o.shade[2] = shading.z;

    double occlusion = 1.0;
    if (C.rp->adsm_on == 1 && C.rp->shading_model != 3) {
        // volume_approximated_deep_shadows(..., thickness)
        double sum = 0.0;
        for (float t = 0.0f; t < 1.0f; t += C.step_size) sum += sample_density(v, mix(surface, C.light_origin, double(t)));
        o.shadow_sum = sum;
        o.diag[1] = sum * C.adsm_k;
        o.adsm = std::pow(1.0 - C.alpha, sum * C.adsm_k);
        occlusion *= o.adsm;
    }
    if (C.rp->shading_model != 2) {
        // local_ambient_occlusion(volume, P, origin, size, 2, radius, exponent, clamp)
        const double kernel_size = 2.0;
        double kernel_radius = (kernel_size - 1.0) / 2.0;
        d3 voxel_space = v.size / d3{double(v.nx), double(v.ny), double(v.nz)};
        double voxel_scaling = C.rp->occlusion_radius / kernel_radius;
        d3 scale = voxel_space * voxel_scaling;
        double dens = 0.0;
        for (float z = -float(kernel_radius); z <= float(kernel_radius); z += 1.0f)
            for (float y = -float(kernel_radius); y <= float(kernel_radius); y += 1.0f)
                for (float x = -float(kernel_radius); x <= float(kernel_radius); x += 1.0f) {
                    d3 sp3 = surface + d3{x, y, z} * scale;
                    double s = sample_density(v, sp3);
                    dens += s < C.rp->ao_clamp ? s : double(C.rp->ao_clamp);
                }
        o.lao = std::pow(1.0 - dens / (kernel_size * kernel_size * kernel_size), double(C.rp->ao_exponent));
        occlusion *= o.lao;
    }
    double c4[4];
    xform(C.vp, surface, c4);
    o.depth = c4[2] / c4[3];
    double col[4] = {shading.x * occlusion, shading.y * occlusion, shading.z * occlusion, coverage};
    for (int c = 0; c < 4; ++c) o.rgba[c] = col[c] < 0.0 ? 0.0 : (col[c] > 1.0 ? 1.0 : col[c]);
}

// Scatter c * (exact trilinear weights at world point P) into grad.
void scatter(const Volume& v, d3 P, double c, double* grad) {
    d3 uvw = (P - v.origin) / v.size;
    double x = uvw.x * v.nx - 0.5, y = uvw.y * v.ny - 0.5, z = uvw.z * v.nz - 0.5;
    if (!(std::fabs(x) < 1e6 && std::fabs(y) < 1e6 && std::fabs(z) < 1e6)) return;
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int i = int(fx), j = int(fy), k = int(fz);
    double a = x - fx, b = y - fy, cc = z - fz;
    for (int n = 0; n < 8; ++n) {
        int dx = n & 1, dy = (n >> 1) & 1, dz = n >> 2;
        int tx = i + dx, ty = j + dy, tz = k + dz;
        if (tx < 0 || ty < 0 || tz < 0 || tx >= v.nx || ty >= v.ny || tz >= v.nz) continue;
        double w = (dx ? a : 1 - a) * (dy ? b : 1 - b) * (dz ? cc : 1 - cc);
        grad[size_t(tx) + size_t(v.nx) * (size_t(ty) + size_t(v.ny) * size_t(tz))] += c * w;
    }
}

}  // namespace

extern "C" {

// Per pixel outputs: rgba (clamped to [0,1], as packUnorm4x8 consumes it;
// zero where no fragment), depth (1 where none), info[4] = {0: ray misses the box, 1: fragment,
// 2: marched but discarded; surface step (-1: the vec3(0) surface of the shader), entry step (-1 none),
// found (0/1)}, diag[4] = {accumulated density, ADSM strands, min |acc - iso|
// over the nonzero samples, |cos(tangent, light)|}.
// Threads of the OpenMP team used by the pixel loops (the CPU baseline times
// one and all cores).
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

// pixels: list of linear pixel indices to evaluate (n_pixels of them).
// forced: nullable; per evaluated pixel, {state, surface step, entry step,
// found} decisions to use instead of the oracle's own (fragment or not, and
// the samples the shading is evaluated at, in float64).
void oracle_raymarch(const uint8_t* density, const int8_t* tangent, const float* density_f32, uint32_t nx,
                     uint32_t ny, uint32_t nz, const vkhr_camera* cam, const vkhr_light* light,
                     const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width, uint32_t height,
                     const float* depth_buffer, const int64_t* pixels, uint64_t n_pixels, const int32_t* forced,
                     double* rgba, double* depth, int32_t* info, double* diag) {
    Ctx C = make_ctx(density, tangent, density_f32, nx, ny, nz, cam, light, sp, rp, width, height, depth_buffer);
    // pixels are independent: the OpenMP team (oracle_set_threads) splits them
#pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t q = 0; q < n_pixels; ++q) {
        PixelOut o;
        march_pixel(C, pixels[q], forced ? forced + 4 * q : nullptr, o);
        for (int c = 0; c < 4; ++c) {
            rgba[4 * q + c] = o.rgba[c];
            info[4 * q + c] = o.info[c];
            diag[4 * q + c] = o.diag[c];
        }
        depth[q] = o.depth;
    }
}

// The differentiable form: forward with parameters (alpha, thickness, p) in
// place of hair_opacity / 11 / hair_shininess, then the analytic gradient of
// L given dL/d(rgba) per pixel (d_rgba, W*H*4): per-sample scatter of the
// per-ray weights into d_density (float64, nx*ny*nz, accumulated) and the
// parameter gradients into d_params[3].  Optionally returns rgba (W*H*4).
void oracle_raymarch_grad(const int8_t* tangent, const float* density_f32, uint32_t nx, uint32_t ny, uint32_t nz,
                          const vkhr_camera* cam, const vkhr_light* light, const vkhr_strand_params* sp,
                          const vkhr_render_params* rp, const float* diff, uint32_t width, uint32_t height,
                          const float* d_rgba, double* rgba_out, double* d_density, double* d_params) {
    Ctx C = make_ctx(nullptr, tangent, density_f32, nx, ny, nz, cam, light, sp, rp, width, height, nullptr);
    C.alpha = diff[0];
    C.adsm_k = diff[1];
    C.shininess = diff[2];
    const bool adsm_on = rp->adsm_on == 1 && rp->shading_model != 3, lao_on = rp->shading_model != 2;
    const double ln_t = std::log(1.0 - C.alpha);
    PixelOut o;
    for (int64_t pix = 0; pix < int64_t(width) * height; ++pix) {
        march_pixel(C, pix, nullptr, o);
        if (rgba_out)
            for (int c = 0; c < 4; ++c) rgba_out[4 * pix + c] = o.rgba[c];
        if (o.info[0] != 1) continue;
        const float* g = d_rgba + 4 * pix;
        double G = 0.0;
        for (int c = 0; c < 3; ++c) {
            double rgb = o.shade[c] * o.adsm * o.lao;
            if (rgb > 0.0 && rgb < 1.0) {
                G += g[c] * rgb;
                if (rp->shading_model == 0 && o.spec_base > 0.0)
                    d_params[2] += g[c] * o.adsm * o.lao * (c == 0 ? C.intensity.x : c == 1 ? C.intensity.y : C.intensity.z) *
                                   std::pow(o.spec_base, C.shininess) * std::log(o.spec_base);
            }
        }
        double cov = o.coverage_raw;
        if (cov > 0.0 && cov < 1.0) {
            double cp = g[3] * C.lod * C.alpha / C.iso;
            d_params[0] += g[3] * C.lod * o.acc / C.iso;
            float t = 0.0f;
            for (int i = 0; i < o.n_primary; ++i, t += C.step_size) scatter(C.v, mix(o.start, o.end, double(t)), cp, d_density);
        }
        if (adsm_on) {
            double cs = G * ln_t * C.adsm_k;
            d_params[1] += G * ln_t * o.shadow_sum;
            d_params[0] += G * (C.adsm_k * o.shadow_sum) * (-1.0 / (1.0 - C.alpha));
            for (float t = 0.0f; t < 1.0f; t += C.step_size)
                scatter(C.v, mix(o.surface, C.light_origin, double(t)), cs, d_density);
        }
        if (lao_on && G != 0.0) {
            d3 voxel_space = C.v.size / d3{double(nx), double(ny), double(nz)};
            d3 scale = voxel_space * (double(rp->occlusion_radius) / 0.5);
            double vals[8], D = 0.0;
            d3 taps[8];
            int n = 0;
            for (float z = -0.5f; z <= 0.5f; z += 1.0f)
                for (float y = -0.5f; y <= 0.5f; y += 1.0f)
                    for (float x = -0.5f; x <= 0.5f; x += 1.0f, ++n) {
                        taps[n] = o.surface + d3{x, y, z} * scale;
                        vals[n] = sample_density(C.v, taps[n]);
                        D += vals[n] < rp->ao_clamp ? vals[n] : double(rp->ao_clamp);
                    }
            double cl = G * (-double(rp->ao_exponent)) / (8.0 - D);
            for (int q = 0; q < 8; ++q)
                if (vals[q] < rp->ao_clamp) scatter(C.v, taps[q], cl, d_density);
        }
    }
}

}  // extern "C"
