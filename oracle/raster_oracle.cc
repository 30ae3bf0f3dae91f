// ORACLE -- test infrastructure only.  Float64 transcription of the
// reference's strand rasterizer passes, used by tests/ as the checker; never
// linked into the product library.
//
// Transcribed (paths under share/shaders/ unless noted):
//   self-shadowing/depth_map.vert:9-12 + src/vkhr/rasterizer.cc:293-315
//                                   shadow-map depth pass (depth test LESS, writes)
//   strands/strand.vert:21-31, strands/strand.frag:34-87
//                                   the strand colour pass
//   anti-aliasing/gpaa.glsl:33-51   line coverage
//   shading/kajiya-kay.glsl:7-24    (unnormalized interpolated tangent, as the shader)
//   self-shadowing/approximate_deep_shadows.glsl:12-21,41-73, tex2Dproj.glsl:5-10
//                                   3x3 Gaussian PCF of the shadow map
//   volumes/local_ambient_occlusion.glsl:9-30
//   transparency/ppll.glsl:31-33    packUnorm4x8 of the node colour
//   transparency/resolve.comp:15-102  the k-buffer resolve
// Fixed-function parts made explicit with the build's line rule (declared in
// include/vkhr_hip.h): clip to 0 <= z <= w, window mapping, pixel centre c
// covered when 0 <= (c-a).(b-a) < |b-a|^2 and cross^2 <= (w/2)^2 |b-a|^2;
// here every pixel of the segment's bounding box is tested (brute force).
// Depth interpolates linearly in t, varyings perspective-correctly.  The
// depth map is sampled bilinearly with clamp-to-edge
// (src/vkhr/rasterizer/depth_map.cc:53-59).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/vkhr_hip.h"

namespace {

inline void xform4(const double* m, const double p[3], double o[4]) {
    for (int r = 0; r < 4; ++r) o[r] = m[0 * 4 + r] * p[0] + m[1 * 4 + r] * p[1] + m[2 * 4 + r] * p[2] + m[3 * 4 + r];
}

void matmul(const float* a, const float* b, double* out) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += double(a[k * 4 + r]) * double(b[c * 4 + k]);
            out[c * 4 + r] = s;
        }
}

struct Line {
    double ax, ay, az, bx, by, bz, wa, wb, s0, s1, vx, vy, l2;
};

// Clip to 0 <= z <= w and map to window coordinates; false when culled.
bool setup(const double* m, const double pa[3], const double pb[3], int width, int height, Line& L) {
    double ca[4], cb[4];
    xform4(m, pa, ca);
    xform4(m, pb, cb);
    double s0 = 0.0, s1 = 1.0;
    double fa[2] = {ca[2], ca[3] - ca[2]}, fb[2] = {cb[2], cb[3] - cb[2]};
    for (int k = 0; k < 2; ++k) {
        if (fa[k] < 0.0 && fb[k] < 0.0) return false;
        if (fa[k] < 0.0) s0 = std::fmax(s0, fa[k] / (fa[k] - fb[k]));
        if (fb[k] < 0.0) s1 = std::fmin(s1, fa[k] / (fa[k] - fb[k]));
    }
    if (!(s0 < s1)) return false;
    double c0[4], c1[4];
    for (int r = 0; r < 4; ++r) {
        c0[r] = ca[r] + s0 * (cb[r] - ca[r]);
        c1[r] = ca[r] + s1 * (cb[r] - ca[r]);
    }
    if (!(c0[3] > 0.0 && c1[3] > 0.0)) return false;
    L.s0 = s0; L.s1 = s1; L.wa = c0[3]; L.wb = c1[3];
    L.ax = (c0[0] / c0[3] + 1.0) * (0.5 * width);
    L.ay = (c0[1] / c0[3] + 1.0) * (0.5 * height);
    L.az = c0[2] / c0[3];
    L.bx = (c1[0] / c1[3] + 1.0) * (0.5 * width);
    L.by = (c1[1] / c1[3] + 1.0) * (0.5 * height);
    L.bz = c1[2] / c1[3];
    L.vx = L.bx - L.ax;
    L.vy = L.by - L.ay;
    L.l2 = L.vx * L.vx + L.vy * L.vy;
    return L.l2 > 0.0 && L.l2 < 1e30;
}

template <class F>
void for_each_fragment(const Line& L, int width, int height, double line_width, F emit) {
    double hw = 0.5 * line_width;
    int x0 = int(std::floor(std::fmin(L.ax, L.bx) - hw)) - 1, x1 = int(std::ceil(std::fmax(L.ax, L.bx) + hw)) + 1;
    int y0 = int(std::floor(std::fmin(L.ay, L.by) - hw)) - 1, y1 = int(std::ceil(std::fmax(L.ay, L.by) + hw)) + 1;
    x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0;
    x1 = x1 > width - 1 ? width - 1 : x1; y1 = y1 > height - 1 ? height - 1 : y1;
    for (int py = y0; py <= y1; ++py)
        for (int px = x0; px <= x1; ++px) {
            double cx = px + 0.5 - L.ax, cy = py + 0.5 - L.ay;
            double dp = cx * L.vx + cy * L.vy;
            if (!(dp >= 0.0 && dp < L.l2)) continue;
            double cr = L.vx * cy - L.vy * cx;
            if (!(cr * cr <= hw * hw * L.l2)) continue;
            emit(px, py, dp / L.l2);
        }
}

inline double* at3(const float* p, uint32_t i, double o[3]) {
    o[0] = p[3 * size_t(i)]; o[1] = p[3 * size_t(i) + 1]; o[2] = p[3 * size_t(i) + 2];
    return o;
}

double density_texel(const uint8_t* d, const float* f, int nx, int ny, int nz, int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.0;
    size_t i = size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * size_t(z));
    return f ? double(f[i]) : double(d[i]) / 255.0;
}

double trilinear(const uint8_t* d, const float* f, int nx, int ny, int nz, double x, double y, double z) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    if (!(std::fabs(x) < 1e9 && std::fabs(y) < 1e9 && std::fabs(z) < 1e9)) return 0.0;
    int i = int(fx), j = int(fy), k = int(fz);
    double a = x - fx, b = y - fy, c = z - fz, r = 0.0;
    for (int n = 0; n < 8; ++n) {
        int dx = n & 1, dy = (n >> 1) & 1, dz = n >> 2;
        double w = (dx ? a : 1 - a) * (dy ? b : 1 - b) * (dz ? c : 1 - c);
        if (w != 0.0) r += w * density_texel(d, f, nx, ny, nz, i + dx, j + dy, k + dz);
    }
    return r;
}

double shadow_tex(const float* map, int sw, int sh, double u, double v) {
    double x = u * sw - 0.5, y = v * sh - 0.5;
    if (!(std::fabs(x) < 1e9 && std::fabs(y) < 1e9)) return map[0];
    double fx = std::floor(x), fy = std::floor(y);
    double wx = x - fx, wy = y - fy;
    auto cl = [](int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); };
    int x0 = int(fx), y0 = int(fy);
    int xa = cl(x0, sw), xb = cl(x0 + 1, sw), ya = cl(y0, sh), yb = cl(y0 + 1, sh);
    double t00 = map[size_t(ya) * sw + xa], t10 = map[size_t(ya) * sw + xb];
    double t01 = map[size_t(yb) * sw + xa], t11 = map[size_t(yb) * sw + xb];
    double r0 = t00 * (1 - wx) + t10 * wx, r1 = t01 * (1 - wx) + t11 * wx;
    return r0 * (1 - wy) + r1 * wy;
}

inline void normalize3(double v[3]) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double s = l > 0.0 ? 1.0 / l : 0.0;
    v[0] *= s; v[1] *= s; v[2] *= s;
}

}  // namespace

extern "C" {

void oracle_shadow_depth(const float* xyz, uint32_t n_vertices, const uint32_t* indices, uint64_t n_indices,
                         const float* light_vp, float line_width, uint32_t sw, uint32_t sh, float* out) {
    std::vector<double> best(size_t(sw) * sh, 1.0);
    double m[16];
    for (int i = 0; i < 16; ++i) m[i] = light_vp[i];
    for (uint64_t s = 0; s + 1 < n_indices; s += 2) {
        uint32_t ia = indices[s], ib = indices[s + 1];
        if (ia >= n_vertices || ib >= n_vertices) continue;
        double pa[3], pb[3];
        at3(xyz, ia, pa);
        at3(xyz, ib, pb);
        Line L;
        if (!setup(m, pa, pb, int(sw), int(sh), L)) continue;
        for_each_fragment(L, int(sw), int(sh), line_width, [&](int px, int py, double t) {
            double z = L.az + t * (L.bz - L.az);
            double& b = best[size_t(py) * sw + px];
            if (float(z) < float(b)) b = z;
        });
    }
    for (size_t i = 0; i < best.size(); ++i) out[i] = float(best[i]);
}

// Enumerates and shades the fragments of the listed index pairs (all when
// pairs is NULL).  Per fragment: segment (pair number), pixel, rgba before
// packing (float64, unclamped), window depth, and kept (1 when the fragment
// passes the depth test and the coverage test and becomes a node).
// Returns the number of fragments; writes at most capacity of them.
uint64_t oracle_raster(const float* xyz, const float* tangents, const float* thickness, uint32_t n_vertices,
                       const uint32_t* indices, uint64_t n_indices, const uint64_t* pairs, uint64_t n_list,
                       const vkhr_camera* cam, const vkhr_light* light, const vkhr_strand_params* sp,
                       const vkhr_render_params* rp, const uint8_t* density, const float* density_f32, uint32_t nx,
                       uint32_t ny, uint32_t nz, const float* shadow_map, uint32_t sw, uint32_t sh,
                       const float* depth_buffer, uint64_t capacity, uint32_t* out_segment, uint32_t* out_pixel,
                       double* out_rgba, double* out_depth, uint8_t* out_kept) {
    double vp[16], lm[16];
    matmul(cam->projection, cam->view, vp);
    for (int i = 0; i < 16; ++i) lm[i] = light->matrix[i];
    const int W = int(cam->resolution[0]) > 0 ? int(cam->resolution[0]) : 1;
    const int H = int(cam->resolution[1]) > 0 ? int(cam->resolution[1]) : 1;
    const double lw = sp->strand_radius, alpha = sp->hair_opacity;
    const double lod = rp->renderer == 0 ? 0.0 : rp->renderer == 2 ? 1.0 : [&] {
        double t = (double(cam->look_at_distance) - rp->lod_magnified_distance) /
                   (double(rp->lod_minified_distance) - rp->lod_magnified_distance);
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        return t * t * (3.0 - 2.0 * t);
    }();
    const bool adsm = rp->adsm_on == 1 && rp->shading_model != 3, lao = rp->shading_model != 2;
    // PCF kernel (approximate_deep_shadows.glsl:47-63)
    std::vector<double> kx, ky, kw;
    if (adsm) {
        double kwid = rp->adsm_kernel_size, smoothing = rp->adsm_stride_size;
        float kr = float((kwid - 1.0) / 2.0);
        double sigma = (kwid / 2.0) / 2.4, sigma2 = sigma * sigma;
        for (float y = -kr; y <= kr; y += 1.0f)
            for (float x = -kr; x <= kr; x += 1.0f) {
                double exponent = -1.0 * (double(x) * x + double(y) * y) / 2.0 * sigma2;
                kw.push_back(1.0 / (2.0 * M_PI * sigma2) * std::pow(M_E, exponent));
                kx.push_back(double(x) / (double(sw) / smoothing));
                ky.push_back(double(y) / (double(sh) / smoothing));
            }
    }
    uint64_t count = 0;
    uint64_t n_pairs = n_indices / 2;
    uint64_t n_iter = pairs ? n_list : n_pairs;
    for (uint64_t q = 0; q < n_iter; ++q) {
        uint64_t s = pairs ? pairs[q] : q;
        if (s >= n_pairs) continue;
        uint32_t ia = indices[2 * s], ib = indices[2 * s + 1];
        if (ia >= n_vertices || ib >= n_vertices) continue;
        double pa[3], pb[3], ta[3], tb[3];
        at3(xyz, ia, pa);
        at3(xyz, ib, pb);
        at3(tangents, ia, ta);
        at3(tangents, ib, tb);
        double ha = thickness[ia], hb = thickness[ib];
        Line L;
        if (!setup(vp, pa, pb, W, H, L)) continue;
        double P0[3], P1[3], T0[3], T1[3];
        for (int c = 0; c < 3; ++c) {
            P0[c] = pa[c] + L.s0 * (pb[c] - pa[c]);
            P1[c] = pa[c] + L.s1 * (pb[c] - pa[c]);
            T0[c] = ta[c] + L.s0 * (tb[c] - ta[c]);
            T1[c] = ta[c] + L.s1 * (tb[c] - ta[c]);
        }
        double H0 = ha + L.s0 * (hb - ha), H1 = ha + L.s1 * (hb - ha);
        for_each_fragment(L, W, H, lw, [&](int px, int py, double t) {
            uint64_t k = count++;
            if (k >= capacity) return;
            size_t pix = size_t(py) * W + px;
            double z = L.az + t * (L.bz - L.az);
            out_segment[k] = uint32_t(s);
            out_pixel[k] = uint32_t(pix);
            out_depth[k] = z;
            double* rgba = out_rgba + 4 * k;
            rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0.0;
            out_kept[k] = 0;
            if (!(float(z) < (depth_buffer ? depth_buffer[pix] : 1.0f))) return;
            double iw = (1.0 - t) / L.wa + t / L.wb;
            double spar = (t / L.wb) / iw;
            double P[3], T[3];
            for (int c = 0; c < 3; ++c) {
                P[c] = P0[c] + spar * (P1[c] - P0[c]);
                T[c] = T0[c] + spar * (T1[c] - T0[c]);
            }
            double thick = H0 + spar * (H1 - H0);
            // gpaa
            double cl[4];
            xform4(vp, P, cl);
            double sx = (cl[0] / cl[3] + 1.0) * (double(cam->resolution[0]) / 2.0);
            double sy = (cl[1] / cl[3] + 1.0) * (double(cam->resolution[1]) / 2.0);
            double d = std::sqrt((sx - (px + 0.5)) * (sx - (px + 0.5)) + (sy - (py + 0.5)) * (sy - (py + 0.5)));
            double coverage = 1.0 - d / (lw / 2.0);
            coverage *= alpha;
            if (coverage < 0.001) return;  // discard
            coverage *= 1.0 - lod;
            coverage *= thick * (1.0 / 0.042);
            double e[3] = {P[0] - cam->position[0], P[1] - cam->position[1], P[2] - cam->position[2]};
            double l[3] = {light->origin[0] - P[0], light->origin[1] - P[1], light->origin[2] - P[2]};
            normalize3(e);
            normalize3(l);
            double shading[3] = {1.0, 1.0, 1.0};
            if (rp->shading_model == 0) {
                double ctl = T[0] * l[0] + T[1] * l[1] + T[2] * l[2];
                double cte = T[0] * e[0] + T[1] * e[1] + T[2] * e[2];
                double a1 = 1.0 - ctl * ctl, a2 = 1.0 - cte * cte;
                double stl = std::sqrt(a1 > 0 ? a1 : 0.0), ste = std::sqrt(a2 > 0 ? a2 : 0.0);
                double spec = ctl * cte + stl * ste;
                double sp80 = std::pow(spec > 0 ? spec : 0.0, double(sp->hair_shininess));
                for (int c = 0; c < 3; ++c) shading[c] = double(sp->hair_color[c]) * stl + double(light->intensity[c]) * sp80;
            }
            double occlusion = 1.0;
            if (adsm) {
                double ls[4];
                xform4(lm, P, ls);
                double light_depth = ls[2] / ls[3];
                double vis = 0.0, tot = 0.0;
                for (size_t q2 = 0; q2 < kw.size(); ++q2) {
                    double sd = shadow_tex(shadow_map, int(sw), int(sh), ls[0] / ls[3] + kx[q2], ls[1] / ls[3] + ky[q2]);
                    double depth = std::fmax(light_depth - sd, 0.0);
                    double n = depth * 15000.0;
                    if (depth > 1e-5) n += 1.0;
                    vis += std::pow(1.0 - alpha, n) * kw[q2];
                    tot += kw[q2];
                }
                occlusion *= vis / tot;
            }
            if (lao) {
                double st[3], dens = 0.0;
                const double res3[3] = {double(nx), double(ny), double(nz)};
                for (int c = 0; c < 3; ++c)
                    st[c] = (P[c] - sp->volume_bounds.origin[c]) / sp->volume_bounds.size[c] * res3[c] - 0.5;
                double scale = double(rp->occlusion_radius) / 0.5;
                for (float zz = -0.5f; zz <= 0.5f; zz += 1.0f)
                    for (float yy = -0.5f; yy <= 0.5f; yy += 1.0f)
                        for (float xx = -0.5f; xx <= 0.5f; xx += 1.0f) {
                            double v = trilinear(density, density_f32, int(nx), int(ny), int(nz), st[0] + xx * scale,
                                                 st[1] + yy * scale, st[2] + zz * scale);
                            dens += v < rp->ao_clamp ? v : double(rp->ao_clamp);
                        }
                occlusion *= std::pow(1.0 - dens / 8.0, double(rp->ao_exponent));
            }
            rgba[0] = shading[0] * occlusion;
            rgba[1] = shading[1] * occlusion;
            rgba[2] = shading[2] * occlusion;
            rgba[3] = coverage;
            out_kept[k] = 1;
        });
    }
    return count;
}

// resolve.comp:15-102 on the given lists; out = float64 RGBA per pixel.
void oracle_resolve(const uint32_t* heads, const uint32_t* nodes, uint32_t node_count, uint32_t width,
                    uint32_t height, const float* background, double* out) {
    const uint32_t kNull = 0xffffffffu;
    for (size_t pix = 0; pix < size_t(width) * height; ++pix) {
        double res[4];
        for (int c = 0; c < 4; ++c) res[c] = background ? background[4 * pix + c] : 1.0;
        uint32_t head = heads[pix];
        if (head == kNull) {
            for (int c = 0; c < 4; ++c) out[4 * pix + c] = res[c];
            continue;
        }
        double depth[16];
        uint32_t color[16];
        for (int k = 0; k < 16; ++k) { depth[k] = double(3.e+38f); color[k] = 0; }
        for (int k = 0; k < 16; ++k)
            if (head != kNull && head < node_count) {
                color[k] = nodes[3 * size_t(head)];
                float fd;
                std::memcpy(&fd, &nodes[3 * size_t(head) + 1], 4);
                depth[k] = fd;
                head = nodes[3 * size_t(head) + 2];
            } else {
                head = kNull;
            }
        auto blend = [&](uint32_t c) {
            double f[4] = {(c & 0xff) / 255.0, ((c >> 8) & 0xff) / 255.0, ((c >> 16) & 0xff) / 255.0, (c >> 24) / 255.0};
            for (int i = 0; i < 4; ++i) res[i] = res[i] * (1.0 - f[3]) + f[i] * f[3];
        };
        for (int f = 0; f < 1024; ++f) {
            if (head == kNull || head >= node_count) break;
            double fd_max = 0.0;
            int idx = 0;
            for (int k = 0; k < 16; ++k)
                if (depth[k] > fd_max) { fd_max = depth[k]; idx = k; }
            uint32_t fc = nodes[3 * size_t(head)];
            float fdf;
            std::memcpy(&fdf, &nodes[3 * size_t(head) + 1], 4);
            double fd = fdf;
            uint32_t next = nodes[3 * size_t(head) + 2];
            if (fd < fd_max) {
                std::swap(color[idx], fc);
                std::swap(depth[idx], fd);
            }
            blend(fc);
            head = next;
        }
        for (int r = 0; r < 16; ++r) {
            double fd_max = 0.0;
            int idx = 0;
            for (int k = 0; k < 16; ++k)
                if (depth[k] > fd_max) { fd_max = depth[k]; idx = k; }
            depth[idx] = 0.0;
            blend(color[idx]);
        }
        for (int c = 0; c < 4; ++c) out[4 * pix + c] = res[c];
    }
}

}  // extern "C"
