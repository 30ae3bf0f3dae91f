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
    int n_primary;         // samples taken by the primary march (up to the depth-buffer break or
                           // the end of the volume's support)
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
    float steps_per_unit;      // raycast_steps when it is a power of two (the loop parameter is then
                               // exactly i * step_size), else 0
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

// The parameter interval [lo, hi] outside of which x0 + t*dx (texel
// coordinates) lies outside [-1, n)^3, where every trilinear tap of the
// zero-bordered volume reads 0.  Widened by far more than the rounding of
// the fma that forms the sample position, so that a sample skipped for
// lying outside it is exactly 0; lo > hi when the line misses the support.
__device__ __forceinline__ void support_interval(const RaymarchArgs& a, const double x0[3], const double dx[3],
                                                 double* lo, double* hi) {
    double l = -1e300, h = 1e300;
    const double n[3] = {double(a.nx), double(a.ny), double(a.nz)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (dx[c] == 0.0) {
            if (!(x0[c] >= -1.0 && x0[c] < n[c])) { l = 1.0; h = 0.0; }
            continue;
        }
        const double margin = 1e-9 + 1e-14 * (fabs(x0[c]) + fabs(dx[c]) + n[c]) / fabs(dx[c]);
        const double u = (-1.0 - x0[c]) / dx[c], v = (n[c] - x0[c]) / dx[c];
        l = fmax(l, fmin(u, v) - margin);
        h = fmin(h, fmax(u, v) + margin);
    }
    *lo = l;
    *hi = h;
}

// A march over the power-of-two sample grid t_i = i / steps (exact in float
// and double): sample i lies at texel coordinate + 1  x1s + i * dxs, the
// same point as x = xs + t_i * dx (dxs = dx / steps exactly).  ea / eb give
// the exits of empty brick regions (region_exit); pos bit c is set when the
// march moves up along axis c.
struct GridRay {
    double x1s[3], dxs[3];
    double ea[3], eb[3];
    uint32_t pos;
};

__device__ __forceinline__ GridRay grid_ray(const double xs[3], const double dx[3], double step, bool skip,
                                            double margin = 1e-6) {
    GridRay g;
    g.pos = 0u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g.x1s[c] = xs[c] + 1.0;
        g.dxs[c] = dx[c] * step;
        g.ea[c] = 0.0;
        g.eb[c] = INFINITY;   // parallel to the planes: never leaves a region along c
        if (skip && g.dxs[c] != 0.0) {
            const double inv = 1.0 / g.dxs[c];
            g.ea[c] = 8.0 * inv;
            // the plane `margin` texels inside the region (by default 1e-6:
            // float64 rounding of a sample's position cannot cross it)
            g.eb[c] = (g.dxs[c] > 0.0 ? -margin - g.x1s[c] : margin - g.x1s[c]) * inv;
            if (g.dxs[c] > 0.0) g.pos |= 1u << c;
        }
    }
    return g;
}

// Sample index (real-valued) where the march leaves the region of bricks
// within chessboard radius r of brick bk: along axis c it spans x1 in
// [8 (b - r), 8 (b + r + 1)), left at the far plane in the direction of
// travel, K = b + r + 1 moving up, K = b - r moving down, at index
// (8 K -+ 1e-6 - x1s) / dxs = K * ea + eb.  Every sample whose index does
// not pass it lies inside the region.
__device__ __forceinline__ double region_exit(const GridRay& g, uint3 bk, int r) {
    const int b[3] = {int(bk.x), int(bk.y), int(bk.z)};
    double te = INFINITY;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int k = ((g.pos >> c) & 1u) ? b[c] + r + 1 : b[c] - r;
        te = fmin(te, fma(double(k), g.ea[c], g.eb[c]));
    }
    return te;
}

// The power-of-two sample grid of the corner-packed cells in 32.32 fixed
// point: sample i of a ray lies at texel coordinate + 1  P(i) = p0 + i * d
// (per axis, in units of 2^-32 texel), with p0 and d the float64 x1s and
// dxs of grid_ray rounded to 2^-32, so P(i) is within (i + 1) 2^-33 texel
// of the float64 position (6e-8 texel at 512 steps).  The cell of a sample
// is the high word of P and its trilinear weight the low word, and the
// ranges of sample indices -- the volume's support, the extent of an empty
// brick region -- are exact integer solutions, so a march over them needs
// no per-sample range test and jumps exactly to the first sample past a
// region.  Used while |x1s| < 2^17 and |dxs| * steps < 2^18 texels (then
// every P and every j * d below stays under 2^52, exact in float64 too).
struct FixedRay {
    long long p0[3], d[3];
};

constexpr double kFixedOne = 4294967296.0;  // 2^32

__device__ __forceinline__ bool fixed_ray_fits(const double x1s[3], const double dxs[3], double steps) {
    bool fits = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) fits = fits && fabs(x1s[c]) < 131072.0 && fabs(dxs[c]) * steps < 262144.0;
    return fits;
}

__device__ __forceinline__ FixedRay fixed_ray(const double x1s[3], const double dxs[3]) {
    FixedRay r;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.p0[c] = llrint(x1s[c] * kFixedOne);   // the products are exact (power-of-two scale)
        r.d[c] = llrint(dxs[c] * kFixedOne);
    }
    return r;
}

// First j in [0, cap] with p + j * d >= k, for d > 0 (cap when no j below
// cap qualifies).  The float64 quotient is within an ulp of the exact one,
// so its ceiling is at most one off, and one integer check each way fixes it.
__device__ __forceinline__ int fixed_first_ge(long long p, long long d, long long k, int cap) {
    if (p >= k) return 0;
    const double q = double(k - p) / double(d);
    if (!(q < double(cap))) return cap;
    long long j = (long long)ceil(q);
    if (j > 0 && p + (j - 1) * d >= k)
        --j;
    else if (p + j * d < k)
        ++j;
    return int(j < cap ? j : cap);
}

// The samples [lo, hi) of a ray (indices below steps) whose cells lie in
// the volume's support, cell indices 0..n along every axis, P in
// [0, (n + 1) 2^32); lo >= hi when there are none.
__device__ __forceinline__ void fixed_support(const FixedRay& r, const int n[3], int steps, int* lo, int* hi) {
    int l = 0, h = steps;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const long long top = (long long)(n[c] + 1) << 32;
        const long long p = r.p0[c], d = r.d[c];
        int in = 0, out = steps;
        if (d > 0) {
            in = fixed_first_ge(p, d, 0, steps);
            out = fixed_first_ge(p, d, top, steps);
        } else if (d < 0) {
            in = fixed_first_ge(-p, -d, -top + 1, steps);
            out = fixed_first_ge(-p, -d, 1, steps);
        } else if (!(p >= 0 && p < top)) {
            in = steps;
        }
        l = max(l, in);
        h = min(h, out);
    }
    *lo = l;
    *hi = h;
}

// The plane margin of the region exits on the fixed-point grid: the
// float64 planes of grid_ray taken far enough inside each region that the
// fixed-point positions, within (steps + 1) 2^-33 texel of the float64
// ones, cannot cross them either.
__device__ __forceinline__ double fixed_exit_margin(int steps) {
    return 1e-6 + double(steps + 1) * 1.1641532182693481e-10;  // 2^-33
}

// The trilinear blend (march_tap's) of a gathered cell at fractions (fx, fy, fz)
// in units of 2^-32.
__device__ __forceinline__ float fixed_cells_blend(unsigned long long q, uint32_t fx, uint32_t fy, uint32_t fz) {
    const uint32_t lo = uint32_t(q), hi = uint32_t(q >> 32);
    const float b0 = float(lo & 0xffu), b1 = float((lo >> 8) & 0xffu);
    const float b2 = float((lo >> 16) & 0xffu), b3 = float(lo >> 24);
    const float b4 = float(hi & 0xffu), b5 = float((hi >> 8) & 0xffu);
    const float b6 = float((hi >> 16) & 0xffu), b7 = float(hi >> 24);
    constexpr float kInv = 2.3283064365386963e-10f;  // 2^-32
    const float wx = float(fx) * kInv, wy = float(fy) * kInv, wz = float(fz) * kInv;
    const float c00 = fmaf(wx, b1 - b0, b0);
    const float c10 = fmaf(wx, b3 - b2, b2);
    const float c01 = fmaf(wx, b5 - b4, b4);
    const float c11 = fmaf(wx, b7 - b6, b6);
    const float c0 = fmaf(wy, c10 - c00, c00);
    const float c1 = fmaf(wy, c11 - c01, c01);
    return fmaf(wz, c1 - c0, c0) * (1.0f / 255.0f);
}

// One trilinear tap of the cells at cell (cx, cy, cz) (texel + 1, inside
// the support) with fractions (fx, fy, fz) in units of 2^-32.
__device__ __forceinline__ float fixed_cells_tap(const DensityVolume& v, uint32_t cx, uint32_t cy, uint32_t cz,
                                                 uint32_t fx, uint32_t fy, uint32_t fz) {
    const uint32_t c = cx + __umul24(uint32_t(v.nx) + 1u, cy + __umul24(uint32_t(v.ny) + 1u, cz));
    return fixed_cells_blend(v.cells[c], fx, fy, fz);
}

// The distance nibble of the brick of cell (cx, cy, cz) (0 without the
// staged field: nothing is skipped).
__device__ __forceinline__ uint32_t fixed_brick_distance(const DensityVolume& v, const OccupancyView& occ,
                                                         uint32_t cx, uint32_t cy, uint32_t cz) {
    if (!occ.use) return 0u;
    const uint32_t b = (cx >> kBrickShift) +
                       __umul24(v.bx, (cy >> kBrickShift) + __umul24(v.by, cz >> kBrickShift));
    return brick_distance(occ, b);
}

// The ADSM march of the cells over the fixed-point grid: the sum of the
// taps of samples [0, steps) from texel coordinate + 1 x1s by dxs per
// sample, jumping over empty brick regions.
__device__ __forceinline__ double fixed_shadow_sum(const DensityVolume& v, const OccupancyView& occ,
                                                   const double x1s[3], const double dxs[3], int steps) {
    const FixedRay r = fixed_ray(x1s, dxs);
    const int n[3] = {v.nx, v.ny, v.nz};
    int j, j_end;
    fixed_support(r, n, steps, &j, &j_end);
    // region exits on the float64 grid of the same samples (x1s - 1 + j dxs)
    const double xs[3] = {x1s[0] - 1.0, x1s[1] - 1.0, x1s[2] - 1.0};
    const GridRay g = grid_ray(xs, dxs, 1.0, occ.use, fixed_exit_margin(steps));
    long long P[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) P[c] = r.p0[c] + (long long)j * r.d[c];
    double sum = 0.0;
    while (j < j_end) {
        const uint32_t cx = uint32_t(P[0] >> 32), cy = uint32_t(P[1] >> 32), cz = uint32_t(P[2] >> 32);
        const uint32_t nib = fixed_brick_distance(v, occ, cx, cy, cz);
        if (nib == 0u) {
            // sample j + 1 too when it lies in the same (occupied) brick
            // before the end: both gathers are in flight together
            const long long Q0 = P[0] + r.d[0], Q1 = P[1] + r.d[1], Q2 = P[2] + r.d[2];
            const uint32_t qx = uint32_t(Q0 >> 32), qy = uint32_t(Q1 >> 32), qz = uint32_t(Q2 >> 32);
            const bool pair = (j + 1 < j_end) & (((qx ^ cx) | (qy ^ cy) | (qz ^ cz)) < (1u << kBrickShift));
            const uint32_t nx1 = uint32_t(v.nx) + 1u, ny1 = uint32_t(v.ny) + 1u;
            const unsigned long long g0 = v.cells[cx + __umul24(nx1, cy + __umul24(ny1, cz))];
            const unsigned long long g1 = pair ? v.cells[qx + __umul24(nx1, qy + __umul24(ny1, qz))] : 0ull;
            sum += double(fixed_cells_blend(g0, uint32_t(P[0]), uint32_t(P[1]), uint32_t(P[2])));
            if (pair) {
                sum += double(fixed_cells_blend(g1, uint32_t(Q0), uint32_t(Q1), uint32_t(Q2)));
                P[0] = Q0 + r.d[0];
                P[1] = Q1 + r.d[1];
                P[2] = Q2 + r.d[2];
                j += 2;
            } else {
                P[0] = Q0;
                P[1] = Q1;
                P[2] = Q2;
                ++j;
            }
        } else {
            // the last sample inside the region (planes taken inside by the
            // margin), then the next one
            const uint3 bk = make_uint3(cx >> kBrickShift, cy >> kBrickShift, cz >> kBrickShift);
            const int next = max(j + 1, int(floor(fmin(region_exit(g, bk, int(nib) - 1), double(j_end)))) + 1);
            const long long k = next - j;
#pragma unroll
            for (int c = 0; c < 3; ++c) P[c] += k * r.d[c];
            j = next;
        }
    }
    return sum;
}

// First grid index i >= 0 with i / steps >= t (steps when there is none
// below steps).
__device__ __forceinline__ int grid_first(double t, double steps) {
    if (!(t > 0.0)) return 0;
    const double q = t * steps;
    return q >= steps ? int(steps) : int(ceil(q));
}

// Number of grid indices i < steps with i / steps <= t.
__device__ __forceinline__ int grid_count(double t, double steps) {
    const double q = t * steps;
    if (!(q >= 0.0)) return 0;
    return q >= steps ? int(steps) : int(floor(q)) + 1;
}

// One pixel of the raymarch pass (volume.frag:31-102).  Every lane of a
// wave must call it (the optional PPLL insert ballots the wave); lanes
// outside the image take no fragment.  occ: the LDS copy of the brick
// occupancy (stage_occupancy), if any; kF32: the float density volume
// (a.vol.f32) instead of the cells.
template <bool kF32>
__device__ __forceinline__ void raymarch_pixel(const RaymarchArgs& a, int px, int py, OccupancyView occ) {
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
        // Past the end of the volume's support every sample is 0 and can
        // change neither the sum nor the surface choices: the march stops
        // there (the depth test before it changes nothing either).
        double t_lo, t_hi;
        support_interval(a, xs, dx, &t_lo, &t_hi);
        // Below t_depth the depth test cannot end the march (z - db*w is
        // affine in t and negative there, by a margin far above its
        // rounding), so it is evaluated only from there on.
        double t_depth = -1.0;
        {
            const double f0 = zs - db * ws, slope = dz - db * dw;
            if (f0 < 0.0) {
                t_depth = 2.0;
                if (slope > 0.0)
                    t_depth = -f0 / slope -
                              (1e-9 + 1e-13 * (fabs(zs) + fabs(db * ws) + fabs(dz) + fabs(db * dw)) / slope);
            }
        }
        // loop invariants in registers (not re-read from the argument block)
        const DensityVolume vol = a.vol;
        const double iso = a.iso;
        const float step = a.step_size;
        int i = 0;
        bool fixed = false;
        if constexpr (!kF32) {
            if (a.steps_per_unit > 0.0f) {
                double x1s[3], dxs[3];
                const double st = double(step);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    x1s[c] = xs[c] + 1.0;
                    dxs[c] = dx[c] * st;
                }
                fixed = fixed_ray_fits(x1s, dxs, double(a.steps_per_unit));
                if (fixed) {
                    // The fixed-point grid: the samples [i_lo, i_end) in the
                    // volume's support, gathered without range tests; one in
                    // an empty brick region jumps to the first sample past
                    // the region, but not past i_depth, from where the depth
                    // test runs on every sample (as on the float64 grid).
                    const int steps = int(a.steps_per_unit);
                    const FixedRay r = fixed_ray(x1s, dxs);
                    const int n[3] = {vol.nx, vol.ny, vol.nz};
                    int i_lo, i_end;
                    fixed_support(r, n, steps, &i_lo, &i_end);
                    const GridRay g = grid_ray(xs, dx, st, occ.use, fixed_exit_margin(steps));
                    const int i_depth = grid_first(t_depth, double(steps));
                    const double dzs = dz * st, dws = dw * st;
                    bool behind = false;
                    // the depth test over samples before the support (they read 0)
                    for (i = i_depth; i < i_lo; ++i) {
                        const double di = double(i);
                        if (db * fma(di, dws, ws) < fma(di, dzs, zs)) {
                            behind = true;
                            break;
                        }
                    }
                    i = i_lo;
                    long long P[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) P[c] = r.p0[c] + (long long)i * r.d[c];
                    while (!behind && i < i_end) {
                        if (i >= i_depth) {
                            const double di = double(i);
                            // depth test: db < z/w  <=>  db*w < z  (w > 0 in front of the eye)
                            if (db * fma(di, dws, ws) < fma(di, dzs, zs)) break;
                        }
                        const uint32_t cx = uint32_t(P[0] >> 32), cy = uint32_t(P[1] >> 32),
                                       cz = uint32_t(P[2] >> 32);
                        const uint32_t nib = fixed_brick_distance(vol, occ, cx, cy, cz);
                        float d = 0.0f;
                        int next = i + 1;
                        if (nib == 0u) {
                            // sample i + 1 too when it lies in the same (occupied)
                            // brick before the end and the depth test: both
                            // gathers are in flight together
                            const long long Q0 = P[0] + r.d[0], Q1 = P[1] + r.d[1], Q2 = P[2] + r.d[2];
                            const uint32_t qx = uint32_t(Q0 >> 32), qy = uint32_t(Q1 >> 32), qz = uint32_t(Q2 >> 32);
                            const bool pair = (next < i_end) & (next < i_depth) &
                                              (((qx ^ cx) | (qy ^ cy) | (qz ^ cz)) < (1u << kBrickShift));
                            const uint32_t nx1 = uint32_t(vol.nx) + 1u, ny1 = uint32_t(vol.ny) + 1u;
                            const uint32_t c0 = cx + __umul24(nx1, cy + __umul24(ny1, cz));
                            const uint32_t c1 = qx + __umul24(nx1, qy + __umul24(ny1, qz));
                            const unsigned long long g0 = vol.cells[c0];
                            const unsigned long long g1 = pair ? vol.cells[c1] : 0ull;
                            d = fixed_cells_blend(g0, uint32_t(P[0]), uint32_t(P[1]), uint32_t(P[2]));
                            if (pair) {
                                const float d1 = fixed_cells_blend(g1, uint32_t(Q0), uint32_t(Q1), uint32_t(Q2));
                                acc += double(d);
                                const bool nz0 = d != 0.0f;
                                surf = (nz0 & (acc <= iso)) ? i : surf;
                                found = found | (nz0 & (acc >= iso));
                                entry = (nz0 & (entry < 0)) ? i : entry;
                                d = d1;
                                ++i;
                                next = i + 1;
                                P[0] = Q0 + r.d[0];
                                P[1] = Q1 + r.d[1];
                                P[2] = Q2 + r.d[2];
                            } else {
                                P[0] = Q0;
                                P[1] = Q1;
                                P[2] = Q2;
                            }
                        } else {
                            if (next < i_depth) {
                                // the last sample inside the region (planes taken
                                // inside by the margin), then the next one
                                const uint3 bk = make_uint3(cx >> kBrickShift, cy >> kBrickShift, cz >> kBrickShift);
                                const int j = int(floor(fmin(region_exit(g, bk, int(nib) - 1), double(i_end))));
                                next = max(next, min(j + 1, i_depth));
                            }
                            const long long k = next - i;
#pragma unroll
                            for (int c = 0; c < 3; ++c) P[c] += k * r.d[c];
                        }
                        acc += double(d);
                        // surface / entry bookkeeping as selects
                        const bool nonzero = d != 0.0f;
                        const bool below = nonzero & (acc <= iso);
                        surf = below ? i : surf;
                        found = found | (nonzero & (acc >= iso));
                        entry = (nonzero & (entry < 0)) ? i : entry;
                        i = next;
                    }
                    t_surf = float(surf) * step;
                    t_entry = float(entry) * step;
                }
            }
        }
        // the cells path takes the float64 grid only where the fixed-point
        // one does not fit, and then the shader's loop below (same samples)
        if (fixed) {
            // marched above
        } else if (kF32 && a.steps_per_unit > 0.0f) {
            // The loop parameter is exactly i * step (a power-of-two step
            // count): the march runs over integer sample indices, bounded by
            // the support interval, with the depth test from the first index
            // at or past t_depth.  A sample in an empty brick region (every
            // sample in it reads 0 and changes neither the sum nor the
            // surface choices) jumps to the first sample past the region, but
            // not past the first index of the depth test, so the march ends
            // at the same sample as without the jumps.
            const double steps = a.steps_per_unit, st = double(step);
            const GridRay g = grid_ray(xs, dx, st, occ.use);
            const double dzs = dz * st, dws = dw * st;
            const int i_end = grid_count(t_hi, steps);
            const int i_depth = grid_first(t_depth, steps);
            i = min(grid_first(t_lo, steps), i_depth);
            while (i < i_end) {
                const double di = double(i);
                if (i >= i_depth) {
                    // depth test: db < z/w  <=>  db*w < z  (w > 0 in front of the eye)
                    if (db * fma(di, dws, ws) < fma(di, dzs, zs)) break;
                }
                int region;
                uint3 bk;
                const float d = grid_tap<kF32>(vol, fma(di, g.dxs[0], g.x1s[0]), fma(di, g.dxs[1], g.x1s[1]),
                                               fma(di, g.dxs[2], g.x1s[2]), occ, &region, &bk);
                int next = i + 1;
                if (region >= 0 && next < i_depth) {
                    const int j = int(floor(fmin(region_exit(g, bk, region), double(i_end))));
                    next = max(next, min(j + 1, i_depth));
                }
                acc += double(d);
                // surface / entry bookkeeping as selects
                const bool nonzero = d != 0.0f;
                const bool below = nonzero & (acc <= iso);
                surf = below ? i : surf;
                found = found | (nonzero & (acc >= iso));
                entry = (nonzero & (entry < 0)) ? i : entry;
                i = next;
            }
            i = min(i, i_end);
            t_surf = float(surf) * step;
            t_entry = float(entry) * step;
        } else {
            // The shader's float loop counter t += step (raycast_steps not a
            // power of two), sample by sample.
            for (float t = 0.0f; t < 1.0f; t += step) {
                const double td = double(t);
                // depth test: db < z/w  <=>  db*w < z  (w > 0 in front of the eye);
                // evaluated as a predicate (no branch around the products)
                const bool behind = (td >= t_depth) & (db * fma(td, dw, ws) < fma(td, dz, zs));
                if (behind | (td > t_hi)) break;
                const float d = march_tap<kF32>(vol, fma(td, dx[0], xs[0]), fma(td, dx[1], xs[1]),
                                                fma(td, dx[2], xs[2]), occ);
                acc += double(d);
                const bool nonzero = d != 0.0f;
                const bool below = nonzero & (acc <= iso);
                surf = below ? i : surf;
                t_surf = below ? t : t_surf;
                found = found | (nonzero & (acc >= iso));
                const bool first = nonzero & (entry < 0);
                entry = first ? i : entry;
                t_entry = first ? t : t_entry;
                ++i;
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
                // only the samples inside the volume's support add anything
                double t_lo, t_hi;
                support_interval(a, sx, xl, &t_lo, &t_hi);
                const DensityVolume vol = a.vol;
                const float step = a.step_size;
                bool adsm_fixed = false;
                if constexpr (!kF32) {
                    if (a.steps_per_unit > 0.0f) {
                        double x1s[3], dxs[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            x1s[c] = sx[c] + 1.0;
                            dxs[c] = xl[c] * double(step);
                        }
                        adsm_fixed = fixed_ray_fits(x1s, dxs, double(a.steps_per_unit));
                        if (adsm_fixed) strands = fixed_shadow_sum(vol, occ, x1s, dxs, int(a.steps_per_unit));
                    }
                }
                if (adsm_fixed) {
                    // summed on the fixed-point grid above
                } else if (kF32 && a.steps_per_unit > 0.0f) {
                    // over the sample grid from the support's start, jumping
                    // over empty brick regions (their samples read 0)
                    const double steps = a.steps_per_unit;
                    const GridRay g = grid_ray(sx, xl, double(step), occ.use);
                    const int j_end = grid_count(t_hi, steps);
                    for (int j = grid_first(t_lo, steps); j < j_end;) {
                        const double dj = double(j);
                        int region;
                        uint3 bk;
                        strands += double(grid_tap<kF32>(vol, fma(dj, g.dxs[0], g.x1s[0]), fma(dj, g.dxs[1], g.x1s[1]),
                                                         fma(dj, g.dxs[2], g.x1s[2]), occ, &region, &bk));
                        int next = j + 1;
                        if (region >= 0)
                            next = max(next, int(floor(fmin(region_exit(g, bk, region), double(j_end)))) + 1);
                        j = next;
                    }
                } else {
                    for (float t = 0.0f; t < 1.0f; t += step) {
                        const double td = double(t);
                        if (td > t_hi) break;
                        if (td >= t_lo)
                            strands += double(march_tap<kF32>(vol, fma(td, xl[0], sx[0]), fma(td, xl[1], sx[1]),
                                                              fma(td, xl[2], sx[2]), occ));
                    }
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

void launch_raymarch(const RaymarchArgs& a, hipStream_t stream);

int make_raymarch_args(const vkhr_volume* vol, const vkhr_camera* cam, const vkhr_light* light,
                       const vkhr_strand_params* sp, const vkhr_render_params* rp, uint32_t width, uint32_t height,
                       RaymarchArgs* a);

}  // namespace vkhr
