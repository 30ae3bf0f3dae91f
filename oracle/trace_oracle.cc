// ORACLE (test infrastructure only): CPU restatement of the strand ray
// tracer -- the reference's Embree path tracer of flat linear curves
// (src/vkhr/ray_tracer.cc:76-169, ray_tracer/hair_style.cc:18-44,50-110,
// ray_tracer/ray.cc:89-111) -- by brute force over every segment.
//
// Embree's own intersection code is not available here (only its headers
// and prebuilt Windows libraries ship with the reference), so the curve rule
// below restates the published geometry of RTC_GEOMETRY_TYPE_FLAT_LINEAR_CURVE:
// a ribbon facing the ray whose radius varies linearly along the segment.
// The device tracer evaluates the same float32 expressions in the same
// order (no contraction), so hits (t, u, segment) agree bit for bit; the
// match to Embree's rounding is unpinned.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

struct Hit {
    float t, u;
    uint32_t seg;
    bool hit;
};

// The random draws of one pixel and sample: a 32-bit integer mix
// (C. Wellons' "lowbias32", 2018) of the seed, pixel, sample and draw index.
uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t draw(uint32_t seed, uint32_t pixel, uint32_t frame, uint32_t k) {
    return mix32(mix32(mix32(seed ^ 0x9e3779b9u) ^ pixel) ^ (frame * 8u + k));
}

// Raytracer::sample (ray_tracer.cc:246-251): (r / UINT32_MAX) * range + min
// in float, UINT32_MAX rounding to 2^32.
float sample(uint32_t r, float lo, float hi) {
    return (float(r) / 4294967296.0f) * (hi - lo) + lo;
}

float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Flat linear curve: the segment's axis and radius projected on the plane
// through the ray origin perpendicular to d; u minimises the projected
// distance to the ray (clamped to the segment), the hit needs that distance
// within the radius at u and the axis depth t = z(u) in [tnear, tfar].
bool flat_curve(const float o[3], const float d[3], float dd, const float p0[3], float r0, const float p1[3],
                float r1, float tnear, float tfar, float* t_out, float* u_out) {
    float q0[3], q1[3];
    for (int c = 0; c < 3; ++c) {
        q0[c] = p0[c] - o[c];
        q1[c] = p1[c] - o[c];
    }
    const float z0 = dot3(q0, d) / dd, z1 = dot3(q1, d) / dd;
    float x0[3], v[3];
    for (int c = 0; c < 3; ++c) {
        x0[c] = q0[c] - z0 * d[c];
        v[c] = (q1[c] - z1 * d[c]) - x0[c];
    }
    const float vv = dot3(v, v);
    float u = 0.0f;
    if (vv > 0.0f) {
        u = -dot3(x0, v) / vv;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
    }
    float p[3];
    for (int c = 0; c < 3; ++c) p[c] = x0[c] + u * v[c];
    const float z = z0 + u * (z1 - z0);
    const float r = r0 + u * (r1 - r0);
    if (!(dot3(p, p) <= r * r)) return false;
    if (!(z >= tnear && z <= tfar)) return false;
    *t_out = z;
    *u_out = u;
    return true;
}

struct Scene {
    const float* xyz;
    const float* radius;
    const uint32_t* indices;
    uint64_t n_pairs;
};

// Closest hit, ties to the lowest segment index.
Hit closest(const Scene& s, const float o[3], const float d[3], float tnear, float tfar, uint64_t skip = ~0ull) {
    Hit best{0.0f, 0.0f, 0u, false};
    const float dd = dot3(d, d);
    for (uint64_t k = 0; k < s.n_pairs; ++k) {
        if (k == skip) continue;
        const uint32_t a = s.indices[2 * k], b = s.indices[2 * k + 1];
        float t, u;
        if (!flat_curve(o, d, dd, s.xyz + 3 * size_t(a), s.radius[a], s.xyz + 3 * size_t(b), s.radius[b], tnear,
                        best.hit ? best.t : tfar, &t, &u))
            continue;
        if (!best.hit || t < best.t) best = Hit{t, u, uint32_t(k), true};
    }
    return best;
}

bool any_hit(const Scene& s, const float o[3], const float d[3], float tnear, float tfar, uint64_t skip = ~0ull) {
    const float dd = dot3(d, d);
    for (uint64_t k = 0; k < s.n_pairs; ++k) {
        if (k == skip) continue;
        const uint32_t a = s.indices[2 * k], b = s.indices[2 * k + 1];
        float t, u;
        if (flat_curve(o, d, dd, s.xyz + 3 * size_t(a), s.radius[a], s.xyz + 3 * size_t(b), s.radius[b], tnear, tfar,
                       &t, &u))
            return true;
    }
    return false;
}

void normalize3(float v[3]) {
    const float l = std::sqrt(dot3(v, v));
    for (int c = 0; c < 3; ++c) v[c] = v[c] / l;
}

// embree::HairStyle::kajiya_kay (ray_tracer/hair_style.cc:86-107) in
// float; a negative 1 - cos^2 (rounding) gives sin 0 and a negative
// cosine sum gives no specular (the reference would produce NaN there).
void kajiya_kay(const float diffuse[3], const float specular[3], float p, const float t[3], const float l[3],
                const float e[3], float out[3]) {
    const float ctl = dot3(l, t), cte = dot3(e, t);
    const float a = 1.0f - ctl * ctl, b = 1.0f - cte * cte;
    const float stl = std::sqrt(a > 0.0f ? a : 0.0f), ste = std::sqrt(b > 0.0f ? b : 0.0f);
    const float base = ctl * cte + stl * ste;
    float spec = base > 0.0f ? std::pow(base, p) : 0.0f;
    spec = spec < 0.0f ? 0.0f : (spec > 1.0f ? 1.0f : spec);
    for (int c = 0; c < 3; ++c) out[c] = diffuse[c] * stl + specular[c] * spec;
}

constexpr float kEpsilon = 0.000001f;  // Ray::Epsilon (ray_tracer/ray.hh:15)

}  // namespace

extern "C" {

uint32_t oracle_trace_draw(uint32_t seed, uint32_t pixel, uint32_t frame, uint32_t k) {
    return draw(seed, pixel, frame, k);
}

int oracle_flat_curve(const float* o, const float* d, const float* p0, float r0, const float* p1, float r1,
                      float tnear, float tfar, float* t, float* u) {
    return flat_curve(o, d, dot3(d, d), p0, r0, p1, r1, tnear, tfar, t, u) ? 1 : 0;
}

// rays: n x 8 floats (origin xyz, tnear, direction xyz, tfar) -> t, u,
// segment (0xffffffff: miss).
void oracle_trace_closest(const float* xyz, const float* radius, const uint32_t* indices, uint64_t n_indices,
                          const float* rays, uint64_t n, float* t_out, float* u_out, uint32_t* seg_out) {
    const Scene s{xyz, radius, indices, n_indices / 2};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        const float* r = rays + 8 * i;
        const Hit h = closest(s, r, r + 4, r[3], r[7]);
        t_out[i] = h.hit ? h.t : 0.0f;
        u_out[i] = h.hit ? h.u : 0.0f;
        seg_out[i] = h.hit ? h.seg : 0xffffffffu;
    }
}

void oracle_trace_any(const float* xyz, const float* radius, const uint32_t* indices, uint64_t n_indices,
                      const float* rays, uint64_t n, uint8_t* out) {
    const Scene s{xyz, radius, indices, n_indices / 2};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        const float* r = rays + 8 * i;
        out[i] = any_hit(s, r, r + 4, r[3], r[7]) ? 1 : 0;
    }
}

// One sample per pixel of Raytracer::draw (ray_tracer.cc:76-169) added to
// accum (W x H x 3 doubles, row j from the top):
//   plane:  ViewingPlane x, y, z, point (12 floats, camera.cc:251-266)
//   light:  spotlight origin (3), intensity (3)
//   method: 0 shaded, 1 combined shadows, 2 direct shadows, 3 ambient
//           occlusion (Raytracer::VisualizationMethod)
void oracle_raytrace_frame(const float* xyz, const float* radius, const float* tangents, const uint32_t* indices,
                           uint64_t n_indices, const float* plane, const float* light, const float* diffuse,
                           float exponent, int method, int shadows_on, float ao_radius, uint32_t seed,
                           uint32_t frame, uint32_t width, uint32_t height, int ignore_self, double* accum) {
    const Scene s{xyz, radius, indices, n_indices / 2};
    const float* px = plane;
    const float* py = plane + 3;
    const float* pz = plane + 6;
    const float* eye = plane + 9;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < int64_t(height); ++j)
        for (uint32_t i = 0; i < width; ++i) {
            const uint32_t pix = i + uint32_t(j) * width;
            float color[3] = {1.0f, 1.0f, 1.0f};
            const float jx = sample(draw(seed, pix, frame, 0), 0.0f, 1.0f);
            const float jy = sample(draw(seed, pix, frame, 1), 0.0f, 1.0f);
            const float fx = float(i) + jx, fy = float(j) + jy;
            float d[3];
            for (int c = 0; c < 3; ++c) d[c] = (fx * px[c] + fy * py[c]) + pz[c];
            const Hit h = closest(s, eye, d, 0.0f, INFINITY);
            if (h.hit) {
                float P[3];
                for (int c = 0; c < 3; ++c) P[c] = eye[c] + d[c] * h.t;
                const uint64_t skip = ignore_self ? uint64_t(h.seg) : ~0ull;
                // light_shading (ray_tracer.cc:116-140): the shadow ray's
                // direction is the jittered light position itself
                float sd[3];
                for (int c = 0; c < 3; ++c) sd[c] = light[c] + sample(draw(seed, pix, frame, 2 + c), -16.0f, 16.0f);
                float L[3] = {1.0f, 1.0f, 1.0f};
                if (method != 3) {
                    const bool occluded = any_hit(s, P, sd, kEpsilon, INFINITY, skip);
                    if (!occluded || !shadows_on) {
                        if (method == 0) {
                            const uint32_t a = indices[2 * size_t(h.seg)], b = indices[2 * size_t(h.seg) + 1];
                            float T[3], l[3], e[3];
                            for (int c = 0; c < 3; ++c) {
                                const float t0 = tangents[3 * size_t(a) + c], t1 = tangents[3 * size_t(b) + c];
                                T[c] = t0 + h.u * (t1 - t0);
                                l[c] = light[c] - P[c];
                                e[c] = P[c] - eye[c];
                            }
                            normalize3(l);
                            normalize3(e);
                            kajiya_kay(diffuse, light + 3, exponent, T, l, e, L);
                        }
                    } else {
                        L[0] = L[1] = L[2] = 0.0f;
                    }
                }
                for (int c = 0; c < 3; ++c) color[c] = L[c];
                if (method != 2) {
                    // ambient_occlusion (ray_tracer.cc:142-160): one ray in a
                    // random direction; occluded when its closest hit lies
                    // within ao_radius of the surface point
                    float rd[3];
                    for (int c = 0; c < 3; ++c) rd[c] = sample(draw(seed, pix, frame, 5 + c), -1.0f, 1.0f);
                    const float len = std::sqrt(dot3(rd, rd));
                    // every hit within ao_radius has t <= ao_radius / |rd|
                    const float t_max = len > 0.0f ? (ao_radius / len) * 1.001f + kEpsilon : INFINITY;
                    const Hit a = closest(s, P, rd, kEpsilon, t_max, skip);
                    float ao = 2.0f;
                    if (a.hit) {
                        float q[3];
                        for (int c = 0; c < 3; ++c) q[c] = (P[c] + rd[c] * a.t) - P[c];
                        if (!(std::sqrt(dot3(q, q)) > ao_radius)) ao = 0.0f;
                    }
                    for (int c = 0; c < 3; ++c) color[c] *= ao;
                }
            }
            for (int c = 0; c < 3; ++c) accum[3 * size_t(pix) + c] += double(color[c]);
        }
}

}  // extern "C"
