// ORACLE -- test infrastructure only.  CPU restatement of the reference's
// load-time voxelization path, used by tests/ and by bench.py's CPU-baseline
// leg as the checker; never linked into or called by the product library.
//
// Follows src/vkhr/scene_graph/hair_style.cc op for op in IEEE single
// precision (built with -O2 -fno-fast-math -ffp-contract=off on x86-64 SSE):
//   oracle_aabb                 generate_bounding_box + get_bounding_box  :215-255
//   oracle_generate_indices     generate_indices                          :196-213
//   oracle_generate_tangents    generate_tangents                         :171-194
//   oracle_voxelize_vertices    voxelize_vertices                         :257-294
//   oracle_voxelize_segments    voxelize_segments                         :296-342
//   oracle_quantize_tangents    the i8vec3 quantization loops             :283-291, 331-339
//   oracle_normalize            Volume::normalize                         :344-357
// glm semantics used (glm is not vendored; its documented definitions):
//   min(x, y) = y < x ? y : x;  max(x, y) = x < y ? y : x;
//   normalize(v) = v * (1 / sqrt(dot(v, v))), dot summed x, y, then z;
//   compMax(v) = max(max(v.x, v.y), v.z); floor = std::floor;
//   i8vec3(vec3) converts each component by truncation toward zero.
// Where the reference is undefined, the oracle defines it as the GPU build
// does: voxel coordinates below 0 (reachable only through float rounding
// when a vertex lies on the box's lower face) clamp to 0; 0/0 tangent
// averages of empty voxels quantize to 0; NaN quantizes to 0; a constant
// volume normalizes to all zero.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct vec3 {
    float x, y, z;
};
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
inline float gmin(float x, float y) { return y < x ? y : x; }
inline float gmax(float x, float y) { return x < y ? y : x; }
inline vec3 gmin(vec3 a, vec3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline vec3 gfloor(vec3 a) { return {std::floor(a.x), std::floor(a.y), std::floor(a.z)}; }
inline vec3 gabs(vec3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float comp_max(vec3 a) { return gmax(gmax(a.x, a.y), a.z); }
inline float dot(vec3 a, vec3 b) {
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    float s = x + y;
    return s + z;
}
inline vec3 normalize(vec3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return v * inv;
}
inline vec3 at(const float* p, uint64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

struct AABB {
    vec3 origin;
    float radius;
    vec3 size;
    float volume;
};
inline AABB from_array(const float* a) { return {{a[0], a[1], a[2]}, a[3], {a[4], a[5], a[6]}, a[7]}; }

int8_t to_i8(float v) {
    if (!(v == v)) return 0;
    return static_cast<int8_t>(static_cast<int>(v));
}

// Position of a voxel sample in the count arrays: glm::min(floor(p), res-1),
// then the float linear index of hair_style.cc:276 / :321.
uint64_t voxel_index(vec3 p, vec3 res_m1, uint64_t w, uint64_t h) {
    vec3 v = gmin(gfloor(p), res_m1);
    if (!(v.x > 0.0f)) v.x = 0.0f;  // lower clamp: undefined in the reference
    if (!(v.y > 0.0f)) v.y = 0.0f;
    if (!(v.z > 0.0f)) v.z = 0.0f;
    float fw = static_cast<float>(w), fh = static_cast<float>(h);
    float yw = v.y * fw;
    float zwh = (v.z * fw) * fh;
    float s = v.x + yw;
    s = s + zwh;
    return static_cast<uint64_t>(s);
}

}  // namespace

extern "C" {

void oracle_aabb(const float* xyz, uint64_t n, float* out) {
    vec3 mn = {0.0f, 0.0f, 0.0f}, mx = {0.0f, 0.0f, 0.0f};
    for (uint64_t i = 0; i < n; ++i) {
        vec3 p = at(xyz, i);
        mn.x = gmin(p.x, mn.x); mn.y = gmin(p.y, mn.y); mn.z = gmin(p.z, mn.z);
        mx.x = gmax(p.x, mx.x); mx.y = gmax(p.y, mx.y); mx.z = gmax(p.z, mx.z);
    }
    vec3 size = mx - mn;
    out[0] = mn.x; out[1] = mn.y; out[2] = mn.z;
    out[3] = std::sqrt(dot(size, size));  // glm::length
    out[4] = size.x; out[5] = size.y; out[6] = size.z;
    float xy = size.x * size.y;
    out[7] = xy * size.z;
}

void oracle_generate_indices(const uint32_t* offsets, uint32_t n_strands, uint32_t* out) {
    uint64_t k = 0;
    for (uint32_t s = 0; s < n_strands; ++s) {
        uint32_t seg = offsets[s + 1] - offsets[s] - 1;
        uint32_t vertex = offsets[s];
        for (uint32_t i = 0; i < seg; ++i) {
            out[k++] = vertex++;
            out[k++] = vertex;
        }
    }
}

void oracle_generate_tangents(const float* xyz, const uint32_t* offsets, uint32_t n_strands, float* out) {
    for (uint32_t s = 0; s < n_strands; ++s) {
        uint32_t seg = offsets[s + 1] - offsets[s] - 1;
        uint64_t v = offsets[s];
        vec3 last = {0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < seg; ++i, ++v) {
            last = normalize(at(xyz, v + 1) - at(xyz, v));
            out[3 * v] = last.x; out[3 * v + 1] = last.y; out[3 * v + 2] = last.z;
        }
        out[3 * v] = last.x; out[3 * v + 1] = last.y; out[3 * v + 2] = last.z;
    }
}

// counts: unsaturated uint32 counts (the GPU build's output).
// dens_u8: the reference's saturating uint8 densities.
// sums_ref: tangent sums with the reference's "only while != 255" rule.
// sums_all: tangent sums over every contribution (nullable each).
void oracle_voxelize_vertices(const float* xyz, uint64_t n, const float* tangents, const float* aabb,
                              uint32_t w, uint32_t h, uint32_t d, uint32_t* counts, uint8_t* dens_u8,
                              float* sums_ref, float* sums_all) {
    AABB b = from_array(aabb);
    vec3 res = {float(w), float(h), float(d)};
    vec3 voxel_size = b.size / res;
    vec3 res_m1 = {res.x - 1.0000f, res.y - 1.0000f, res.z - 1.0000f};
    uint64_t nv = uint64_t(w) * h * d;
    std::vector<uint8_t> dens(nv, 0);
    std::memset(counts, 0, nv * sizeof(uint32_t));
    if (sums_ref) std::memset(sums_ref, 0, 3 * nv * sizeof(float));
    if (sums_all) std::memset(sums_all, 0, 3 * nv * sizeof(float));
    for (uint64_t i = 0; i < n; ++i) {
        vec3 voxel = (at(xyz, i) - b.origin) / voxel_size;
        uint64_t pos = voxel_index(voxel, res_m1, w, h);
        counts[pos] += 1;
        if (tangents && sums_all) {
            vec3 t = at(tangents, i);
            sums_all[3 * pos] += t.x; sums_all[3 * pos + 1] += t.y; sums_all[3 * pos + 2] += t.z;
        }
        if (dens[pos] != 255) {
            if (tangents && sums_ref) {
                vec3 t = at(tangents, i);
                sums_ref[3 * pos] += t.x; sums_ref[3 * pos + 1] += t.y; sums_ref[3 * pos + 2] += t.z;
            }
            dens[pos] += 1;
        }
    }
    if (dens_u8) std::memcpy(dens_u8, dens.data(), nv);
}

void oracle_voxelize_segments(const float* xyz, const uint32_t* indices, uint64_t n_indices, const float* tangents,
                              const float* aabb, uint32_t w, uint32_t h, uint32_t d, uint32_t* counts,
                              uint8_t* dens_u8, float* sums_ref, float* sums_all) {
    AABB b = from_array(aabb);
    vec3 res = {float(w), float(h), float(d)};
    vec3 voxel_size = b.size / res;
    vec3 res_m1 = {res.x - 1.0f, res.y - 1.0f, res.z - 1.0f};
    uint64_t nv = uint64_t(w) * h * d;
    std::vector<uint8_t> dens(nv, 0);
    std::memset(counts, 0, nv * sizeof(uint32_t));
    if (sums_ref) std::memset(sums_ref, 0, 3 * nv * sizeof(float));
    if (sums_all) std::memset(sums_all, 0, 3 * nv * sizeof(float));
    for (uint64_t i = 0; i + 1 < n_indices; i += 2) {
        vec3 root = (at(xyz, indices[i]) - b.origin) / voxel_size;
        vec3 tip = (at(xyz, indices[i + 1]) - b.origin) / voxel_size;
        vec3 direction = tip - root;
        float steps = comp_max(gabs(direction));
        direction = direction / steps;  // [-1, 1]
        while (steps-- > 0.0f) {
            uint64_t pos = voxel_index(root, res_m1, w, h);
            counts[pos] += 1;
            if (tangents && sums_all) {
                vec3 t = at(tangents, indices[i]);
                sums_all[3 * pos] += t.x; sums_all[3 * pos + 1] += t.y; sums_all[3 * pos + 2] += t.z;
            }
            if (dens[pos] != 255) {
                if (tangents && sums_ref) {
                    vec3 t = at(tangents, indices[i]);
                    sums_ref[3 * pos] += t.x; sums_ref[3 * pos + 1] += t.y; sums_ref[3 * pos + 2] += t.z;
                }
                dens[pos] += 1;
            }
            root += direction;
        }
    }
    if (dens_u8) std::memcpy(dens_u8, dens.data(), nv);
}

// Count-only segment DDA split over the OpenMP team (atomic increments):
// the same samples as oracle_voxelize_segments, a multi-core CPU timing
// variant the reference does not have (its loop is serial).
void oracle_voxelize_segments_counts_omp(const float* xyz, const uint32_t* indices, uint64_t n_indices,
                                         const float* bounds8, uint32_t w, uint32_t h, uint32_t d,
                                         uint32_t* counts) {
    AABB b = from_array(bounds8);
    vec3 res = {float(w), float(h), float(d)};
    vec3 voxel_size = b.size / res;
    vec3 res_m1 = {res.x - 1.0f, res.y - 1.0f, res.z - 1.0f};
    const uint64_t nv = uint64_t(w) * h * d;
    std::memset(counts, 0, nv * sizeof(uint32_t));
#pragma omp parallel for schedule(static, 4096)
    for (int64_t p = 0; p < int64_t(n_indices / 2); ++p) {
        const uint64_t i = 2 * uint64_t(p);
        vec3 root = (at(xyz, indices[i]) - b.origin) / voxel_size;
        vec3 tip = (at(xyz, indices[i + 1]) - b.origin) / voxel_size;
        vec3 direction = tip - root;
        float steps = comp_max(gabs(direction));
        direction = direction / steps;
        while (steps-- > 0.0f) {
            uint64_t pos = voxel_index(root, res_m1, w, h);
#pragma omp atomic
            counts[pos] += 1;
            root += direction;
        }
    }
}

// i8vec3(sum / float(denominator) * 127), w = 0.  Over the OpenMP team with
// dynamic scheduling, as the reference's quantization loop
// (hair_style.cc:333-339); each element is independent, so the result does
// not depend on the team.
void oracle_quantize_tangents(const float* sums, const uint32_t* denominators, uint64_t n, int8_t* out) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < int64_t(n); ++i) {
        float den = static_cast<float>(denominators[i]);
        out[4 * i + 0] = to_i8(sums[3 * i + 0] / den * 127.0f);
        out[4 * i + 1] = to_i8(sums[3 * i + 1] / den * 127.0f);
        out[4 * i + 2] = to_i8(sums[3 * i + 2] / den * 127.0f);
        out[4 * i + 3] = 0;
    }
}

void oracle_normalize(uint8_t* densities, uint64_t n) {
    unsigned char data_min = 255, data_max = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (densities[i] > data_max) data_max = densities[i];
        if (densities[i] < data_min) data_min = densities[i];
    }
    if (data_max == data_min) {  // 255/0 -> NaN products: defined as zero
        std::memset(densities, 0, n);
        return;
    }
    float scaling = 255.0f / (data_max - data_min);
    for (uint64_t i = 0; i < n; ++i) {
        densities[i] -= data_min;
        densities[i] = static_cast<unsigned char>(densities[i] * scaling);
    }
}

}  // extern "C"
