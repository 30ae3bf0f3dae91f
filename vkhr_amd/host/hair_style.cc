// Host C++17 side of the hair-style call surface (include/vkhr/hair_style.hh)
// over the C ABI.  The public vectors stay the reference's data model; the
// generators, voxelizers and renderers run on the device.
#include "vkhr/hair_style.hh"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>

namespace vkhr {

namespace {

void check(int rc, const char* what) {
    if (rc != VKHR_OK) throw Error(std::string(what) + ": " + vkhr_last_error(), rc);
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e), VKHR_ERR_HIP);
}

}  // namespace

// ---------------------------------------------------------------- DeviceBuffer

DeviceBuffer::DeviceBuffer(size_t bytes) : bytes_(bytes) {
    if (bytes) check_hip(hipMalloc(&ptr_, bytes), "hipMalloc");
}

DeviceBuffer::DeviceBuffer(const void* host, size_t bytes) : DeviceBuffer(bytes) {
    if (bytes) upload(host, bytes);
}

DeviceBuffer::~DeviceBuffer() {
    if (ptr_) (void)hipFree(ptr_);
}

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = o.ptr_;
        bytes_ = o.bytes_;
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

void DeviceBuffer::download(void* host, size_t bytes) const {
    check_hip(hipMemcpy(host, ptr_, bytes, hipMemcpyDeviceToHost), "download");
}

void DeviceBuffer::upload(const void* host, size_t bytes) {
    check_hip(hipMemcpy(ptr_, host, bytes, hipMemcpyHostToDevice), "upload");
}

// ------------------------------------------------------------------ HairStyle

unsigned HairStyle::get_strand_count() const {
    return has_segments() ? static_cast<unsigned>(segments.size()) : file_header_.strand_count;
}
unsigned HairStyle::get_vertex_count() const { return static_cast<unsigned>(vertices.size()); }
unsigned HairStyle::get_segment_count() const { return get_vertex_count() - get_strand_count(); }

// First vertex of every strand (+ the end): from the segment array, or the
// header's default segment count for every strand.
std::vector<uint32_t> HairStyle::host_strand_offsets() const {
    const unsigned n = get_strand_count();
    std::vector<uint32_t> off(size_t(n) + 1, 0);
    for (unsigned s = 0; s < n; ++s) {
        const unsigned count = has_segments() ? segments[s] : file_header_.default_segment_count;
        off[s + 1] = off[s] + count + 1u;
    }
    return off;
}

void HairStyle::generate_thickness(float radius) {
    thickness.clear();
    thickness.reserve(get_vertex_count());
    const std::vector<uint32_t> off = host_strand_offsets();
    for (size_t s = 0; s + 1 < off.size(); ++s) {
        for (uint32_t i = off[s]; i + 1 < off[s + 1]; ++i) thickness.push_back(radius);
        thickness.push_back(0.0f);  // the tip
    }
    d_thickness_ = DeviceBuffer();
}

DeviceBuffer HairStyle::strand_offsets() const {
    const std::vector<uint32_t> off = host_strand_offsets();
    return DeviceBuffer(off.data(), off.size() * sizeof(uint32_t));
}

void HairStyle::generate_tangents() {
    DeviceBuffer off = strand_offsets();
    DeviceBuffer out(vertices.size() * sizeof(vec3));
    check(vkhr_generate_tangents(device_vertices().as<float>(), off.as<uint32_t>(), get_strand_count(),
                                 out.as<float>(), stream),
          "vkhr_generate_tangents");
    tangents.resize(vertices.size());
    out.download(tangents.data(), tangents.size() * sizeof(vec3));
    d_tangents_ = std::move(out);
}

void HairStyle::generate_indices() {
    DeviceBuffer off = strand_offsets();
    size_t n = 2 * size_t(get_segment_count());
    DeviceBuffer out(n ? n * sizeof(uint32_t) : sizeof(uint32_t));
    check(vkhr_generate_indices(off.as<uint32_t>(), get_strand_count(), out.as<uint32_t>(), stream),
          "vkhr_generate_indices");
    indices.resize(n);
    if (n) out.download(indices.data(), n * sizeof(uint32_t));
    d_indices_ = std::move(out);
}



void HairStyle::sort_indices_spatially() {
    const AABB box = get_bounding_box();
    const vkhr_aabb b{{box.origin.x, box.origin.y, box.origin.z}, box.radius, {box.size.x, box.size.y, box.size.z},
                      box.volume};
    const size_t n = indices.size();
    DeviceBuffer out(n ? n * sizeof(uint32_t) : sizeof(uint32_t));
    check(vkhr_sort_segments(device_vertices().as<float>(), get_vertex_count(), device_indices().as<uint32_t>(),
                             uint32_t(n), &b, out.as<uint32_t>(), stream),
          "vkhr_sort_segments");
    if (n) out.download(indices.data(), n * sizeof(uint32_t));
    d_indices_ = std::move(out);
}

const DeviceBuffer& HairStyle::device_vertices() const {
    if (!d_vertices_.get() && !vertices.empty()) d_vertices_ = DeviceBuffer(vertices.data(), vertices.size() * sizeof(vec3));
    return d_vertices_;
}
const DeviceBuffer& HairStyle::device_tangents() const {
    if (!d_tangents_.get() && !tangents.empty()) d_tangents_ = DeviceBuffer(tangents.data(), tangents.size() * sizeof(vec3));
    return d_tangents_;
}
const DeviceBuffer& HairStyle::device_thickness() const {
    if (!d_thickness_.get() && !thickness.empty())
        d_thickness_ = DeviceBuffer(thickness.data(), thickness.size() * sizeof(float));
    return d_thickness_;
}
const DeviceBuffer& HairStyle::device_indices() const {
    if (!d_indices_.get() && !indices.empty()) d_indices_ = DeviceBuffer(indices.data(), indices.size() * sizeof(unsigned));
    return d_indices_;
}
void HairStyle::invalidate_device() const {
    d_vertices_ = DeviceBuffer();
    d_tangents_ = DeviceBuffer();
    d_thickness_ = DeviceBuffer();
    d_indices_ = DeviceBuffer();
}

namespace {

HairStyle::Volume make_volume(const AABB& box, std::size_t w, std::size_t h, std::size_t d) {
    HairStyle::Volume v;
    v.resolution = vec3{float(w), float(h), float(d)};
    v.bounds = box;
    size_t n = w * h * d;
    v.counts = DeviceBuffer(n * sizeof(uint32_t));
    v.tangent_sums = DeviceBuffer(2 * n * sizeof(int64_t));
    v.density = DeviceBuffer(n);
    v.tangent = DeviceBuffer(4 * n);
    v.cells = DeviceBuffer((w + 1) * (h + 1) * (d + 1) * sizeof(uint64_t));
    v.occupancy = DeviceBuffer(vkhr_occupancy_words(uint32_t(w), uint32_t(h), uint32_t(d)) * sizeof(uint32_t));
    return v;
}

vkhr_aabb to_c(const AABB& b) {
    vkhr_aabb a;
    a.origin[0] = b.origin.x; a.origin[1] = b.origin.y; a.origin[2] = b.origin.z;
    a.radius = b.radius;
    a.size[0] = b.size.x; a.size[1] = b.size.y; a.size[2] = b.size.z;
    a.volume = b.volume;
    return a;
}

void finalize(HairStyle::Volume& v, int normalize, vkhr_stream stream) {
    check(vkhr_volume_finalize(v.counts.as<uint32_t>(), v.tangent_sums.as<int64_t>(), uint32_t(v.resolution.x),
                               uint32_t(v.resolution.y), uint32_t(v.resolution.z), normalize, v.density.as<uint8_t>(),
                               v.tangent.as<int8_t>(), nullptr, v.cells.as<uint64_t>(),
                               v.occupancy.as<uint32_t>(), stream),
          "vkhr_volume_finalize");
}

}  // namespace

HairStyle::Volume HairStyle::voxelize_vertices(std::size_t width, std::size_t height, std::size_t depth) const {
    Volume v = make_volume(get_bounding_box(), width, height, depth);
    vkhr_aabb b = to_c(v.bounds);
    check(vkhr_voxelize_vertices(device_vertices().as<float>(), get_vertex_count(), device_tangents().as<float>(), &b,
                                 uint32_t(width), uint32_t(height), uint32_t(depth), v.counts.as<uint32_t>(),
                                 v.tangent_sums.as<int64_t>(), stream),
          "vkhr_voxelize_vertices");
    finalize(v, 0, stream);
    return v;
}

HairStyle::Volume HairStyle::voxelize_segments(std::size_t width, std::size_t height, std::size_t depth) const {
    Volume v = make_volume(get_bounding_box(), width, height, depth);
    vkhr_aabb b = to_c(v.bounds);
    check(vkhr_voxelize_segments(device_vertices().as<float>(), get_vertex_count(), device_indices().as<uint32_t>(),
                                 uint32_t(indices.size()), device_tangents().as<float>(), &b, uint32_t(width),
                                 uint32_t(height), uint32_t(depth), v.counts.as<uint32_t>(), v.tangent_sums.as<int64_t>(),
                                 stream),
          "vkhr_voxelize_segments");
    finalize(v, 0, stream);
    return v;
}

void HairStyle::Volume::normalize() {
    finalize(*this, 1, nullptr);
    normalized = true;
}

void HairStyle::Volume::download() {
    size_t n = size_t(resolution.x) * size_t(resolution.y) * size_t(resolution.z);
    densities.resize(n);
    tangents.resize(n);
    density.download(densities.data(), n);
    tangent.download(tangents.data(), 4 * n);
}

vkhr_volume HairStyle::Volume::view() const {
    vkhr_volume v;
    std::memset(&v, 0, sizeof(v));
    v.density = density.as<uint8_t>();
    v.cells = cells.as<uint64_t>();
    v.occupancy = occupancy.as<uint32_t>();
    v.tangents = tangent.as<int8_t>();
    v.nx = uint32_t(resolution.x);
    v.ny = uint32_t(resolution.y);
    v.nz = uint32_t(resolution.z);
    return v;
}

// ------------------------------------------------------------------ renderers

void Raymarcher::draw(const HairStyle::Volume& volume, const vkhr_camera& camera, const vkhr_light& light,
                      const vkhr_strand_params& strand, const vkhr_render_params& params, uint32_t width,
                      uint32_t height, float* rgba, float* depth, vkhr_stream stream) const {
    vkhr_volume v = volume.view();
    check(vkhr_raymarch(&v, &camera, &light, &strand, &params, width, height, nullptr, rgba, depth, nullptr, stream),
          "vkhr_raymarch");
}

void Raymarcher::draw(const HairStyle::Volume& volume, const vkhr_camera& camera, const vkhr_light& light,
                      const vkhr_strand_params& strand, const vkhr_render_params& params, LinkedList& ppll,
                      const float* depth_buffer, vkhr_stream stream) const {
    vkhr_volume v = volume.view();
    const vkhr_ppll& p = ppll.view();
    check(vkhr_raymarch_ppll(&v, &camera, &light, &strand, &params, p.width, p.height, depth_buffer, &p, stream),
          "vkhr_raymarch_ppll");
}

LinkedList::LinkedList(uint32_t width, uint32_t height, std::size_t node_count)
    : heads_(size_t(width) * height * sizeof(uint32_t)),
      nodes_(node_count * NodeSize),
      counter_(sizeof(uint32_t)) {
    ppll_.heads = heads_.as<uint32_t>();
    ppll_.nodes = nodes_.as<uint32_t>();
    ppll_.counter = counter_.as<uint32_t>();
    ppll_.node_count = uint32_t(node_count);
    ppll_.width = width;
    ppll_.height = height;
}

void LinkedList::clear(vkhr_stream stream) { check(vkhr_ppll_clear(&ppll_, stream), "vkhr_ppll_clear"); }

void LinkedList::resolve(float* out_rgba, const float* background, vkhr_stream stream) const {
    check(vkhr_ppll_resolve(&ppll_, background, out_rgba, stream), "vkhr_ppll_resolve");
}

uint32_t LinkedList::node_count_used() const {
    uint32_t n = 0;
    check_hip(hipDeviceSynchronize(), "synchronize");
    counter_.download(&n, sizeof(n));
    return n;
}

void StrandRasterizer::draw_depth(const HairStyle& style, const float light_vp[16], float line_width, uint32_t size,
                                  float* depth_map, vkhr_stream stream) const {
    check(vkhr_shadow_depth(style.device_vertices().as<float>(), style.get_vertex_count(),
                            style.device_indices().as<uint32_t>(), uint32_t(style.indices.size()), light_vp,
                            line_width, size, size, depth_map, stream),
          "vkhr_shadow_depth");
}

void StrandRasterizer::draw(const HairStyle& style, const HairStyle::Volume& volume, const vkhr_camera& camera,
                            const vkhr_light& light, const vkhr_strand_params& strand,
                            const vkhr_render_params& params, const float* shadow_map, uint32_t shadow_size,
                            LinkedList& ppll, vkhr_stream stream) const {
    vkhr_volume v = volume.view();
    check(vkhr_raster_strands(style.device_vertices().as<float>(), style.device_tangents().as<float>(),
                              style.device_thickness().as<float>(), style.get_vertex_count(),
                              style.device_indices().as<uint32_t>(), uint32_t(style.indices.size()), &camera, &light,
                              &strand, &params, &v, shadow_map, shadow_size, shadow_size, nullptr, &ppll.view(),
                              nullptr, stream),
          "vkhr_raster_strands");
}

void HybridRenderer::update(const vkhr_camera& camera, const vkhr_render_params& params) {
    // glm::smoothstep(edge0, edge1, x) in float
    const float e0 = params.lod_magnified_distance, e1 = params.lod_minified_distance;
    float t = (camera.look_at_distance - e0) / (e1 - e0);
    t = std::min(std::max(t, 0.0f), 1.0f);
    level_of_detail = t * t * (3.0f - 2.0f * t);
}

bool HybridRenderer::rasterizer_enabled(const vkhr_render_params& params) const {
    return params.renderer == Renderer::Rasterizer ||
           (params.renderer == Renderer::Hybrid_LoD && level_of_detail != 1.0f);
}

bool HybridRenderer::raymarcher_enabled(const vkhr_render_params& params) const {
    return params.renderer == Renderer::Raymarcher ||
           (params.renderer == Renderer::Hybrid_LoD && level_of_detail != 0.0f);
}

void HybridRenderer::draw_color(const HairStyle& style, const HairStyle::Volume& volume, const vkhr_camera& camera,
                                const vkhr_light& light, const vkhr_strand_params& strand,
                                const vkhr_render_params& params, const float* shadow_map, uint32_t shadow_size,
                                LinkedList& ppll, float* out_rgba, const float* background,
                                vkhr_stream stream) const {
    ppll.clear(stream);
    if (rasterizer_enabled(params))
        strands_.draw(style, volume, camera, light, strand, params, shadow_map, shadow_size, ppll, stream);
    if (raymarcher_enabled(params)) raymarcher_.draw(volume, camera, light, strand, params, ppll, nullptr, stream);
    ppll.resolve(out_rgba, background, stream);
}

}  // namespace vkhr
