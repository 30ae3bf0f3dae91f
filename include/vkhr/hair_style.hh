// Host C++17 mirror of the reference's hair-style call surface
// (include/vkhr/scene_graph/hair_style.hh:16-210 and the GPU-side classes of
// include/vkhr/rasterizer/{hair_style,volume,linked_list}.hh), implemented on
// the MI355X C ABI (vkhr_hip.h).  Names and argument meanings follow the
// reference; glm types are replaced by small PODs (glm is not a dependency).
// Errors: the voxelizers and renderers throw vkhr::Error on a failed C-ABI
// call (the reference's voxelizers have no error path; its Vulkan wrapper
// throws vkpp::Exception).
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "vkhr_hip.h"

namespace vkhr {

struct vec3 {
    float x, y, z;
};

// struct AABB (hair_style.hh:16-21)
struct AABB {
    vec3 origin;
    float radius;
    vec3 size;
    float volume;
};

struct Error : std::runtime_error {
    int code;
    Error(const std::string& what, int c) : std::runtime_error(what), code(c) {}
};

// A device allocation owned by its holder (move-only), as the reference's
// vk::DeviceImage / vk::StorageBuffer RAII objects.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    DeviceBuffer(const void* host, size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept;
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    void* get() const { return ptr_; }
    template <class T> T* as() const { return static_cast<T*>(ptr_); }
    size_t size() const { return bytes_; }
    void download(void* host, size_t bytes) const;
    void upload(const void* host, size_t bytes);

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

class HairStyle final {
public:
    // HairStyle::Error (hair_style.hh:28-54): the state of the last load/save.
    enum class Error {
        None,

        OpeningFile,
        ReadingFileHeader,
        WritingFileHeader,

        InvalidSignature,

        ReadingSegments,
        ReadingVertices,
        ReadingThickness,
        ReadingTransparency,
        ReadingColor,
        ReadingTangents,
        ReadingIndices,

        WritingSegments,
        WritingVertices,
        WritingThickness,
        WritingTransparency,
        WritingColor,
        WritingTangents,
        WritingIndices,

        InvalidFormat
    };

    HairStyle() = default;
    explicit HairStyle(const std::string& file_path);    // load(file_path)

    explicit operator bool() const;                        // no error state
    Error get_last_error_state() const;

    // The .hair format (Cem Yuksel's HAIR files with vkhr's 128-byte header:
    // hair_style.hh:147-174, hair_style.cc:24-70, 576-665): the header, then
    // the arrays its bit field announces, in the order segments (uint16 per
    // strand), vertices, thickness, transparency, colour, tangents, indices.
    bool load(const std::string& file_path);
    bool save(const std::string& file_path) const;
    std::size_t get_size() const;                          // bytes save() writes

    // The strand order shuffle and the random strand reduction
    // (hair_style.cc:371-449), driven by xorshift64 of `seed` (the reference
    // seeds it from std::random_device; here it is a member the caller sets).
    void shuffle();
    void reduce(float ratio);
    std::uint64_t seed = 0x2545f4914f6cdd1dull;

    bool has_segments() const { return !segments.empty(); }
    bool has_vertices() const { return !vertices.empty(); }
    bool has_thickness() const { return !thickness.empty(); }
    bool has_transparency() const { return !transparency.empty(); }
    bool has_color() const { return !color.empty(); }
    bool has_tangents() const { return !tangents.empty(); }
    bool has_indices() const { return !indices.empty(); }
    bool has_bounding_box() const;                         // from the header

    unsigned get_default_segment_count() const;
    void set_default_segment_count(unsigned default_segment_count);
    float get_default_thickness() const;
    void set_default_thickness(float default_thickness);
    float get_default_transparency() const;
    void set_default_transparency(float default_transparency);
    vec3 get_default_color() const;
    void set_default_color(const vec3& default_color);
    const char* get_information() const;
    void set_information(const std::string& information);
    void set_strand_count(unsigned strand_count);          // for styles without a segment array

    // HairStyle::Volume (hair_style.hh:89-101): the host vectors of the
    // reference plus the device volumes the renderers read.
    struct Volume {
        vec3 resolution;
        AABB bounds;
        std::vector<unsigned char> densities;            // filled by download()
        std::vector<std::array<int8_t, 4>> tangents;     // filled by download()

        void normalize();                                // Volume::normalize (hair_style.cc:344-357)
        void download();                                 // device -> the host vectors
        vkhr_volume view() const;                        // what the kernels take

        DeviceBuffer counts;        // uint32 unsaturated counts
        DeviceBuffer tangent_sums;  // fixed-point tangent accumulator (2 int64 per voxel)
        DeviceBuffer density;       // uint8 (min(count, 255), then normalized)
        DeviceBuffer tangent;       // int8x4
        DeviceBuffer cells;         // corner-packed density (raymarch layout)
        DeviceBuffer occupancy;     // brick occupancy of the cells (empty bricks are skipped)
        bool normalized = false;
    };

    unsigned get_strand_count() const;     // segments.size(), or the header's without a segment array
    unsigned get_segment_count() const;    // vertices - strands
    unsigned get_vertex_count() const;

    void generate_thickness(float radius);   // hair_style.cc:155-169
    void generate_tangents();                // hair_style.cc:171-194 (on the device)
    void generate_indices();                 // hair_style.cc:196-213 (on the device)
    void generate_bounding_box();            // hair_style.cc:215-234 (into the header, as the reference)
    // The index pairs in the Morton order of the segment midpoints in the
    // bounding box (vkhr_sort_segments; no counterpart in the reference): a
    // permutation on which every pass gives the same results, with the
    // strand passes faster for styles stored in random strand order.  After
    // generate_indices and generate_bounding_box.
    void sort_indices_spatially();
    AABB get_bounding_box() const;           // hair_style.cc:236-255 (from the header)

    Volume voxelize_vertices(std::size_t width, std::size_t height, std::size_t depth) const;
    Volume voxelize_segments(std::size_t width, std::size_t height, std::size_t depth) const;

    // Public data, as in the reference (hair_style.hh:126-139).
    std::vector<unsigned short> segments;
    std::vector<vec3> vertices;
    std::vector<float> thickness;
    std::vector<float> transparency;
    std::vector<vec3> color;
    std::vector<vec3> tangents;
    std::vector<unsigned> indices;

    // Device copies, uploaded on demand from the vectors above.
    const DeviceBuffer& device_vertices() const;
    const DeviceBuffer& device_tangents() const;
    const DeviceBuffer& device_thickness() const;
    const DeviceBuffer& device_indices() const;
    void invalidate_device() const;          // after editing the public vectors

    vkhr_stream stream = nullptr;

    // The 128-byte file header (hair_style.hh:147-174); the bit field is
    // kept as its uint32 word (bit 0 has_segments ... bit 7
    // has_bounding_box, the order of the reference's field declaration).
    struct FileHeader {
        char signature[4];
        std::uint32_t strand_count, vertex_count;
        std::uint32_t field;
        std::uint32_t default_segment_count;
        float default_thickness, default_transparency;
        float default_color[3];
        char information[64];
        float bounding_box_min[3];
        float bounding_box_max[3];
    };
    static_assert(sizeof(FileHeader) == 128, "the .hair header is 128 bytes");
    enum FieldBit : std::uint32_t {
        kHasSegments = 1u << 0, kHasVertices = 1u << 1, kHasThickness = 1u << 2, kHasTransparency = 1u << 3,
        kHasColor = 1u << 4, kHasTangents = 1u << 5, kHasIndices = 1u << 6, kHasBoundingBox = 1u << 7,
    };
    const FileHeader& header() const { return file_header_; }

private:
    DeviceBuffer strand_offsets() const;
    std::vector<std::uint32_t> host_strand_offsets() const;
    bool set_error_state(Error e) const;
    bool valid_signature() const;
    bool format_is_valid() const;
    void complete_header() const;
    mutable FileHeader file_header_{};
    mutable Error error_state_ = Error::None;
    mutable DeviceBuffer d_vertices_, d_tangents_, d_thickness_, d_indices_;
};

// vkhr::Renderer::Type (include/vkhr/renderer.hh:13-18): the values of
// vkhr_render_params::renderer.
struct Renderer {
    enum Type : int { Rasterizer = 0, Ray_Tracer = 1, Raymarcher = 2, Hybrid_LoD = 3 };
};

class LinkedList;

// The raymarcher pass (vulkan::Volume::draw + volume.frag) for one style.
class Raymarcher {
public:
    // Writes W*H float4 (clamped fragment colour) and W*H depth into device
    // buffers; depth may be null.
    void draw(const HairStyle::Volume& volume, const vkhr_camera& camera, const vkhr_light& light,
              const vkhr_strand_params& strand, const vkhr_render_params& params, uint32_t width, uint32_t height,
              float* rgba, float* depth, vkhr_stream stream = nullptr) const;
    // Links one node per raymarched fragment into the PPLL (the raymarch
    // subpass of Rasterizer::draw_color, rasterizer.cc:254-258, with
    // volume.frag:95-100); depth_buffer (W*H, nullable) is the depth test input.
    void draw(const HairStyle::Volume& volume, const vkhr_camera& camera, const vkhr_light& light,
              const vkhr_strand_params& strand, const vkhr_render_params& params, LinkedList& ppll,
              const float* depth_buffer = nullptr, vkhr_stream stream = nullptr) const;
};

// vulkan::LinkedList (linked_list.hh:16-45): PPLL storage, clear, resolve.
class LinkedList {
public:
    LinkedList(uint32_t width, uint32_t height, std::size_t node_count);
    void clear(vkhr_stream stream = nullptr);
    void resolve(float* out_rgba, const float* background = nullptr, vkhr_stream stream = nullptr) const;
    uint32_t node_count_used() const;     // synchronous read of the counter
    const vkhr_ppll& view() const { return ppll_; }

    static constexpr std::size_t AverageFragmentsPerPixel = 32;
    static constexpr std::size_t NodeSize = 12;

private:
    DeviceBuffer heads_, nodes_, counter_;
    vkhr_ppll ppll_{};
};

// The strand passes: Rasterizer::draw_depth and vulkan::HairStyle::draw.
class StrandRasterizer {
public:
    void draw_depth(const HairStyle& style, const float light_vp[16], float line_width, uint32_t size,
                    float* depth_map, vkhr_stream stream = nullptr) const;
    void draw(const HairStyle& style, const HairStyle::Volume& volume, const vkhr_camera& camera,
              const vkhr_light& light, const vkhr_strand_params& strand, const vkhr_render_params& params,
              const float* shadow_map, uint32_t shadow_size, LinkedList& ppll, vkhr_stream stream = nullptr) const;
};

// The colour pass of vkhr::Rasterizer for one style, with the hybrid LoD.
// update() takes the level of detail from the camera distance: glm::smoothstep
// of the two LoD distances in float (rasterizer.cc:162-168).  draw_color()
// clears the PPLL, draws the strands when the rasterizer is enabled at that
// level, raymarches into the same list when the raymarcher is, and resolves
// (rasterizer.cc:232-282).  The enable rules are Interface::rasterizer_enabled
// and raymarcher_enabled (interface.cc:534-540): the rasterizer draws for
// Renderer::Rasterizer, or for Hybrid_LoD below level 1; the raymarcher draws
// for Renderer::Raymarcher, or for Hybrid_LoD above level 0.  Each pass
// scales its own coverage by lod() (scheme.glsl:7-11): strands by 1 - lod
// (strand.frag:42), the raymarch by lod (volume.frag:48).
class HybridRenderer {
public:
    void update(const vkhr_camera& camera, const vkhr_render_params& params);
    bool rasterizer_enabled(const vkhr_render_params& params) const;
    bool raymarcher_enabled(const vkhr_render_params& params) const;
    // The colour pass into out_rgba (W*H float4, the PPLL's size); background
    // as vkhr_ppll_resolve (nullable: opaque white).
    void draw_color(const HairStyle& style, const HairStyle::Volume& volume, const vkhr_camera& camera,
                    const vkhr_light& light, const vkhr_strand_params& strand, const vkhr_render_params& params,
                    const float* shadow_map, uint32_t shadow_size, LinkedList& ppll, float* out_rgba,
                    const float* background = nullptr, vkhr_stream stream = nullptr) const;

    float level_of_detail = 0.0f;

private:
    StrandRasterizer strands_;
    Raymarcher raymarcher_;
};

}  // namespace vkhr
