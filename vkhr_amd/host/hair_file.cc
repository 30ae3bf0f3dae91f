// The .hair file side of the host HairStyle: load / save of the 128-byte
// vkhr header and the arrays it announces (src/vkhr/scene_graph/
// hair_style.cc:24-70, 531-665; include/vkhr/scene_graph/hair_style.hh:
// 28-57, 147-174), the header accessors, the bounding box kept in the
// header (hair_style.cc:215-255), and the strand shuffle / reduction
// (hair_style.cc:371-449).  Host-only: nothing here touches the GPU.
#include "vkhr/hair_style.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

namespace vkhr {

namespace {

// Reads n elements; fails (without allocating) when the file holds fewer.
template <class T>
bool read_array(std::ifstream& file, std::vector<T>& v, std::size_t n) {
    const std::streampos here = file.tellg();
    file.seekg(0, std::ios::end);
    const std::streamoff left = file.tellg() - here;
    file.seekg(here);
    if (!file || left < 0 || std::uint64_t(left) / sizeof(T) < n) return false;
    v.resize(n);
    if (n == 0) return true;
    return bool(file.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(T))));
}

template <class T>
bool write_array(std::ofstream& file, const std::vector<T>& v) {
    if (v.empty()) return true;
    return bool(file.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T))));
}

// xorshift64 (hair_style.cc:679-685).
std::uint64_t xorshift64(std::uint64_t* state) {
    std::uint64_t x = *state;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return *state = x;
}

}  // namespace

HairStyle::HairStyle(const std::string& file_path) { load(file_path); }

HairStyle::operator bool() const { return error_state_ == Error::None; }

HairStyle::Error HairStyle::get_last_error_state() const { return error_state_; }

bool HairStyle::set_error_state(Error e) const {
    error_state_ = e;
    return e == Error::None;
}

bool HairStyle::valid_signature() const {
    const char* s = file_header_.signature;
    return s[0] == 'H' && s[1] == 'A' && s[2] == 'I' && s[3] == 'R';
}

// hair_style.cc:537-544: vertices present, per-vertex arrays of vertex length.
bool HairStyle::format_is_valid() const {
    if (!has_vertices()) return false;
    if (!valid_signature()) return false;
    if (has_thickness() && thickness.size() != vertices.size()) return false;
    if (has_transparency() && transparency.size() != vertices.size()) return false;
    if (has_color() && color.size() != vertices.size()) return false;
    return true;
}

// hair_style.cc:546-567: signature, the bit field from the arrays present
// (the bounding-box bit is left as it is), the counts.
void HairStyle::complete_header() const {
    std::memcpy(file_header_.signature, "HAIR", 4);
    std::uint32_t f = file_header_.field & kHasBoundingBox;
    if (has_segments()) f |= kHasSegments;
    if (has_vertices()) f |= kHasVertices;
    if (has_thickness()) f |= kHasThickness;
    if (has_transparency()) f |= kHasTransparency;
    if (has_color()) f |= kHasColor;
    if (has_tangents()) f |= kHasTangents;
    if (has_indices()) f |= kHasIndices;
    file_header_.field = f;  // future_extension bits cleared
    file_header_.strand_count = get_strand_count();
    file_header_.vertex_count = get_vertex_count();
}

bool HairStyle::load(const std::string& file_path) {
    std::ifstream file(file_path, std::ios::binary);
    if (!file) return set_error_state(Error::OpeningFile);
    FileHeader h;
    if (!file.read(reinterpret_cast<char*>(&h), sizeof(FileHeader))) return set_error_state(Error::ReadingFileHeader);
    file_header_ = h;
    if (!valid_signature()) return set_error_state(Error::InvalidSignature);
    segments.clear(); vertices.clear(); thickness.clear(); transparency.clear();
    color.clear(); tangents.clear(); indices.clear();
    invalidate_device();
    const std::uint32_t f = h.field;
    if ((f & kHasSegments) && !read_array(file, segments, h.strand_count)) return set_error_state(Error::ReadingSegments);
    if ((f & kHasVertices) && !read_array(file, vertices, h.vertex_count)) return set_error_state(Error::ReadingVertices);
    if ((f & kHasThickness) && !read_array(file, thickness, h.vertex_count))
        return set_error_state(Error::ReadingThickness);
    if ((f & kHasTransparency) && !read_array(file, transparency, h.vertex_count))
        return set_error_state(Error::ReadingTransparency);
    if ((f & kHasColor) && !read_array(file, color, h.vertex_count)) return set_error_state(Error::ReadingColor);
    if ((f & kHasTangents) && !read_array(file, tangents, h.vertex_count)) return set_error_state(Error::ReadingTangents);
    // the index count follows from the counts read so far (hair_style.cc:628-633);
    // fewer vertices than strands leave no valid count
    if ((f & kHasIndices) && (get_vertex_count() < get_strand_count() ||
                              !read_array(file, indices, std::size_t(get_segment_count()) * 2)))
        return set_error_state(Error::ReadingIndices);
    if (!format_is_valid()) return set_error_state(Error::InvalidFormat);
    return set_error_state(Error::None);
}

bool HairStyle::save(const std::string& file_path) const {
    complete_header();
    if (!format_is_valid()) return set_error_state(Error::InvalidFormat);
    std::ofstream file(file_path, std::ios::binary);
    if (!file) return set_error_state(Error::OpeningFile);
    if (!file.write(reinterpret_cast<const char*>(&file_header_), sizeof(FileHeader)))
        return set_error_state(Error::WritingFileHeader);
    if (!write_array(file, segments)) return set_error_state(Error::WritingSegments);
    if (!write_array(file, vertices)) return set_error_state(Error::WritingVertices);
    if (!write_array(file, thickness)) return set_error_state(Error::WritingThickness);
    if (!write_array(file, transparency)) return set_error_state(Error::WritingTransparency);
    if (!write_array(file, color)) return set_error_state(Error::WritingColor);
    if (!write_array(file, tangents)) return set_error_state(Error::WritingTangents);
    if (!write_array(file, indices)) return set_error_state(Error::WritingIndices);
    return set_error_state(Error::None);
}

// The bytes save() writes.  (The reference's get_size, hair_style.cc:
// 664-675, leaves the transparency array out.)
std::size_t HairStyle::get_size() const {
    return sizeof(FileHeader) + segments.size() * sizeof(segments[0]) + vertices.size() * sizeof(vertices[0]) +
           thickness.size() * sizeof(thickness[0]) + transparency.size() * sizeof(transparency[0]) +
           color.size() * sizeof(color[0]) + tangents.size() * sizeof(tangents[0]) +
           indices.size() * sizeof(indices[0]);
}

bool HairStyle::has_bounding_box() const { return (file_header_.field & kHasBoundingBox) != 0; }

unsigned HairStyle::get_default_segment_count() const { return file_header_.default_segment_count; }
void HairStyle::set_default_segment_count(unsigned n) { file_header_.default_segment_count = n; }
float HairStyle::get_default_thickness() const { return file_header_.default_thickness; }
void HairStyle::set_default_thickness(float t) { file_header_.default_thickness = t; }
float HairStyle::get_default_transparency() const { return file_header_.default_transparency; }
void HairStyle::set_default_transparency(float t) { file_header_.default_transparency = t; }
vec3 HairStyle::get_default_color() const {
    return vec3{file_header_.default_color[0], file_header_.default_color[1], file_header_.default_color[2]};
}
void HairStyle::set_default_color(const vec3& c) {
    file_header_.default_color[0] = c.x;
    file_header_.default_color[1] = c.y;
    file_header_.default_color[2] = c.z;
}
const char* HairStyle::get_information() const { return file_header_.information; }
// hair_style.cc:146-152: zero-filled, truncated to 64 bytes (not terminated
// when it fills them).
void HairStyle::set_information(const std::string& information) {
    std::memset(file_header_.information, 0, sizeof(file_header_.information));
    std::memcpy(file_header_.information, information.data(),
                std::min(sizeof(file_header_.information), information.size()));
}
void HairStyle::set_strand_count(unsigned strand_count) { file_header_.strand_count = strand_count; }

// hair_style.cc:215-234: component min / max over the vertices with both
// started at the origin (so the box contains it), stored in the header.
void HairStyle::generate_bounding_box() {
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    for (const vec3& p : vertices) {
        const float c[3] = {p.x, p.y, p.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(c[k], lo[k]);  // glm::min(x, y) = y < x ? y : x
            hi[k] = std::max(c[k], hi[k]);
        }
    }
    std::memcpy(file_header_.bounding_box_min, lo, sizeof(lo));
    std::memcpy(file_header_.bounding_box_max, hi, sizeof(hi));
    file_header_.field |= kHasBoundingBox;
}

// hair_style.cc:236-255.
AABB HairStyle::get_bounding_box() const {
    const float* lo = file_header_.bounding_box_min;
    const float* hi = file_header_.bounding_box_max;
    const vec3 size{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const float dot = size.x * size.x + size.y * size.y + size.z * size.z;
    return AABB{vec3{lo[0], lo[1], lo[2]}, std::sqrt(dot), size, size.x * size.y * size.z};
}

void HairStyle::shuffle() { reduce(1.0f); }

// hair_style.cc:375-449: strands drawn at random (swap with the last, pop)
// into a new order; ratio 1 keeps every strand (the count expression wraps),
// ratio r < 1 keeps (count - ceil(count * r)) - 1 strands, as the
// reference's loop does.  A draw of exactly 1.0 (which would index one past
// the end in the reference) takes the last strand.  The index array is then
// rebuilt from the new strand order.
void HairStyle::reduce(float ratio) {
    const unsigned n = get_strand_count();
    unsigned strands_left = static_cast<unsigned>(float(n) - std::ceil(float(n) * ratio));
    std::vector<std::size_t> strand_offset(n);
    {
        std::vector<uint32_t> off = host_strand_offsets();
        for (unsigned s = 0; s < n; ++s) strand_offset[s] = off[s];
    }
    std::vector<unsigned short> r_segments;
    std::vector<vec3> r_vertices, r_tangents, r_color;
    std::vector<float> r_thickness, r_transparency;
    while (--strands_left && !strand_offset.empty()) {
        const double random = double(xorshift64(&seed)) / double(std::numeric_limits<std::uint64_t>::max());
        std::size_t pick = static_cast<std::size_t>(random * double(strand_offset.size()));
        pick = std::min(pick, strand_offset.size() - 1);
        unsigned count = file_header_.default_segment_count;
        if (has_segments()) {
            count = segments[pick];
            std::swap(segments[pick], segments.back());
            segments.pop_back();
        }
        r_segments.push_back(static_cast<unsigned short>(count));
        const std::size_t a = strand_offset[pick], b = a + count + 1u;
        auto take = [a, b](const auto& src, auto& dst) {
            if (!src.empty()) dst.insert(dst.end(), src.begin() + a, src.begin() + b);
        };
        take(vertices, r_vertices);
        take(thickness, r_thickness);
        take(tangents, r_tangents);
        take(transparency, r_transparency);
        take(color, r_color);
        std::swap(strand_offset[pick], strand_offset.back());
        strand_offset.pop_back();
    }
    segments = std::move(r_segments);
    vertices = std::move(r_vertices);
    thickness = std::move(r_thickness);
    tangents = std::move(r_tangents);
    transparency = std::move(r_transparency);
    color = std::move(r_color);
    // generate_indices (hair_style.cc:196-213, called by the reference's
    // reduce), on the host
    indices.clear();
    uint32_t v = 0;
    for (unsigned short count : segments) {
        for (unsigned k = 0; k < count; ++k) {
            indices.push_back(v);
            indices.push_back(v + 1);
            ++v;
        }
        ++v;
    }
    invalidate_device();
}

}  // namespace vkhr
