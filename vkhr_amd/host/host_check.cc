// Drives the host C++ surface end to end on one style, the way the
// reference's load and frame do (HairStyle generators -> voxelize_segments ->
// Volume::normalize -> raymarch; shadow depth -> strands -> PPLL resolve),
// and writes the results for comparison with the oracle.
//
//   host_check <input.bin> <output-dir>
//   host_check --save <input.bin> <out.hair>
//   host_check --hair <style.hair> <seed> <params.bin> <output-dir>
//
// input.bin (little endian): u32 n_strands, n_vertices, volume_n, width,
// height, shadow_size; u16 segments[n_strands]; f32 xyz[3*n_vertices];
// then the raw vkhr_camera, vkhr_light, vkhr_strand_params and
// vkhr_render_params blocks (the strand block's bounds are replaced by the
// style's own AABB), then f32 light_vp[16] (unbiased light view-projection,
// column-major) for the shadow-map pass; optionally u32 n, then n pairs of
// f32 (lod_magnified_distance, lod_minified_distance), each drawn as one
// hybrid LoD colour pass into hybrid<k>.bin, with hybrid<k>.txt holding the
// level of detail, the two enable flags and the PPLL nodes used.
//
// --save writes the style of input.bin (segments and vertices) as a .hair
// file with HairStyle::save.  --hair loads a .hair file with
// HairStyle(path) and prepares it as SceneGraph::add_style does
// (src/vkhr/scene_graph.cc:222-245: shuffle with the given seed, then the
// tangents, thickness, indices and bounding box the file does not hold),
// then runs the pipeline above; params.bin holds u32 volume_n, width,
// height, shadow_size and the blocks that follow the vertices in input.bin.
// It also writes the prepared style's vertices, tangents and indices.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "vkhr/hair_style.hh"

namespace {

template <class T>
void read(std::ifstream& f, T* p, size_t n) {
    if (!f.read(reinterpret_cast<char*>(p), sizeof(T) * n)) throw std::runtime_error("short input file");
}

void write(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f.write(static_cast<const char*>(p), bytes)) throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    const bool save = mode == "--save", hair = mode == "--hair";
    if (!((argc == 3 && !save && !hair) || (save && argc == 4) || (hair && argc == 6))) {
        std::fprintf(stderr,
                     "usage: %s <input.bin> <output-dir> | --save <input.bin> <out.hair> | "
                     "--hair <style.hair> <seed> <params.bin> <output-dir>\n",
                     argv[0]);
        return 2;
    }
    try {
        vkhr::HairStyle style;
        uint32_t n = 0, W = 0, H = 0, S = 0;
        std::ifstream in;
        std::string out;
        if (hair) {
            style = vkhr::HairStyle(argv[2]);
            if (!style) throw std::runtime_error(std::string("cannot load ") + argv[2]);
            // SceneGraph::add_style preprocessing (src/vkhr/scene_graph.cc:232-242)
            style.seed = std::strtoull(argv[3], nullptr, 0);
            style.shuffle();
            if (!style.has_tangents()) style.generate_tangents();
            if (!style.has_thickness()) style.generate_thickness(0.042f);
            if (!style.has_indices()) style.generate_indices();
            if (!style.has_bounding_box()) style.generate_bounding_box();
            in.open(argv[4], std::ios::binary);
            uint32_t hdr[4];
            read(in, hdr, 4);
            n = hdr[0], W = hdr[1], H = hdr[2], S = hdr[3];
            out = argv[5];
            write(out + "/vertices.bin", style.vertices.data(), style.vertices.size() * 12);
            write(out + "/style_tangents.bin", style.tangents.data(), style.tangents.size() * 12);
        } else {
            in.open(argv[save ? 2 : 1], std::ios::binary);
            uint32_t hdr[6];
            read(in, hdr, 6);
            const uint32_t n_strands = hdr[0], n_vertices = hdr[1];
            n = hdr[2], W = hdr[3], H = hdr[4], S = hdr[5];
            style.segments.resize(n_strands);
            style.vertices.resize(n_vertices);
            read(in, style.segments.data(), n_strands);
            read(in, style.vertices.data(), n_vertices);
            if (save) {
                if (!style.save(argv[3])) throw std::runtime_error(std::string("cannot save ") + argv[3]);
                std::printf("host_check: saved %u strands, %u vertices\n", style.get_strand_count(),
                            style.get_vertex_count());
                return 0;
            }
            out = argv[2];
            // SceneGraph::add_style preprocessing order (src/vkhr/scene_graph.cc:235-242)
            style.generate_tangents();
            style.generate_thickness(0.042f);
            style.generate_indices();
            style.generate_bounding_box();
        }
        vkhr_camera cam;
        vkhr_light light;
        vkhr_strand_params sp;
        vkhr_render_params rp;
        read(in, &cam, 1);
        read(in, &light, 1);
        read(in, &sp, 1);
        read(in, &rp, 1);

        // vulkan::HairStyle::load (src/vkhr/rasterizer/hair_style.cc:75-77)
        vkhr::HairStyle::Volume volume = style.voxelize_segments(n, n, n);
        std::vector<uint32_t> counts(size_t(n) * n * n);
        volume.counts.download(counts.data(), counts.size() * 4);
        volume.normalize();
        volume.download();

        vkhr::AABB b = style.get_bounding_box();
        sp.volume_bounds = vkhr_aabb{{b.origin.x, b.origin.y, b.origin.z}, b.radius, {b.size.x, b.size.y, b.size.z},
                                     b.volume};
        write(out + "/aabb.bin", &sp.volume_bounds, sizeof(vkhr_aabb));
        write(out + "/counts.bin", counts.data(), counts.size() * 4);
        write(out + "/density.bin", volume.densities.data(), volume.densities.size());
        write(out + "/tangents.bin", volume.tangents.data(), volume.tangents.size() * 4);
        write(out + "/indices.bin", style.indices.data(), style.indices.size() * 4);

        // raymarcher pass
        vkhr::DeviceBuffer rgba(size_t(W) * H * 16);
        vkhr::Raymarcher().draw(volume, cam, light, sp, rp, W, H, rgba.as<float>(), nullptr);
        std::vector<float> img(size_t(W) * H * 4);
        rgba.download(img.data(), img.size() * 4);
        write(out + "/raymarch.bin", img.data(), img.size() * 4);

        // strand pass: shadow map, strands into the PPLL, resolve
        vkhr::DeviceBuffer shadow(size_t(S) * S * 4);
        vkhr::StrandRasterizer raster;
        float light_vp[16];
        in.read(reinterpret_cast<char*>(light_vp), sizeof(light_vp));
        raster.draw_depth(style, light_vp, sp.strand_radius, S, shadow.as<float>());
        vkhr::LinkedList ppll(W, H, vkhr::LinkedList::AverageFragmentsPerPixel * W * H);
        ppll.clear();
        vkhr_render_params raster_params = rp;
        raster_params.renderer = 0;
        raster.draw(style, volume, cam, light, sp, raster_params, shadow.as<float>(), S, ppll);
        std::vector<float> smap(size_t(S) * S);
        shadow.download(smap.data(), smap.size() * 4);
        write(out + "/shadow.bin", smap.data(), smap.size() * 4);
        ppll.resolve(rgba.as<float>());
        rgba.download(img.data(), img.size() * 4);
        write(out + "/resolve.bin", img.data(), img.size() * 4);
        uint32_t used = ppll.node_count_used();

        // the load-time spatial order of the pairs (a permutation)
        style.sort_indices_spatially();
        write(out + "/indices_sorted.bin", style.indices.data(), style.indices.size() * 4);

        // hybrid colour passes (Rasterizer::update + draw_color), one per
        // (magnified, minified) LoD distance pair of the optional tail
        uint32_t n_hybrid = 0;
        if (in.read(reinterpret_cast<char*>(&n_hybrid), sizeof(n_hybrid))) {
            for (uint32_t k = 0; k < n_hybrid; ++k) {
                float dist[2];
                read(in, dist, 2);
                vkhr_render_params hp = rp;
                hp.renderer = vkhr::Renderer::Hybrid_LoD;
                hp.lod_magnified_distance = dist[0];
                hp.lod_minified_distance = dist[1];
                vkhr::HybridRenderer hybrid;
                hybrid.update(cam, hp);
                hybrid.draw_color(style, volume, cam, light, sp, hp, shadow.as<float>(), S, ppll, rgba.as<float>());
                rgba.download(img.data(), img.size() * 4);
                const std::string tag = out + "/hybrid" + std::to_string(k);
                write(tag + ".bin", img.data(), img.size() * 4);
                const std::string info = std::to_string(hybrid.level_of_detail) + " " +
                                         std::to_string(int(hybrid.rasterizer_enabled(hp))) + " " +
                                         std::to_string(int(hybrid.raymarcher_enabled(hp))) + " " +
                                         std::to_string(ppll.node_count_used()) + "\n";
                write(tag + ".txt", info.data(), info.size());
            }
        }
        std::printf("host_check: %u strands, %u segments, %u nodes\n", style.get_strand_count(),
                    style.get_segment_count(), used);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "host_check: %s\n", e.what());
        return 1;
    }
}
