// Command-line driver of the host HairStyle's file side (no GPU needed):
//   hair_tool info    <in.hair>                      -> one JSON line
//   hair_tool copy    <in.hair> <out.hair>           (load + save)
//   hair_tool shuffle <in.hair> <out.hair> <seed>    (load, shuffle, save)
//   hair_tool reduce  <in.hair> <out.hair> <seed> <ratio>
//   hair_tool bbox    <in.hair> <out.hair>           (load, generate_bounding_box, save)
// Every command prints {"ok": ..., "error": <HairStyle::Error>, ...} and
// exits 0 when the operation succeeded.
#include "vkhr/hair_style.hh"

#include <cstdio>
#include <cstdlib>
#include <string>

using vkhr::HairStyle;

namespace {

int report(const HairStyle& h, bool ok) {
    const auto& f = h.header();
    const vkhr::AABB b = h.get_bounding_box();
    std::printf("{\"ok\": %s, \"error\": %d, \"strands\": %u, \"vertices\": %u, \"segments\": %u, \"field\": %u, "
                "\"size\": %zu, \"default_segment_count\": %u, \"has_bounding_box\": %s, "
                "\"bbox_min\": [%.9g, %.9g, %.9g], \"bbox_max\": [%.9g, %.9g, %.9g], "
                "\"aabb\": {\"origin\": [%.9g, %.9g, %.9g], \"radius\": %.9g, \"size\": [%.9g, %.9g, %.9g], "
                "\"volume\": %.9g}}\n",
                ok ? "true" : "false", int(h.get_last_error_state()), h.get_strand_count(), h.get_vertex_count(),
                h.get_vertex_count() >= h.get_strand_count() ? h.get_segment_count() : 0u, f.field, h.get_size(),
                h.get_default_segment_count(), h.has_bounding_box() ? "true" : "false", f.bounding_box_min[0],
                f.bounding_box_min[1], f.bounding_box_min[2], f.bounding_box_max[0], f.bounding_box_max[1],
                f.bounding_box_max[2], b.origin.x, b.origin.y, b.origin.z, b.radius, b.size.x, b.size.y, b.size.z,
                b.volume);
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: hair_tool info|copy|shuffle|reduce|bbox <in.hair> [out.hair] [seed] [ratio]\n");
        return 2;
    }
    const std::string cmd = argv[1];
    HairStyle h;
    bool ok = h.load(argv[2]);
    if (cmd == "info" || !ok) return report(h, ok);
    if (argc < 4) return 2;
    if (cmd == "shuffle" || cmd == "reduce") {
        if (argc < 5) return 2;
        h.seed = std::strtoull(argv[4], nullptr, 0);
        if (cmd == "shuffle")
            h.shuffle();
        else if (argc >= 6)
            h.reduce(std::strtof(argv[5], nullptr));
        else
            return 2;
    } else if (cmd == "bbox") {
        h.generate_bounding_box();
    } else if (cmd != "copy") {
        return 2;
    }
    ok = h.save(argv[3]);
    return report(h, ok);
}
