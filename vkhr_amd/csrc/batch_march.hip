// Batched raymarch over (style, tile) work items -- one block per item, or
// persistent blocks pulling items from an atomic queue -- and the host side
// of the batched launch: device tables per group of styles, one hipGraph
// capturing the volume build passes and the raymarch of every group
// (pipelined: a group's raymarch beside the next group's build), replayed
// by vkhr_batch_launch.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace vkhr {
namespace {

// Work item it = one block tile (2x2 wave tiles of tile_w x tile_h pixels)
// of one style; all threads of the block call it with the same it.
__device__ __forceinline__ void batch_item(const BatchDevice& d, const RaymarchArgs* __restrict__ args,
                                           uint32_t it, uint32_t* occ_lds, uint32_t* staged, bool* use_occ) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tw = int(d.tile_w), th = int(d.tile_h);
    const uint32_t s = uint32_t(__builtin_amdgcn_readfirstlane(int(style_of(d.tile_prefix, d.n_styles, it))));
    // a read-only, non-aliased argument: the uniform loads of the style's
    // arguments become scalar loads instead of filling vector registers
    const RaymarchArgs& a = args[s];
    if (s != *staged) {  // block-uniform
        if (*staged != 0xffffffffu) __syncthreads();   // every wave is done with the previous style's field
        *use_occ = stage_distance(a.vol, occ_lds);
        *staged = s;
    }
    const uint32_t local = it - d.tile_prefix[s];
    const int bx = int(local % d.tiles_x), by = int(local / d.tiles_x);
    const int px = bx * 2 * tw + (wave & 1) * tw + lane % tw;
    const int py = by * 2 * th + (wave >> 1) * th + lane / tw;
    raymarch_pixel<false>(a, px, py, OccupancyView{occ_lds, *use_occ});
}

// One block per work item: the hardware dispatcher balances the tiles of all
// styles.  Five waves per SIMD, as the per-style raymarch kernel.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
batch_raymarch_grid_kernel(BatchDevice d, const RaymarchArgs* __restrict__ args) {
    extern __shared__ uint32_t occ_lds[];   // d.march_lds_bytes
    uint32_t staged = 0xffffffffu;
    bool use_occ = false;
    batch_item(d, args, uint32_t(__builtin_amdgcn_readfirstlane(int(blockIdx.x))), occ_lds, &staged, &use_occ);
}

// Persistent blocks pulling work items from an atomic queue.  With the pixel
// body inlined into the loop the compiler allocates 141 registers (3 waves
// per SIMD against the grid form's 5); the autotuner measures both.
__global__ void __launch_bounds__(256)
batch_raymarch_kernel(BatchDevice d, const RaymarchArgs* __restrict__ args) {
    __shared__ uint32_t item;
    extern __shared__ uint32_t occ_lds[];   // d.march_lds_bytes
    uint32_t staged = 0xffffffffu;   // the style whose distance field is in LDS
    bool use_occ = false;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(d.work_counter, 1u);
        __syncthreads();
        const uint32_t it = uint32_t(__builtin_amdgcn_readfirstlane(int(item)));
        __syncthreads();
        if (it >= d.total_tiles) break;
        batch_item(d, args, it, occ_lds, &staged, &use_occ);
    }
}

}  // namespace

// grid == 0: one block per work item; else that many persistent blocks.
void launch_batch_march(const BatchDevice& d, unsigned grid, hipStream_t stream) {
    if (grid == 0)
        batch_raymarch_grid_kernel<<<d.total_tiles, 256, d.march_lds_bytes, stream>>>(d, d.args);
    else
        batch_raymarch_kernel<<<grid, 256, d.march_lds_bytes, stream>>>(d, d.args);
}

}  // namespace vkhr

using namespace vkhr;

struct vkhr_batch {
    std::vector<BatchDevice> groups;    // one device table per group of styles
    std::vector<void*> allocations;
    unsigned march_grid = 0;
    hipStream_t side = nullptr;         // the raymarch branch of the captured pipeline
    std::vector<hipEvent_t> events;     // fork, one per group, join
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <class T>
int upload(vkhr_batch* b, const std::vector<T>& host, T** out) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(host.size() * sizeof(T), 16);
    VKHR_CHECK_HIP(hipMalloc(&p, bytes));
    b->allocations.push_back(p);
    if (!host.empty()) VKHR_CHECK_HIP(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<T*>(p);
    return VKHR_OK;
}

void release(vkhr_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    b->events.clear();
    if (b->side) (void)hipStreamDestroy(b->side);
    b->side = nullptr;
    for (void* p : b->allocations) (void)hipFree(p);
    b->allocations.clear();
}

// The device tables of styles[first, last): per style its build arguments
// (voxel grid, inputs, outputs, and its section of the group's build
// state: bins, status, min/max slots, uint8 view, bin entries), raymarch
// arguments, the workgroup offsets of every pass, tile prefix and distance
// line chunks.
int build_group(vkhr_batch* b, const vkhr_batch_style* styles, uint32_t first, uint32_t last, uint32_t width,
                uint32_t height, const vkhr_tuning& t, BatchDevice* out) {
    const uint32_t n = last - first;
    std::vector<BatchVolStyle> vs(n);
    std::vector<RaymarchArgs> args(n);
    std::vector<uint32_t> tile_prefix(n + 1, 0);
    std::vector<uint32_t> prefix[kPassCount];
    for (auto& p : prefix) p.assign(n + 1, 0);
    std::vector<BatchChunk> lchunks[3];
    std::vector<BuildLayout> layouts(n);
    std::vector<size_t> layout_base(n + 1, 0);
    size_t lds_bytes = 0;
    const uint32_t bw = 2 * t.tile_w, bh = 2 * t.tile_h;
    const uint32_t tiles_x = (width + bw - 1) / bw, tiles_y = (height + bh - 1) / bh;
    int rc = 0;
    for (uint32_t s = 0; s < n; ++s) {
        const vkhr_batch_style& in = styles[first + s];
        VoxelGrid g;
        if ((rc = make_grid(&in.bounds, in.volume, in.volume, in.volume, &g))) return rc;
        if ((rc = check_volume(in.volume, in.volume, in.volume, true, "batch style"))) return rc;
        if (!((in.counts || in.tangent_acc) && in.density && in.cells && in.rgba &&
              (in.n_indices < 2 || (in.xyz && in.indices && in.n_vertices)))) {
            set_error("style %u: missing buffer", first + s);
            return VKHR_ERR_INVALID;
        }
        const auto misaligned = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) != 0; };
        if (misaligned(in.tangent_acc, 16) || misaligned(in.counts, 16) || misaligned(in.density, 4) ||
            misaligned(in.tangent_i8x4, 16)) {
            set_error("style %u: tangent_acc, counts, tangent_i8x4 must be 16-byte and density 4-byte aligned",
                      first + s);
            return VKHR_ERR_INVALID;
        }
        if (in.tangent_acc && !(in.tangents && in.tangent_i8x4)) {
            set_error("style %u: tangent_acc needs tangents and tangent_i8x4", first + s);
            return VKHR_ERR_INVALID;
        }
        if (in.tangent_i8x4 && !in.tangent_acc) {
            set_error("style %u: tangent_i8x4 needs tangent_acc", first + s);
            return VKHR_ERR_INVALID;
        }
        const uint32_t n_pairs = in.n_indices / 2;  // the reference loops while i < size-1
        const BrickGrid bg = brick_grid(g.nx, g.ny, g.nz);
        const size_t n_bricks = size_t(bg.bx) * bg.by * bg.bz;
        const size_t nvox = size_t(g.nx) * g.ny * g.nz;
        const size_t capacity = build_capacity(n_pairs);
        layouts[s] = build_layout(n_bricks, nvox, true, capacity);
        layout_base[s + 1] = layout_base[s] + layouts[s].words;
        BuildArgs& a = vs[s].a;
        std::memset(&a, 0, sizeof(a));
        a.g = g;
        a.b = bg;
        a.xyz = in.xyz;
        a.indices = in.indices;
        a.tangents = in.tangents;
        a.n_vertices = in.n_vertices;
        a.n_pairs = n_pairs;
        a.n_bricks = uint32_t(n_bricks);
        a.pairs_per_block = build_pairs_per_block(n_pairs);
        a.capacity = uint32_t(capacity);
        a.occupancy = in.occupancy;
        a.occupancy_words = in.occupancy ? occupancy_words(g.nx, g.ny, g.nz) : 0;
        a.counts = in.counts;
        a.tangent_i8x4 = in.tangent_i8x4;
        a.acc = reinterpret_cast<long long*>(in.tangent_acc);
        a.acc_counts = in.tangent_acc ? nullptr : in.counts;
        vs[s].cells = reinterpret_cast<unsigned long long*>(in.cells);
        vs[s].density = in.density;
        const uint32_t blocks[kPassCount] = {
            grid_for(std::max(n_bricks, a.occupancy_words), 256, 1024),                           // init
            n_pairs ? (n_pairs + a.pairs_per_block - 1) / a.pairs_per_block : 0u,                 // bins
            brick_blocks(bg),                                                                     // bricks
            grid_for(nvox, 256, 128),                                                             // finalize
            cells_pass_blocks(g.nx, g.ny, g.nz)};                                                 // cells
        for (int p = 0; p < kPassCount; ++p) {
            const uint64_t next = uint64_t(prefix[p][s]) + blocks[p];
            if (next >= (uint64_t(1) << 31)) {
                set_error("batch group: more than 2^31 workgroups in one pass");
                return VKHR_ERR_OVERFLOW;
            }
            prefix[p][s + 1] = uint32_t(next);
        }
        if (in.occupancy && distance_field_fits(g.nx, g.ny, g.nz))
            for (int axis = 0; axis < 3; ++axis) {
                const uint32_t lines = distance_lines(g.nx, g.ny, g.nz, axis);
                for (uint32_t l = 0; l < lines; l += 256 / kWave)
                    lchunks[axis].push_back({s, 0, l, std::min<uint64_t>(lines, l + 256 / kWave)});
            }
        vkhr_volume vol;
        std::memset(&vol, 0, sizeof(vol));
        vol.density = in.density;
        vol.cells = in.cells;
        vol.occupancy = in.occupancy;
        vol.tangents = in.tangent_i8x4;
        vol.nx = vol.ny = vol.nz = in.volume;
        if ((rc = make_raymarch_args(&vol, &in.camera, &in.light, &in.strand, &in.params, width, height, &args[s])))
            return rc;
        args[s].rgba = reinterpret_cast<float4*>(in.rgba);
        lds_bytes = std::max(lds_bytes, distance_lds_bytes(args[s].vol));
        tile_prefix[s + 1] = tile_prefix[s] + tiles_x * tiles_y;
    }
    // the build state of every style of the group, one allocation; the
    // work counter of the persistent raymarch after it
    void* state = nullptr;
    VKHR_CHECK_HIP(hipMalloc(&state, (layout_base[n] + 64) * sizeof(uint32_t)));
    b->allocations.push_back(state);
    uint32_t* words = static_cast<uint32_t*>(state);
    for (uint32_t s = 0; s < n; ++s) {
        uint32_t* w = words + layout_base[s];
        const BuildLayout& l = layouts[s];
        BuildArgs& a = vs[s].a;
        a.bins = w + l.bins;
        a.cursor = w + l.cursor;
        a.status = w + l.status;
        a.minmax = w + l.minmax;
        a.brick_mm = w + l.brick_mm;
        a.density = reinterpret_cast<uint8_t*>(w + l.raw);
        a.entries = reinterpret_cast<BinEntry*>(w + l.entries);
    }
    BatchDevice& d = *out;
    std::memset(&d, 0, sizeof(d));
    BatchVolStyle* dev_styles = nullptr;
    if ((rc = upload(b, vs, &dev_styles)) || (rc = upload(b, args, &d.args)) ||
        (rc = upload(b, tile_prefix, &d.tile_prefix)) || (rc = upload(b, lchunks[0], &d.line_chunks[0])) ||
        (rc = upload(b, lchunks[1], &d.line_chunks[1])) || (rc = upload(b, lchunks[2], &d.line_chunks[2])))
        return rc;
    for (int p = 0; p < kPassCount; ++p) {
        if ((rc = upload(b, prefix[p], &d.block_prefix[p]))) return rc;
        d.blocks[p] = prefix[p][n];
    }
    d.styles = dev_styles;
    d.work_counter = words + layout_base[n];
    d.n_styles = n;
    for (int axis = 0; axis < 3; ++axis) d.n_line_chunks[axis] = uint32_t(lchunks[axis].size());
    d.march_lds_bytes = uint32_t(lds_bytes);
    d.total_tiles = tile_prefix[n];
    d.tile_w = t.tile_w;
    d.tile_h = t.tile_h;
    d.tiles_x = tiles_x;
    return VKHR_OK;
}

}  // namespace

extern "C" int vkhr_batch_create(const vkhr_batch_style* styles, uint32_t n_styles, uint32_t width, uint32_t height,
                                 const vkhr_tuning* tuning, vkhr_batch** out) {
    clear_error();
    int rc = ensure_device();
    if (rc) return rc;
    VKHR_REQUIRE(styles && n_styles && out, VKHR_ERR_INVALID, "no styles");
    VKHR_REQUIRE(width && height, VKHR_ERR_INVALID, "empty image");
    vkhr_tuning t = {8, 8, 0, 0};
    if (tuning) t = *tuning;
    VKHR_REQUIRE(t.tile_w * t.tile_h == 64 && t.tile_w >= 1 && t.tile_h >= 1, VKHR_ERR_INVALID,
                 "wave tile %ux%u is not 64 pixels", t.tile_w, t.tile_h);
    VKHR_REQUIRE(t.blocks_per_cu <= 32, VKHR_ERR_INVALID, "blocks_per_cu %u", t.blocks_per_cu);
    VKHR_REQUIRE(t.groups <= 64, VKHR_ERR_INVALID, "groups %u", t.groups);
    const uint32_t n_groups = std::max(1u, std::min(t.groups, n_styles));
    // Groups of consecutive styles of about equal work (voxels plus segments).
    std::vector<double> cost(n_styles + 1, 0.0);
    for (uint32_t s = 0; s < n_styles; ++s)
        cost[s + 1] = cost[s] + double(styles[s].volume) * styles[s].volume * styles[s].volume +
                      8.0 * double(styles[s].n_indices / 2);
    std::vector<uint32_t> bounds(n_groups + 1, 0);
    bounds[n_groups] = n_styles;
    for (uint32_t g = 1; g < n_groups; ++g) {
        const double target = cost[n_styles] * g / n_groups;
        uint32_t s = bounds[g - 1] + 1;
        while (s < n_styles - (n_groups - g) && cost[s] < target) ++s;
        bounds[g] = s;
    }
    int dev = 0, cus = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    VKHR_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    vkhr_batch* b = new vkhr_batch();
    b->groups.resize(n_groups);
    for (uint32_t g = 0; g < n_groups; ++g) {
        if ((rc = build_group(b, styles, bounds[g], bounds[g + 1], width, height, t, &b->groups[g]))) {
            release(b);
            delete b;
            return rc;
        }
    }
    if (n_groups > 1) {
        hipError_t e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
        for (uint32_t k = 0; e == hipSuccess && k < n_groups + 2; ++k) {
            hipEvent_t ev = nullptr;
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e == hipSuccess) b->events.push_back(ev);
        }
        if (e != hipSuccess) {
            release(b);
            delete b;
            set_error("batch stream/events: %s", hipGetErrorString(e));
            return VKHR_ERR_HIP;
        }
    }
    b->march_grid = unsigned(cus) * t.blocks_per_cu;
    *out = b;
    return VKHR_OK;
}

extern "C" int vkhr_batch_launch(vkhr_batch* b, vkhr_stream stream) {
    clear_error();
    VKHR_REQUIRE(b, VKHR_ERR_INVALID, "null batch");
    VKHR_REQUIRE(stream, VKHR_ERR_INVALID, "the batched launch is graph-captured: pass a created stream, not NULL");
    hipStream_t s = as_stream(stream);
    if (!b->exec) {
        VKHR_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t err = hipSuccess;  // the first failure; the capture is ended either way
        const auto check = [&](hipError_t x) {
            if (err == hipSuccess) err = x;
        };
        const size_t n_groups = b->groups.size();
        if (n_groups == 1) {
            launch_batch_volume(b->groups[0], s);
            launch_batch_march(b->groups[0], b->march_grid, s);
        } else {
            // A software pipeline: the volume passes of group g+1 (atomic and
            // HBM bound) run beside the raymarch of group g (VALU bound).  The
            // volume passes stay on the capturing stream, the raymarches go to
            // the side stream after their group's volume; the side stream
            // forks from and joins back into the capture.
            check(hipEventRecord(b->events[0], s));
            check(hipStreamWaitEvent(b->side, b->events[0], 0));
            for (size_t g = 0; g < n_groups; ++g) {
                launch_batch_volume(b->groups[g], s);
                check(hipEventRecord(b->events[1 + g], s));
                check(hipStreamWaitEvent(b->side, b->events[1 + g], 0));
                launch_batch_march(b->groups[g], b->march_grid, b->side);
            }
            check(hipEventRecord(b->events[1 + n_groups], b->side));
            check(hipStreamWaitEvent(s, b->events[1 + n_groups], 0));
        }
        check(hipGetLastError());
        check(hipStreamEndCapture(s, &b->graph));
        if (err != hipSuccess && b->graph) {  // the next launch captures afresh
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        VKHR_REQUIRE(err == hipSuccess, VKHR_ERR_HIP, "batch graph capture failed: %s", hipGetErrorString(err));
        VKHR_CHECK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    VKHR_CHECK_HIP(hipGraphLaunch(b->exec, s));
    return VKHR_OK;
}

extern "C" int vkhr_batch_destroy(vkhr_batch* b) {
    clear_error();
    if (!b) return VKHR_OK;
    release(b);
    delete b;
    return VKHR_OK;
}
