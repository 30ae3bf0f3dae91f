// Batched raymarch over (style, tile) work items -- one block per item, or
// persistent blocks pulling items from an atomic queue -- and the host side
// of the batched launch: device tables, one hipGraph capturing clear ->
// voxelize -> finalize -> normalize -> cells -> raymarch, replayed by
// vkhr_batch_launch.
#include "batch.hpp"
#include "volume.hpp"
#include "voxel.hpp"

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
        *use_occ = stage_occupancy(a.vol, occ_lds);
        *staged = s;
    }
    const uint32_t local = it - d.tile_prefix[s];
    const int bx = int(local % d.tiles_x), by = int(local / d.tiles_x);
    const int px = bx * 2 * tw + (wave & 1) * tw + lane % tw;
    const int py = by * 2 * th + (wave >> 1) * th + lane / tw;
    raymarch_pixel<false>(a, px, py, OccupancyView{occ_lds, *use_occ});
}

// One block per work item: the hardware dispatcher balances the tiles of all
// styles.
__global__ void __launch_bounds__(256)
batch_raymarch_grid_kernel(BatchDevice d, const RaymarchArgs* __restrict__ args) {
    __shared__ uint32_t occ_lds[kOccupancyLdsWords];
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
    __shared__ uint32_t occ_lds[kOccupancyLdsWords];
    uint32_t staged = 0xffffffffu;   // the style whose occupancy is in LDS
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
        batch_raymarch_grid_kernel<<<d.total_tiles, 256, 0, stream>>>(d, d.args);
    else
        batch_raymarch_kernel<<<grid, 256, 0, stream>>>(d, d.args);
}

}  // namespace vkhr

using namespace vkhr;

struct vkhr_batch {
    BatchDevice d;
    std::vector<void*> allocations;
    unsigned march_grid = 0;
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
    for (void* p : b->allocations) (void)hipFree(p);
    b->allocations.clear();
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
    vkhr_batch* b = new vkhr_batch();
    std::memset(&b->d, 0, sizeof(b->d));
    std::vector<BatchVolStyle> vs(n_styles);
    std::vector<RaymarchArgs> args(n_styles);
    std::vector<uint64_t> pair_prefix(n_styles + 1, 0);
    std::vector<uint32_t> tile_prefix(n_styles + 1, 0);
    std::vector<BatchChunk> vchunks, cchunks;
    const uint64_t kChunk = 1u << 16;   // voxels per finalize chunk
    const uint32_t bw = 2 * t.tile_w, bh = 2 * t.tile_h;
    const uint32_t tiles_x = (width + bw - 1) / bw, tiles_y = (height + bh - 1) / bh;
    for (uint32_t s = 0; s < n_styles; ++s) {
        const vkhr_batch_style& in = styles[s];
        VoxelGrid g;
        if ((rc = make_grid(&in.bounds, in.volume, in.volume, in.volume, &g))) {
            delete b;
            return rc;
        }
        if (!((in.counts || in.tangent_acc) && in.density && in.cells && in.rgba && (in.n_indices < 2 || (in.xyz && in.indices)))) {
            delete b;
            set_error("style %u: missing buffer", s);
            return VKHR_ERR_INVALID;
        }
        const auto misaligned = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) % a) != 0; };
        if (misaligned(in.tangent_acc, 16) || misaligned(in.counts, 16) || misaligned(in.density, 4) ||
            misaligned(in.tangent_i8x4, 16)) {
            delete b;
            set_error("style %u: tangent_acc, counts, tangent_i8x4 must be 16-byte and density 4-byte aligned", s);
            return VKHR_ERR_INVALID;
        }
        BatchVolStyle& v = vs[s];
        v.ox = g.ox; v.oy = g.oy; v.oz = g.oz; v.vx = g.vx; v.vy = g.vy; v.vz = g.vz;
        v.mx = g.mx; v.my = g.my; v.mz = g.mz; v.nx = g.nx; v.ny = g.ny; v.nz = g.nz;
        v.xyz = in.xyz;
        v.indices = in.indices;
        v.tangents = in.tangents;
        v.n_vertices = in.n_vertices;
        v.n_pairs = in.n_indices / 2;
        v.counts = in.counts;
        v.acc = reinterpret_cast<long long*>(in.tangent_acc);
        v.density = in.density;
        v.tangent_q = in.tangent_i8x4;
        v.cells = reinterpret_cast<unsigned long long*>(in.cells);
        v.occupancy = in.occupancy;
        if (v.acc && !(v.tangents && v.tangent_q)) {
            delete b;
            set_error("style %u: tangent_acc needs tangents and tangent_i8x4", s);
            return VKHR_ERR_INVALID;
        }
        pair_prefix[s + 1] = pair_prefix[s] + v.n_pairs;
        uint64_t nvox = uint64_t(in.volume) * in.volume * in.volume;
        for (uint64_t a = 0; a < nvox; a += kChunk) vchunks.push_back({s, 0, a, std::min(nvox, a + kChunk)});
        const uint64_t cb = cell_blocks(in.volume, in.volume, in.volume);
        for (uint64_t k = 0; k < cb; ++k) cchunks.push_back({s, 0, k, k + 1});
        vkhr_volume vol;
        std::memset(&vol, 0, sizeof(vol));
        vol.density = in.density;
        vol.cells = in.cells;
        vol.occupancy = in.occupancy;
        vol.tangents = in.tangent_i8x4;
        vol.nx = vol.ny = vol.nz = in.volume;
        if ((rc = make_raymarch_args(&vol, &in.camera, &in.light, &in.strand, &in.params, width, height, &args[s]))) {
            delete b;
            return rc;
        }
        args[s].rgba = reinterpret_cast<float4*>(in.rgba);
        tile_prefix[s + 1] = tile_prefix[s] + tiles_x * tiles_y;
    }
    BatchDevice& d = b->d;
    std::vector<uint32_t> zeros(2 * n_styles + 1, 0);
    uint32_t* mm = nullptr;
    if ((rc = upload(b, vs, &d.styles)) || (rc = upload(b, args, &d.args)) ||
        (rc = upload(b, pair_prefix, &d.pair_prefix)) || (rc = upload(b, tile_prefix, &d.tile_prefix)) ||
        (rc = upload(b, vchunks, &d.voxel_chunks)) || (rc = upload(b, cchunks, &d.cell_chunks)) ||
        (rc = upload(b, zeros, &mm))) {
        release(b);
        delete b;
        return rc;
    }
    d.minmax = mm;
    d.work_counter = mm + 2 * n_styles;
    d.n_styles = n_styles;
    d.n_voxel_chunks = uint32_t(vchunks.size());
    d.n_cell_chunks = uint32_t(cchunks.size());
    d.total_tiles = tile_prefix[n_styles];
    d.total_pairs = pair_prefix[n_styles];
    d.tile_w = t.tile_w;
    d.tile_h = t.tile_h;
    d.tiles_x = tiles_x;
    int dev = 0, cus = 0;
    VKHR_CHECK_HIP(hipGetDevice(&dev));
    VKHR_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
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
        launch_batch_volume(b->d, s);
        launch_batch_march(b->d, b->march_grid, s);
        hipError_t e = hipStreamEndCapture(s, &b->graph);
        VKHR_CHECK_HIP(e);
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
