/*
 * vkhr_hip.h -- C ABI of the MI355X (gfx950) hybrid-hair hot path.
 *
 * Every entry point is a plain C function over caller-owned device pointers,
 * enqueued on a caller HIP stream (passed as an opaque pointer; NULL is the
 * default stream), and returns an int status: VKHR_OK (0) or a negative
 * VKHR_ERR_* code, with a message available from vkhr_last_error().  No C++
 * exception crosses this boundary.  Calls on distinct buffers are reentrant;
 * the only hidden state is a scratch area per device and stream that the
 * library allocates on first use and keeps until vkhr_release_scratch on
 * that stream (or the process ends): reductions, counters, the adjoint's
 * cell accumulator, the bins of the volume build, used by
 * vkhr_aabb_compute, vkhr_volume_finalize, vkhr_volume_build_segments and
 * vkhr_raymarch_adjoint.  Calls on one stream are ordered by it; calls on
 * different streams use different areas, so they may run concurrently.
 * hipStreamPerThread gets one area per calling thread.  An area that must
 * grow synchronizes its stream and frees the old buffer (during graph
 * capture the old buffer is kept until the release instead).  The one
 * allocation-free entry point (vkhr_volume_finalize_noalloc) uses a static
 * 4 KiB array per device instead, shared by all its calls on that device.
 *
 * Layouts:
 *   xyz           float32 x,y,z per vertex, tightly packed (12 B / vertex)
 *   indices       uint32 line-list pairs (root, tip) as HairStyle::indices
 *   volumes       x fastest, then y, then z: linear index x + y*nx + z*nx*ny
 *   matrices      column-major float[16] (glm memory order)
 *   images        row-major, row 0 at the top (Vulkan framebuffer order),
 *                 float RGBA interleaved
 *
 * Each function names the reference function or shader it replaces
 * (paths relative to the vkhr source tree).
 */
#ifndef VKHR_HIP_H
#define VKHR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VKHR_ABI_VERSION 3

enum vkhr_status {
    VKHR_OK = 0,
    VKHR_ERR_INVALID = -1,     /* bad argument (null pointer, zero size, ...) */
    VKHR_ERR_HIP = -2,         /* a HIP runtime call failed */
    VKHR_ERR_OVERFLOW = -3,    /* a size exceeds what the path supports */
    VKHR_ERR_NOMEM = -4,       /* scratch allocation failed */
    VKHR_ERR_NO_DEVICE = -5    /* no usable gfx950 device */
};

typedef void* vkhr_stream; /* hipStream_t */

/* AABB of a hair style: include/vkhr/scene_graph/hair_style.hh:16-21 */
typedef struct vkhr_aabb {
    float origin[3];
    float radius;
    float size[3];
    float volume;
} vkhr_aabb;

/* Camera uniform block: ViewProjection, include/vkhr/scene_graph/camera.hh:15-22
 * (GLSL twin share/shaders/scene_graph/camera.glsl:4-11). */
typedef struct vkhr_camera {
    float view[16];
    float projection[16];
    float position[3];
    float look_at_distance;
    float znear, zfar;
    float resolution[2];
} vkhr_camera;

/* One light: LightSource::Buffer, include/vkhr/scene_graph/light_source.hh:17-23
 * (GLSL twin share/shaders/scene_graph/lights.glsl:4-13).  matrix is the
 * biased light view-projection used to sample the shadow map. */
typedef struct vkhr_light {
    float vector[4];
    float intensity[4];
    float matrix[16];
    float origin[3];
    float znear;
    float zfar;
    float pad_[3];
} vkhr_light;

/* Per-style parameters: vulkan::HairStyle::Parameters,
 * include/vkhr/rasterizer/hair_style.hh:48-56 (GLSL twin
 * share/shaders/strands/strand.glsl:8-16, where strand_radius is
 * strand_width, hair_opacity is hair_alpha, hair_shininess is hair_exponent). */
typedef struct vkhr_strand_params {
    vkhr_aabb volume_bounds;
    float volume_resolution[3];
    float strand_radius;
    float hair_color[3];
    float hair_opacity;
    float hair_shininess;
    float strand_ratio;
} vkhr_strand_params;

/* Renderer parameters: Interface::Parameters,
 * include/vkhr/rasterizer/interface.hh:60-112 (GLSL twin
 * share/shaders/scene_graph/params.glsl:7-33). */
typedef struct vkhr_render_params {
    int32_t shading_model;          /* 0 Kajiya-Kay, 2 ADSM only, 3 LAO only */
    int32_t adsm_kernel_size;
    int32_t adsm_sampling_type;
    int32_t adsm_stride_size;
    int32_t adsm_on;
    int32_t ctsm_kernel_size;
    int32_t ctsm_sampling_type;
    float ctsm_bias;
    int32_t ctsm_on;
    int32_t shadow_technique;
    float isosurface;
    float raycast_steps;
    float occlusion_radius;
    float ao_exponent;
    float ao_clamp;
    float lod_magnified_distance;
    int32_t renderer;               /* 0 rasterizer, 2 raymarcher, 3 hybrid */
    float lod_minified_distance;
    int32_t benchmarking;
} vkhr_render_params;

/* ---------------------------------------------------------------- library */

int vkhr_abi_version(void);
/* Message for the last error raised on the calling thread ("" if none). */
const char* vkhr_last_error(void);
/* Binds the current device, checks it is gfx950 and sets up the per-device
 * scratch.  Optional: every entry point calls it on first use. */
int vkhr_init(void);
/* Synchronizes the stream, then frees the scratch area the library holds
 * for it on the current device (and on this thread, for
 * hipStreamPerThread).  A later call on the stream allocates a new area.
 * Call it before destroying a stream the library has used, so that a host
 * that creates streams per request does not accumulate areas.  No
 * counterpart in the reference (which keeps no device scratch). */
int vkhr_release_scratch(vkhr_stream stream);
/* Bytes of the scratch area held for this stream on the current device (0
 * when none). */
size_t vkhr_scratch_bytes(vkhr_stream stream);

/* ------------------------------------------------------ style preparation */

/* HairStyle::generate_bounding_box + get_bounding_box
 * (src/vkhr/scene_graph/hair_style.cc:215-255): component min/max over the
 * vertices with both initialised at the origin, origin = min, size = max-min,
 * radius = |size|, volume = size.x*size.y*size.z.  Writes the host struct
 * *out; synchronises the stream. */
int vkhr_aabb_compute(const float* xyz, uint32_t n_vertices, vkhr_aabb* out,
                      vkhr_stream stream);

/* HairStyle::generate_indices (hair_style.cc:196-213): pairs (v, v+1) within
 * each strand.  strand_offsets holds n_strands+1 first-vertex offsets;
 * indices receives 2*(strand_offsets[n]-n_strands) entries. */
int vkhr_generate_indices(const uint32_t* strand_offsets, uint32_t n_strands,
                          uint32_t* indices, vkhr_stream stream);

/* Spatial order of the segments (no counterpart in the reference, whose
 * .hair assets store strands in scalp order; a load-time layout step after
 * generate_indices).  sorted receives the n_indices/2 pairs of indices
 * ordered by the 30-bit Morton code of each segment's midpoint in bounds
 * (10 bits per axis: q = trunc(clamp((m - origin) / size * 1024, 0, 1023)),
 * float arithmetic without contraction), equal codes in input order, pairs
 * with an index >= n_vertices last.  A permutation of the pairs: voxel
 * counts, fragments and shadow depths are the same on it, while the strand
 * passes gather from a compact region per workgroup (the rasterizer's
 * 1080p pass at 136k strands of random root order: 2.24 -> 1.73 ms).  Both
 * buffers 8-byte aligned and disjoint.  Allocates and frees temporaries;
 * synchronises the stream. */
int vkhr_sort_segments(const float* xyz, uint32_t n_vertices, const uint32_t* indices, uint32_t n_indices,
                       const vkhr_aabb* bounds, uint32_t* sorted, vkhr_stream stream);

/* The same kind of load-time layout step, strand by strand: a strand is a
 * run of consecutive pairs each starting where the previous one ends (as
 * vkhr_generate_indices emits them); the runs are ordered by the 30-bit
 * Morton code of their root vertex (the run's first index; coded as
 * vkhr_sort_segments codes a midpoint), equal codes in input order, and a
 * run keeps its pairs in order.  A permutation of the pairs: the volume
 * build's passes then read each strand's vertices contiguously while their
 * workgroups still see few distinct bricks (the one-call build at 136k
 * strands / 256^3: 0.227 ms on the midpoint order, 0.203 ms on this one).
 * Same buffers and conditions as vkhr_sort_segments. */
int vkhr_sort_strands(const float* xyz, uint32_t n_vertices, const uint32_t* indices, uint32_t n_indices,
                      const vkhr_aabb* bounds, uint32_t* sorted, vkhr_stream stream);

/* HairStyle::generate_tangents (hair_style.cc:171-194): t_v =
 * normalize(x_{v+1}-x_v), the last vertex of a strand repeats the previous
 * tangent.  Strands of a single vertex get a zero tangent. */
int vkhr_generate_tangents(const float* xyz, const uint32_t* strand_offsets,
                           uint32_t n_strands, float* tangents,
                           vkhr_stream stream);

/* ------------------------------------------------------------ voxelization */

/* HairStyle::voxelize_vertices (hair_style.cc:257-294): one count per vertex
 * in voxel min(floor((v-o)/(size/res)), res-1), linear index formed in float.
 * counts (nx*ny*nz uint32) is zeroed by the call and receives unsaturated
 * counts; the reference's uint8 view is min(count, 255).  tangent_acc
 * (nullable, 2 int64 per voxel) receives the per-voxel sum of the tangents
 * of all contributions in fixed point, 2^-16 per unit, packed with the
 * count as word0 = x*2^32 + y and word1 = z*2^32 + count (sign-extended
 * fields).  With tangent_acc, counts may be NULL: the counts then stay
 * packed (vkhr_volume_finalize unpacks them); otherwise they are unpacked
 * into counts at the end of the call (one more pass over tangent_acc);
 * integer sums, so the result does not depend on thread order (the
 * reference sums floats serially).  tangents must then be non-null.  Requires nx*ny*nz <= 2^24
 * (the float index is exact only up to there). */
int vkhr_voxelize_vertices(const float* xyz, uint32_t n_vertices,
                           const float* tangents, const vkhr_aabb* bounds,
                           uint32_t nx, uint32_t ny, uint32_t nz,
                           uint32_t* counts, int64_t* tangent_acc,
                           vkhr_stream stream);

/* HairStyle::voxelize_segments (hair_style.cc:296-342): per index pair, the
 * reference's DDA -- root and tip in voxel units, steps = max |tip-root|,
 * direction = (tip-root)/steps, ceil(steps) samples at root, root+dir, ...
 * accumulated sequentially in float; each sample counts voxel
 * min(floor(p), res-1).  Bit-exact with the reference's float arithmetic
 * evaluated in IEEE single precision without contraction.  Segments whose
 * samples leave the volume (possible only with an AABB that does not contain
 * the vertices) are clamped at 0 from below instead of indexing out of
 * bounds.  Same outputs and limits as vkhr_voxelize_vertices; the tangent of
 * a segment is the tangent of its root vertex. */
int vkhr_voxelize_segments(const float* xyz, uint32_t n_vertices,
                           const uint32_t* indices, uint32_t n_indices,
                           const float* tangents, const vkhr_aabb* bounds,
                           uint32_t nx, uint32_t ny, uint32_t nz,
                           uint32_t* counts, int64_t* tangent_acc,
                           vkhr_stream stream);

/* Turns raw counts into the volumes the renderers sample, as the reference's
 * load path does (rasterizer/hair_style.cc:75-77, hair_style.cc:283-291,
 * 331-357):
 * The counts are read from counts, or, when it is NULL, from the packed
 * tangent_acc of the voxelizer.
 *   density_u8   = min(count, 255), then, if normalize != 0,
 *                  HairStyle::Volume::normalize (min/max stretch to 0..255,
 *                  float scale, truncation; a constant volume becomes 0);
 *   tangent_i8x4 = trunc(sum / count * 127) per component, w = 0, from the
 *                  fixed-point sums of tangent_acc; 0 for empty voxels and
 *                  for voxels of 32768 or more contributions (nullable;
 *                  needs tangent_acc);
 *   density_f32  = density_u8 / 255 (nullable; the R8_UNORM value);
 *                  density_u8 must be 4-byte and counts, tangent_acc,
 *                  density_f32 and tangent_i8x4 16-byte aligned (device
 *                  allocations are);
 *   cells        = corner-packed density (nullable): for every cell
 *                  (i,j,k), i,j,k in [-1, n-1], the 8 texels of its
 *                  trilinear footprint as one uint64 (byte b = texel
 *                  (i+(b&1), j+(b>>1&1), k+(b>>2)), 0 outside the volume);
 *                  (nx+1)*(ny+1)*(nz+1) entries, index (i+1) + (j+1)*(nx+1)
 *                  + (k+1)*(nx+1)*(ny+1).  The raymarcher reads this layout:
 *                  one 8-byte gather per trilinear tap;
 *   occupancy    = the brick occupancy of the cells (nullable; needs cells;
 *                  vkhr_occupancy_words(nx, ny, nz) uint32 words).  Bricks
 *                  are 8x8x8 cells, brick (bx,by,bz) = cell (i+1,j+1,k+1) >> 3
 *                  with index b = bx + BX*(by + BY*bz), BX = (nx+8)/8, BY =
 *                  (ny+8)/8, BZ = (nz+8)/8, B = BX*BY*BZ bricks.  Words
 *                  [0, ceil(B/32)): bit b set when a cell of brick b holds a
 *                  nonzero texel; then ceil(B/8) words of 4-bit distances:
 *                  brick b at bits 4*(b%8) of word b/8, 0 for an occupied
 *                  brick, else the chessboard distance in bricks to the
 *                  nearest occupied brick, capped at 15 (every brick within
 *                  distance - 1 is empty; all 0 when a side exceeds 2040, and
 *                  then nothing is skipped); then ceil(B/4) words the finalize
 *                  works in.  The raymarcher skips the gathers of samples in
 *                  empty bricks and, on the power-of-two sample grid, jumps
 *                  over empty brick regions (their taps read 0: the result is
 *                  unchanged). */
int vkhr_volume_finalize(const uint32_t* counts, const int64_t* tangent_acc,
                         uint32_t nx, uint32_t ny, uint32_t nz, int normalize,
                         uint8_t* density_u8, int8_t* tangent_i8x4,
                         float* density_f32, uint64_t* cells, uint32_t* occupancy,
                         vkhr_stream stream);

/* vkhr_volume_finalize that allocates nothing: same arguments, same
 * results, but the min/max slots of the normalize (2 KiB) live in a static
 * per-device array of VKHR_STATIC_SCRATCH_BYTES (4 KiB), resident from
 * library load until the process ends, instead of the per-stream scratch.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight per-frame serving
 * loops, harnesses that forbid hipMalloc.  The call makes no hipMalloc or
 * hipFree and keeps no per-stream area (vkhr_scratch_bytes stays 0).  Every
 * call on a device shares the one array (its slots are re-initialised by a
 * kernel at the start of each call), so two calls that overlap on one
 * device -- concurrent streams or threads -- race; order them.  Returns
 * VKHR_ERR_OVERFLOW if the scratch needed exceeds the array. */
#define VKHR_STATIC_SCRATCH_BYTES 4096
int vkhr_volume_finalize_noalloc(const uint32_t* counts, const int64_t* tangent_acc,
                                 uint32_t nx, uint32_t ny, uint32_t nz, int normalize,
                                 uint8_t* density_u8, int8_t* tangent_i8x4,
                                 float* density_f32, uint64_t* cells, uint32_t* occupancy,
                                 vkhr_stream stream);

/* HairStyle::voxelize_segments (hair_style.cc:296-342) followed by
 * HairStyle::Volume::normalize (:344-357) -- the reference's load-time
 * volume build (rasterizer/hair_style.cc:75-77) -- in one call, with the
 * outputs of vkhr_voxelize_segments + vkhr_volume_finalize and the same
 * bits: counts (nullable, uint32 per voxel), density_u8, tangent_i8x4
 * (nullable; needs tangents), density_f32, cells and occupancy (nullable),
 * as documented there.  The voxel sums are kept in LDS: the index pairs
 * are binned by the 8x8x8-voxel bricks their samples visit, and one
 * workgroup per brick accumulates and finalizes its brick, so nothing of
 * the size of the volume is cleared or read back beyond the outputs.  The
 * bins live in the per-stream scratch (4 entries per pair plus 65536);
 * workspace (16 bytes per voxel, 16-byte aligned, contents undefined
 * afterwards) is used only when the pairs visit more bricks than that
 * (long segments): a flag set on the device then switches to the packed
 * accumulator path of vkhr_voxelize_segments, with the same results.
 * density_u8 is 4-byte and counts, tangent_i8x4 and density_f32 16-byte
 * aligned, as for vkhr_volume_finalize.
 * Same limits as vkhr_voxelize_segments (nx*ny*nz <= 2^24). */
int vkhr_volume_build_segments(const float* xyz, uint32_t n_vertices,
                               const uint32_t* indices, uint32_t n_indices,
                               const float* tangents, const vkhr_aabb* bounds,
                               uint32_t nx, uint32_t ny, uint32_t nz, int normalize,
                               uint32_t* counts, uint8_t* density_u8, int8_t* tangent_i8x4,
                               float* density_f32, uint64_t* cells, uint32_t* occupancy,
                               int64_t* workspace, vkhr_stream stream);

/* Words of the brick occupancy buffer of an nx x ny x nz volume (bitmask,
 * distance nibbles and working space, see vkhr_volume_finalize). */
size_t vkhr_occupancy_words(uint32_t nx, uint32_t ny, uint32_t nz);

/* HairStyle::Volume::downsample (include/vkhr/scene_graph/hair_style.hh:
 * 228-257) with one of these reductions as its filter functor: voxel (i,j,k)
 * of out (nx/2 x ny/2 x nz/2) reduces the 2x2x2 neighbourhood (2i+x, 2j+y,
 * 2k+z) of in (nx x ny x nz, even sizes).  MEAN is the sum / 8 truncated. */
enum { VKHR_DOWNSAMPLE_MAX = 0, VKHR_DOWNSAMPLE_MIN = 1, VKHR_DOWNSAMPLE_MEAN = 2 };
int vkhr_volume_downsample(const uint8_t* in, uint32_t nx, uint32_t ny, uint32_t nz, int filter, uint8_t* out,
                           vkhr_stream stream);

/* filter_volume (share/shaders/volumes/sample_volume.glsl:12-35) evaluated
 * at every texel centre of a float volume, where its taps fall on texel
 * centres: out = sum_o w(o) in[v + o] / sum_o w(o) over the kernel_width^3
 * offsets o (kernel_width odd, zero border), with the shader's weight
 * exp(-(x^2+y^2+z^2) / 2 * s^2), s = kernel_width / 2 / 2.4 (s^2 multiplies,
 * as written in the shader). */
int vkhr_volume_filter(const float* in, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t kernel_width, float* out,
                       vkhr_stream stream);

/* --------------------------------------------------------------- raymarch */

/* The volumes one hair style is rendered from (device pointers).  Its world
 * placement is vkhr_strand_params.volume_bounds, as in the Strand uniform
 * block of the shaders. */
typedef struct vkhr_volume {
    const uint8_t* density;      /* nx*ny*nz R8_UNORM texels (nullable if cells given) */
    const uint64_t* cells;       /* corner-packed density (vkhr_volume_finalize) */
    const int8_t* tangents;      /* nx*ny*nz R8G8B8A8_SNORM texels */
    const float* density_f32;    /* nullable: float densities, used instead of cells when set */
    uint32_t nx, ny, nz;
    uint32_t reserved_;
    const uint32_t* occupancy;   /* nullable: brick occupancy and distances of the cells (vkhr_volume_finalize) */
} vkhr_volume;

/* Per-pixel volume raymarch of one hair style: the raymarcher pass
 * (share/shaders/volumes/volume.frag:31-102 with volume_rendering.glsl:19-52,
 * sample_volume.glsl:7-9, approximate_deep_shadows.glsl:24-36,
 * local_ambient_occlusion.glsl:9-30, kajiya-kay.glsl:7-24, scheme.glsl:7-11)
 * as drawn by vulkan::Volume::draw over the AABB proxy
 * (src/vkhr/rasterizer/volume.cc:104-203).  A pixel gets a fragment when its
 * centre ray from camera->position enters the AABB in front of the eye
 * (CW front faces, back faces culled: no fragments with the eye inside).
 * Trilinear filtering is computed exactly (not by a texture unit).  Volume
 * limits (VKHR_ERR_OVERFLOW beyond): (nx+1)(ny+1)(nz+1) < 2^32,
 * (ny+1)(nz+1) < 2^24 and nx+1 < 2^24.
 *   depth_buffer  nullable per-pixel depth read by the march's early break
 *                 (the subpass depth input); NULL reads 1.0, the cleared value
 *   rgba          W*H float4: the fragment colour clamped to [0,1] (what
 *                 packUnorm4x8 consumes); 0 where no fragment
 *   depth         nullable W*H NDC depth of the surface point; 1 where none
 *   surface_info  nullable W*H int4 {state (0: the ray misses the box,
 *                 1: fragment, 2: marched but discarded), surface step (-1:
 *                 the shader's vec3(0) surface), entry step (-1: none), found} */
int vkhr_raymarch(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                  const vkhr_strand_params* strand, const vkhr_render_params* params, uint32_t width,
                  uint32_t height, const float* depth_buffer, float* rgba, float* depth,
                  int32_t* surface_info, vkhr_stream stream);

/* ------------------------------------------------------ differentiable form */

/* Learnable parameters of the differentiable raymarch: they replace
 * strand->hair_opacity (absorption alpha, coverage and ADSM base),
 * the ADSM strand thickness 11 of volume.frag:78, and
 * strand->hair_shininess (Kajiya-Kay exponent p). */
typedef struct vkhr_diff_params {
    float absorption;
    float adsm_thickness;
    float shininess;
    float reserved_;
} vkhr_diff_params;

/* Bytes of the per-pixel record the differentiable forward leaves for the
 * adjoint (64 per pixel: the surface point and the per-ray scalars; no
 * per-sample state). */
size_t vkhr_march_record_bytes(uint32_t width, uint32_t height);

/* The raymarch (as vkhr_raymarch, on volume->density_f32 or the cells) with
 * the parameters of diff (NULL: the reference's) that also writes record. */
int vkhr_raymarch_record(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                         const vkhr_strand_params* strand, const vkhr_render_params* params,
                         const vkhr_diff_params* diff, uint32_t width, uint32_t height, const float* depth_buffer,
                         float* rgba, float* depth, void* record, vkhr_stream stream);

/* No counterpart in the reference: the adjoint of vkhr_raymarch_record.
 * Given dL/d(rgba) (float4 W*H, the gradient of the clamped output), adds
 * dL/d(density) into d_density (nx*ny*nz floats, volume->density_f32 is
 * required) and dL/d(absorption, adsm_thickness, shininess) into d_params
 * (3 floats); both are accumulated, not cleared.  The surface/isosurface
 * choices are piecewise constant (no gradient); clamps pass gradient strictly
 * inside [0,1].  Rays are re-walked in 32-step chunks from the ray set-up;
 * nothing per sample is stored.  Float atomics: summation order is not
 * deterministic. */
int vkhr_raymarch_adjoint(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                          const vkhr_strand_params* strand, const vkhr_render_params* params,
                          const vkhr_diff_params* diff, uint32_t width, uint32_t height, const void* record,
                          const float* d_rgba, float* d_density, float* d_params, vkhr_stream stream);

/* ------------------------------------------------- strand rasterizer, PPLL */

/* Per-pixel linked list (vulkan::LinkedList, src/vkhr/rasterizer/
 * linked_list.hh:16-30; share/shaders/transparency/ppll.glsl:4-48).
 * nodes holds node_count 12-byte nodes {packUnorm4x8 colour, float depth,
 * uint32 previous node}; heads holds width*height node indices
 * (0xffffffff = end of list); counter is one uint32. */
typedef struct vkhr_ppll {
    uint32_t* heads;
    uint32_t* nodes;
    uint32_t* counter;
    uint32_t node_count;
    uint32_t width, height;
    uint32_t reserved_;
} vkhr_ppll;

/* Optional per-node capture of what the rasterizer inserted (for tests and
 * debugging); each array has node_count entries, any may be NULL. */
typedef struct vkhr_frag_capture {
    uint32_t* segment;   /* index pair that produced the node */
    uint32_t* pixel;     /* linear pixel index */
    float* color;        /* float4 colour before packUnorm4x8 */
} vkhr_frag_capture;

/* vulkan::LinkedList::clear (linked_list.cc:75-98): heads = 0xffffffff,
 * counter = 0. */
int vkhr_ppll_clear(const vkhr_ppll* ppll, vkhr_stream stream);

/* Line rule shared by the shadow-map and colour passes (the reference leaves
 * wide-line coverage to the Vulkan driver; this is the build's definition).
 * A segment is clipped to 0 <= z <= w in clip space and mapped to window
 * coordinates a, b (viewport = the target, depth 0..1); pixel (i, j) with
 * centre c = (i+0.5, j+0.5) gets a fragment when
 *   0 <= t < 1,  t = (c-a).(b-a) / |b-a|^2     (ends owned half-open)
 *   dist(c, line ab) <= line_width / 2
 * Window depth interpolates linearly in t; varyings perspective-correctly. */

/* Shadow-map depth pass (Rasterizer::draw_depth, src/vkhr/rasterizer.cc:
 * 293-315; depth_map.vert:9-12): lines of width line_width under the
 * unbiased light view-projection light_vp (column-major) into a size_w x
 * size_h float depth map, cleared to 1 and kept at the minimum (depth test
 * LESS with writes).  Vertices are world-space (identity model matrix). */
int vkhr_shadow_depth(const float* xyz, uint32_t n_vertices, const uint32_t* indices, uint32_t n_indices,
                      const float* light_vp, float line_width, uint32_t size_w, uint32_t size_h,
                      float* depth_map, vkhr_stream stream);

/* Strand colour pass (vulkan::HairStyle::draw, rasterizer/hair_style.cc:
 * 182-199; strand.vert:21-31; strand.frag:34-87 with gpaa.glsl:33-51,
 * kajiya-kay.glsl:7-24, approximate_deep_shadows.glsl:12-73 (3x3 Gaussian
 * PCF of shadow_map through light->matrix), local_ambient_occlusion.glsl:9-30
 * on the density volume, ppll.glsl:23-48).  Line width is
 * strand->strand_radius; each surviving fragment becomes a PPLL node.
 * Node ids are reserved per block of 256 segments (one counter atomic per
 * block) for every pixel the line rule covers; nodes reserved for fragments
 * the shading then discards (coverage below 0.001, the depth test) are
 * never linked, so the counter can exceed the nodes linked (by 0.3 % at
 * 136k strands / 1080p).  Nodes beyond ppll->node_count are dropped (the
 * reference accepts index == node_count, one past its pool).  depth_buffer
 * (nullable, W*H) is the depth test input (LESS, no writes); NULL reads
 * 1.0.  capture is nullable (entries of unlinked nodes are not written).
 * Limits: shadow maps up to 16384 x 16384 (VKHR_ERR_INVALID beyond), volumes
 * with (nx+1)(ny+1)(nz+1) < 2^32 cells (VKHR_ERR_OVERFLOW beyond). */
int vkhr_raster_strands(const float* xyz, const float* tangents, const float* thickness, uint32_t n_vertices,
                        const uint32_t* indices, uint32_t n_indices, const vkhr_camera* camera,
                        const vkhr_light* light, const vkhr_strand_params* strand,
                        const vkhr_render_params* params, const vkhr_volume* volume, const float* shadow_map,
                        uint32_t shadow_w, uint32_t shadow_h, const float* depth_buffer, const vkhr_ppll* ppll,
                        const vkhr_frag_capture* capture, vkhr_stream stream);

/* PPLL resolve (share/shaders/transparency/resolve.comp:15-102): per pixel a
 * k-buffer of the K=16 newest nodes, up to 1024 further nodes swapped in by
 * depth and blended unsorted, then the k-buffer blended back to front, all
 * onto the background (float4 W*H, nullable: opaque white, the cleared
 * target).  out (float4 W*H) receives the blended colour before the rgba8
 * store; pixels without nodes receive the background.  Launch: 16x16-pixel
 * workgroups that each reserve 64 KB of LDS, so at most two share a CU
 * (fewer lists walked at once keep their node lines in L2). */
int vkhr_ppll_resolve(const vkhr_ppll* ppll, const float* background, float* out, vkhr_stream stream);

/* Raymarch fragments into the PPLL (the raymarch pass's ppll_node_data /
 * ppll_link_node, volume.frag:95-100), alongside the float outputs. */
int vkhr_raymarch_ppll(const vkhr_volume* volume, const vkhr_camera* camera, const vkhr_light* light,
                       const vkhr_strand_params* strand, const vkhr_render_params* params, uint32_t width,
                       uint32_t height, const float* depth_buffer, const vkhr_ppll* ppll, vkhr_stream stream);

/* ------------------------------------------------------------ batched form */

/* One hair style of a batch (no counterpart in the reference, which draws
 * styles one by one: Rasterizer::voxelize / strand_dvr loop over them,
 * src/vkhr/rasterizer.cc:216-230, 326-333).  Inputs and outputs are device
 * buffers the caller owns, laid out as for the per-style entry points. */
typedef struct vkhr_batch_style {
    const float* xyz;
    uint32_t n_vertices;
    uint32_t n_indices;
    const uint32_t* indices;
    const float* tangents;        /* nullable without tangent_acc */
    vkhr_aabb bounds;             /* the style's AABB (vkhr_aabb_compute) */
    uint32_t volume;              /* n of the n^3 volume */
    uint32_t reserved_;
    uint32_t* counts;             /* n^3; nullable with tangent_acc (without it, also the
                                     workspace of the build's overflow path); 16-byte aligned */
    int64_t* tangent_acc;         /* 2 n^3, nullable: the workspace of the build's overflow path
                                     (contents unspecified after a launch); with it, tangents and
                                     tangent_i8x4 are required; 16-byte aligned */
    uint8_t* density;             /* n^3, normalized; 4-byte aligned */
    int8_t* tangent_i8x4;         /* 4 n^3, nullable; needs tangent_acc; 16-byte aligned */
    uint64_t* cells;              /* (n+1)^3 */
    uint32_t* occupancy;          /* vkhr_occupancy_words(n, n, n), nullable */
    float* rgba;                  /* width*height float4 */
    vkhr_camera camera;
    vkhr_light light;
    vkhr_strand_params strand;    /* volume_bounds should equal bounds */
    vkhr_render_params params;
} vkhr_batch_style;

/* Launch shapes, chosen by the autotuner (vkhr_amd/tune.py). */
typedef struct vkhr_tuning {
    uint32_t tile_w, tile_h;      /* pixels per wave tile, tile_w * tile_h == 64 */
    uint32_t blocks_per_cu;       /* persistent raymarch blocks per compute unit; 0: one block
                                     per (style, tile) work item */
    uint32_t groups;              /* styles split into this many groups of consecutive styles
                                     (about equal voxels + segments), pipelined in the graph: a
                                     group's raymarch runs beside the next group's volume build
                                     on a second stream; 0 or 1: one chain; at most 64 */
} vkhr_tuning;

typedef struct vkhr_batch vkhr_batch;

/* Builds the device tables of a batch: the volume build (voxelize_segments
 * + finalize: normalize, tangents, cells) + raymarch at width x height for
 * every style, and the build state of every style (bins, bin entries for
 * 4 runs per index pair + 65536, the uint8 view).  tuning may be NULL (8x8
 * wave tiles, one block per work item). */
int vkhr_batch_create(const vkhr_batch_style* styles, uint32_t n_styles, uint32_t width, uint32_t height,
                      const vkhr_tuning* tuning, vkhr_batch** out);

/* Runs the whole batch as one hipGraph (captured on the first call on a
 * created, non-NULL stream, replayed after): the passes of
 * vkhr_volume_build_segments (bin count, scan, bin fill, bricks summed in
 * LDS, cells with the normalize, distance field), each one launch over the
 * workgroups of all styles, and a raymarch over (style, tile) work items.
 * A style whose runs pass its bin capacity takes the build's workspace
 * path on its own, as vkhr_volume_build_segments does.
 * Per-style results (counts, density, tangent_i8x4, cells, occupancy, rgba)
 * are bit-identical to vkhr_voxelize_segments +
 * vkhr_volume_finalize(normalize=1) + vkhr_raymarch. */
int vkhr_batch_launch(vkhr_batch* batch, vkhr_stream stream);
int vkhr_batch_destroy(vkhr_batch* batch);

/* ----------------------------------------------------- batched adjoint form */

/* No counterpart in the reference (which has no gradients): one training
 * step of the differentiable raymarch for many renders of several styles --
 * the forward (vkhr_raymarch_record) of every render, the L2 loss against
 * its target image, and the adjoint, with the density / parameter
 * gradients of all renders of a style summed -- as one graph launch.
 *
 * A style: its float density volume (volume.density_f32 and tangents
 * required) and its gradient outputs, overwritten by each launch with the
 * gradient of the sum of its renders' losses: d_density (nx*ny*nz floats)
 * and d_params (3 floats: absorption, ADSM thickness, Kajiya-Kay exponent). */
typedef struct vkhr_adjoint_style {
    vkhr_volume volume;
    float* d_density;
    float* d_params;
} vkhr_adjoint_style;

/* A render of a style: the view and parameters of vkhr_raymarch_record, the
 * target image (width*height float4), the forward image it writes (rgba,
 * width*height float4), its record (vkhr_march_record_bytes) and its loss
 * (1 float, the sum of squared differences over all channels; nullable). */
typedef struct vkhr_adjoint_render {
    uint32_t style;
    uint32_t reserved_;
    vkhr_camera camera;
    vkhr_light light;
    vkhr_strand_params strand;
    vkhr_render_params params;
    vkhr_diff_params diff;
    const float* target;
    float* rgba;
    void* record;
    float* loss;
} vkhr_adjoint_render;

typedef struct vkhr_adjoint_batch vkhr_adjoint_batch;

/* Builds the device tables and the per-style cell accumulators (8 floats
 * per cell of each style, owned by the batch). */
int vkhr_adjoint_batch_create(const vkhr_adjoint_style* styles, uint32_t n_styles,
                              const vkhr_adjoint_render* renders, uint32_t n_renders, uint32_t width,
                              uint32_t height, vkhr_adjoint_batch** out);

/* Runs the step as one hipGraph (captured on the first call on a created,
 * non-NULL stream): clear, forward of all renders, adjoint of all renders
 * (the loss gradient 2 (rgba - target) is formed in the adjoint), and the
 * fold of each style's cells into its d_density.  The images and records
 * are bit-identical to per-render vkhr_raymarch_record; the gradients equal
 * the sums of per-render vkhr_raymarch_adjoint up to float summation order. */
int vkhr_adjoint_batch_launch(vkhr_adjoint_batch* batch, vkhr_stream stream);
int vkhr_adjoint_batch_destroy(vkhr_adjoint_batch* batch);

/* ------------------------------------------------------------------------
 * Strand ray tracer: the reference's Embree renderer (vkhr::Raytracer,
 * src/vkhr/ray_tracer.cc:28-169; embree::HairStyle, ray_tracer/hair_style.cc:
 * 11-110), a ground truth for the raymarcher's shadows and AO.  Segments
 * are flat linear curves (ray-facing ribbons, radius linear along the
 * segment) -- RTC_GEOMETRY_TYPE_FLAT_LINEAR_CURVE with vertex positions and
 * the per-vertex thickness as radius (HairStyle::create_position_thickness_
 * data, scene_graph/hair_style.cc:451-464).  A ray (o, d) hits segment
 * (p0, r0)-(p1, r1) when, with the segment projected on the plane through o
 * perpendicular to d, the point at u = clamp(-x0.(x1-x0) / |x1-x0|^2, 0, 1)
 * lies within r0 + u (r1 - r0) of o and its depth t = (p(u)-o).d / d.d is in
 * [tnear, tfar]; the closest hit breaks ties to the lowest segment index.
 * ---------------------------------------------------------------------- */

/* Camera::ViewingPlane (include/vkhr/scene_graph/camera.hh:70-75, built by
 * camera.cc:251-266): pixel (i, j) of the frame looks along
 * (i + jx) x + (j + jy) y + z from point.  x is the camera's left, as in the
 * reference, so column 0 of the traced frame is the right edge of the
 * rasterized one. */
typedef struct vkhr_viewing_plane {
    float x[3], y[3], z[3], point[3];
} vkhr_viewing_plane;

typedef struct vkhr_trace_params {
    uint32_t method;     /* Raytracer::VisualizationMethod: 0 shaded, 1 combined shadows,
                            2 direct shadows, 3 ambient occlusion */
    int32_t shadows_on;  /* Raytracer::shadows_on */
    float ao_radius;     /* Raytracer::ao_radius (2.5 in the reference) */
    float exponent;      /* embree::HairStyle::hair_exponent (50) */
    float diffuse[3];    /* hair_diffuse: the style's default colour */
    uint32_t seed;       /* random draws: a counter-based hash of (seed, pixel, frame, draw) */
    uint32_t frame;      /* sample index */
    int32_t ignore_self; /* 0: the reference's rule (shadow and AO rays may hit the segment they
                            start on, beyond tnear = 1e-6); 1: they skip it */
} vkhr_trace_params;

typedef struct vkhr_bvh vkhr_bvh;

/* Builds the BVH of the segments of one style on the device (a linear BVH:
 * Morton-sorted leaves, binary radix tree, bottom-up boxes); the positions
 * and radii are copied into it, so the inputs may change afterwards.
 * Synchronises the stream (load-time call); allocates. */
int vkhr_bvh_build(const float* xyz, const float* radius, uint32_t n_vertices, const uint32_t* indices,
                   uint32_t n_indices, vkhr_stream stream, vkhr_bvh** out);
int vkhr_bvh_destroy(vkhr_bvh* bvh);

/* Segment count and the box of the segment centroids (min xyz, max xyz). */
int vkhr_bvh_info(const vkhr_bvh* bvh, uint32_t* n_segments, float* bounds);

/* Copies the tree (for structural checks; each pointer nullable):
 *   nodes         (n-1) x 16 floats: left box lo xyz + left child, left box
 *                 hi xyz + right child, right box lo xyz + 0, right box hi
 *                 xyz + 0 (children as uint32 bits, 0x80000000 | leaf for
 *                 leaves; node 0 is the root)
 *   parent        (n-1) internal then n leaf parents (0xffffffff at the root)
 *   leaf_segment  n: the index pair of each leaf */
int vkhr_bvh_nodes(const vkhr_bvh* bvh, float* nodes, uint32_t* parent, uint32_t* leaf_segment,
                   vkhr_stream stream);

/* Traces n rays (8 floats each: origin xyz, tnear, direction xyz, tfar):
 * the closest hit (any_hit == 0: t, u, segment) or whether any hit exists
 * (any_hit != 0: segment of some hit).  Misses give segment 0xffffffff; t_out
 * and u_out are nullable. */
int vkhr_trace_rays(const vkhr_bvh* bvh, const float* rays, uint32_t n_rays, int any_hit, float* t_out,
                    float* u_out, uint32_t* segment_out, vkhr_stream stream);

/* Adds one sample per pixel of Raytracer::draw (ray_tracer.cc:76-113) to
 * accum (width x height x 3 doubles, row 0 at the top: the back buffer):
 * a jittered primary ray; at its hit a shadow ray (light_shading,
 * :115-140 -- its direction is the jittered light position itself, as in the
 * reference) and, except for direct shadows, one ambient-occlusion ray in a
 * random direction that scores 2 when its closest hit is farther than
 * ao_radius and 0 otherwise (:142-160); a miss scores (1, 1, 1).  The
 * average over samples is the reference's frame.  Differences: the draws
 * come from a seeded counter hash (the reference shares one unsynchronised
 * xorshift state among threads), and where the reference would produce NaN
 * (a negative 1 - cos^2 from rounding, or a negative cosine sum raised to
 * the exponent) the sine or the specular term is 0.  indices are the
 * style's index pairs (segment -> vertices, for the tangent at the hit);
 * tangents are needed for method 0. */
int vkhr_raytrace(const vkhr_bvh* bvh, const float* tangents, const uint32_t* indices,
                  const vkhr_viewing_plane* plane, const vkhr_light* light, const vkhr_trace_params* params,
                  uint32_t width, uint32_t height, double* accum, vkhr_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* VKHR_HIP_H */
