/* zt_gpu.h -- the GPU boundary of zigtrace's path-tracing core (gfx950 / MI355X).
 *
 * This is the C ABI a host (the reference's Zig `render.render`, or the C++
 * host in zt_host.h) calls instead of running the CPU tile loop. Everything
 * below `render.zig:152` -- camera ray generation, BVH traversal, shading,
 * next-event estimation with MIS, Russian roulette and the per-pixel
 * accumulation -- runs on the device. Scene construction (parse, Bvh.build,
 * Lights.build) stays on the host and is handed over as a flat view.
 *
 * Conventions
 *   - Every function returns an int status from zt_types.h (ZT_OK == 0) unless
 *     stated otherwise. No C++ exception crosses this boundary.
 *   - Calls are blocking and synchronous on the context's HIP stream. A
 *     context is not re-entrant: use one per host thread.
 *   - Host buffers passed in are copied; the caller keeps ownership.
 */
#ifndef ZT_GPU_H
#define ZT_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "zt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zt_ctx zt_ctx;

/* A built scene (scene.zig:39-52 after `Scene.build`, scene.zig:92-96)
 * flattened into plain arrays. `shapes` are the bounded shapes in
 * `Scene.shapes` order, `planes` the unbounded ones (scene.zig:71-78).
 * `nodes`/`indices` are the host-built `Bvh` (bvh.zig:35-38); num_nodes == 0
 * with num_shapes > 0 means "no BVH" and selects the brute-force path of
 * scene.zig:109-112. `lights`/`light_cdf`/`light_total` are `Lights`
 * (light.zig:59-63). */
typedef struct zt_scene_view {
  const zt_material* materials;
  uint64_t num_materials;
  const zt_shape* shapes;
  uint64_t num_shapes;
  const zt_shape* planes;
  uint64_t num_planes;
  const zt_bvh_node* nodes;
  uint64_t num_nodes;
  const uint32_t* indices; /* num_shapes entries when num_nodes > 0 */
  const zt_light* lights;
  uint64_t num_lights;
  const double* light_cdf; /* num_lights entries */
  double light_total;
  zt_camera_config camera;
  zt_background background;
} zt_scene_view;

/* GPU-only knobs. None of them changes the image: the same seed gives the
 * same bits for every value (tested). Zero means "default".
 * ZT_METHOD_AUTO (or a NULL tuning) renders with the fastest configuration
 * zt_gpu_autotune has recorded on this context for the scene's size class and
 * the image size, and with the megakernel defaults when none is recorded; the
 * other knobs of an AUTO tuning are then ignored (precision excepted). */
enum { ZT_METHOD_AUTO = 0, ZT_METHOD_MEGAKERNEL = 1, ZT_METHOD_WAVEFRONT = 2 };
enum { ZT_PRECISION_FP32 = 0, ZT_PRECISION_FP64 = 1 };
typedef struct zt_gpu_tuning {
  uint32_t method;      /* ZT_METHOD_* */
  uint32_t precision;   /* ZT_PRECISION_FP32 (product path) or _FP64 (verification path) */
  uint32_t tile_w;      /* megakernel: tile width in pixels; one workgroup per tile (64..256 pixels, 0 = 8x8),
                             except that the wave-scheduled kernel hands tiles of 64..4096 pixels to one wave at
                             a time (0 = 8x8) */
  uint32_t tile_h;      /* megakernel: tile height in pixels */
  uint32_t block_size;    /* wavefront: threads per workgroup of its one-ray-per-lane, shading and
                             shadow-resolution kernels (0 = 256) */
  uint32_t queue_paths;   /* wavefront: path slots in flight (0 = min(pixels, 4M); sample passes: every pixel,
                             up to 64M; schedules 4 and 5 round up to 8 segments of whole 4096-slot ranges) */
  uint32_t variant;       /* megakernel build: 1 = registers unconstrained, 2/3/4 = compiled for >= 4/5/6 waves
                             per SIMD (wave-scheduled over two-child nodes: 2/3/4 waves); 0 = default
                             (wave-scheduled: four-wide 2, two-child fp32 4, fp64 3; one pixel per lane:
                             fp32 4, fp64 1) */
  uint32_t stack_entries; /* traversal stack entries per lane kept in LDS, even, 8..64 (0 = wave-scheduled: 12;
                             one pixel per lane: fp32 8, fp64 16; wavefront: 12 for its wave-scheduled
                             tracer, else 16) */
  uint32_t bvh_width;     /* fp32 traversal: 0 or 4 = four-wide nodes when the scene has them, 2 = two-child
                             nodes (the fp64 path always walks the two-child nodes) */
  uint32_t schedule;      /* megakernel: 1 = one pixel per lane, every lane's ray finished before any lane shades;
                             2 = wave-scheduled: wave-uniform node / leaf passes, lanes shade and take new
                             rays or pixels while other lanes keep traversing (four-wide nodes, or two-child
                             nodes walked in the reference's order: fp64 and fp32 with bvh_width 2);
                             0 = 2 where it applies (not for a scene without a BVH), else 1.
                             wavefront: 1 = one ray per lane; 2 = the queue-fed tracer under the same
                             wave-uniform passes (fp32 over four-wide nodes), with shading and shadow
                             resolution in kernels of their own (four launches per bounce); 3 = that tracer
                             shading each hit itself (two launches per bounce); 4 = the whole bounce (trace,
                             shade, next-event ray) in one launch, slots and pixels in 8 XCD segments and
                             bands; 5 = sample passes: the queue holds pixels and one launch runs a whole
                             sample of each; 0 = 5 where it applies, else 1 */
  uint32_t refill;        /* wave-scheduled: lanes waiting for work before the wave stops to serve them,
                             1..64 (0 = 32; fp64: 40; wavefront sample passes: 56 for one scene, 40 for a
                             batch) */
  uint32_t node_bytes;    /* fp32 four-wide nodes: 0 or 128 = the fp32 boxes themselves, 64 = 64-byte nodes
                             (8-bit child boxes, decoded to boxes that contain the fp32 boxes) */
  uint32_t tile_order;    /* wave-scheduled megakernel: 1 = tiles handed out in scan order within each band,
                             2 = most expensive first (a one-sample probe render without next-event rays counts
                             the extension rays of every tile, then the tiles of each band are sorted by that
                             count), 0 = 2 for a single scene with at least 16 samples per pixel and more
                             tiles than the GPU holds waves at once, else 1.
                             wavefront sample passes: the same order for the pixels the slots take (8x8 tiles),
                             0 = 2 for a single scene with at least 16 samples per pixel, else 1 */
} zt_gpu_tuning;

/* Device-side timing and integer work counters of the last render call. */
typedef struct zt_gpu_stats {
  double kernel_ms;    /* HIP-event time of the render's kernels on the context stream */
  double upload_ms;    /* host wall time of the last scene upload */
  uint64_t launches;   /* kernel launches in the last render */
  uint64_t paths;      /* camera samples traced */
  uint64_t extend_rays;
  uint64_t shadow_rays;
} zt_gpu_stats;

/* Context ---------------------------------------------------------------- */
/* Owns a HIP stream on `device`, the uploaded scenes and scratch buffers.
 * Replaces the thread pool created at render.zig:169-177. */
int zt_gpu_create(int device, zt_ctx** out);
void zt_gpu_destroy(zt_ctx* ctx);
/* Message for the last failing call on this context (never NULL). */
const char* zt_gpu_last_error(const zt_ctx* ctx);
/* The context's HIP stream (a hipStream_t), for callers that time or order work. */
void* zt_gpu_stream(zt_ctx* ctx);
int zt_gpu_get_stats(const zt_ctx* ctx, zt_gpu_stats* out);

/* Device memory owned by the caller (e.g. an output framebuffer that stays
 * in HBM). No reference counterpart: the reference returns a host Image. */
int zt_gpu_device_alloc(zt_ctx* ctx, uint64_t bytes, void** out);
int zt_gpu_device_free(zt_ctx* ctx, void* ptr);
int zt_gpu_copy_to_host(zt_ctx* ctx, void* host, const void* device, uint64_t bytes);
int zt_gpu_copy_to_device(zt_ctx* ctx, void* device, const void* host, uint64_t bytes);

/* Scenes ----------------------------------------------------------------- */
/* Copies a built scene to the device in the GPU layout (64-byte two-child
 * BVH nodes, leaf-ordered primitives, fp32 and on request fp64 copies).
 * Validates every index; returns ZT_ERR_BAD_SCENE on inconsistency. Also
 * builds the four-wide tree (128-byte nodes) unless it would be deeper than
 * the traversal stack allows or hold more than 2^25 nodes; such a scene is
 * rendered over the two-child tree. */
int zt_gpu_scene_upload(zt_ctx* ctx, const zt_scene_view* view, uint32_t* scene_id);
int zt_gpu_scene_release(zt_ctx* ctx, uint32_t scene_id);

/* Rendering -------------------------------------------------------------- */
/* render.zig:152-180. Validation matches render.zig:153 (width, height, spp
 * or tile_size == 0 -> ZT_ERR_INVALID_OPTIONS, checked first). Writes
 * width*height*3 values, row-major y*width+x, RGB: float for
 * ZT_PRECISION_FP32, double for ZT_PRECISION_FP64. `out_on_device` != 0 means
 * `out` is device memory (no host copy). `tuning` may be NULL. */
int zt_gpu_render(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                  void* out, int out_on_device);

/* Many renders in one persistent pass (no reference counterpart: the
 * reference renders one scene per call). Scene i is rendered with opts[i]
 * into outs[i] exactly as zt_gpu_render would (same bits). fp32 scenes that
 * all have four-wide BVHs go through one wave-scheduled megakernel launch
 * over the tiles of every scene (unless tuning asks for ZT_METHOD_WAVEFRONT
 * or schedule 1); otherwise the wavefront renderer's pool of path slots
 * moves across scenes as pixels complete. Either way scenes of uneven size,
 * spp and depth share the GPU without per-scene launches. */
int zt_gpu_render_batch(zt_ctx* ctx, uint32_t n, const uint32_t* scene_ids, const zt_render_options* opts,
                        const zt_gpu_tuning* tuning, void* const* outs, int out_on_device);

/* Autotuner (no reference counterpart). Renders the scene at the options'
 * resolution with min(spp, probe_spp) samples (probe_spp 0 = 16) under each
 * candidate launch configuration of `method` (AUTO: both methods) --
 * megakernel: build variant x LDS stack entries x tile shape; wavefront:
 * block size x path-slot pool x LDS stack entries -- times them with HIP
 * events, records the fastest per method in the context (keyed by method,
 * precision, image size and the scene's primitive count class) and writes the
 * overall fastest to *best and its time to *best_ms (either may be NULL).
 * Later AUTO renders of matching scenes use the record. */
int zt_gpu_autotune(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, uint32_t method,
                    uint32_t precision, uint32_t probe_spp, zt_gpu_tuning* best, double* best_ms);

/* Output.encode (image.zig:30-37) of `count` linear values on the device:
 * the framebuffer of zt_gpu_render (float, or double with
 * ZT_PRECISION_FP64) to RGB8, as zt_encode_rgb8 does on the host (f64
 * arithmetic; the device's pow may differ from the host's in the last ulp,
 * which moves a byte only when the value sits on a rounding boundary).
 * `hdr_on_device` / `rgb_on_device` say where the buffers live. No reference
 * counterpart on the device: it spares the HDR copy when only RGB8 is
 * needed. */
int zt_gpu_encode_rgb8(zt_ctx* ctx, const void* hdr, uint32_t precision, uint64_t count, const zt_output* out,
                       uint8_t* rgb, int hdr_on_device, int rgb_on_device);

/* traceWith (render.zig:61-102) for n caller-given rays, each with its own
 * xoshiro256++ state (4 x u64 per ray, advanced in place as the reference
 * advances `*Rng`). Radiance goes to out_rgb (n*3 doubles). Host buffers. */
int zt_gpu_trace_rays(zt_ctx* ctx, uint32_t scene_id, const zt_ray* rays, uint64_t n, uint64_t* rng_states,
                      uint64_t max_depth, uint32_t light_sampling, uint32_t precision, double* out_rgb);

/* Scene.intersect (scene.zig:108-116) for n rays: mode 0 = BVH + planes as
 * the renderer traverses it by default (fp32: the four-wide BVH in 128-byte
 * nodes when the scene has one), mode 1 = brute force over every primitive
 * (geometry.zig:129-139), mode 2 = the two-child BVH + planes, mode 3 = the
 * four-wide BVH in 64-byte nodes + planes. */
int zt_gpu_intersect(zt_ctx* ctx, uint32_t scene_id, const zt_ray* rays, uint64_t n, double t_min, double t_max,
                     int mode, uint32_t precision, zt_hit* out);

/* Rng.forStream(seed, streams[i]) followed by `draws` calls of next()
 * (rng.zig:22-38) per stream: out[i*draws + k]. kind 0 = next() bits,
 * kind 1 = the fp32 uniform the device path uses, stored as its f32 bit
 * pattern in the low 32 bits. */
int zt_gpu_rng(zt_ctx* ctx, uint64_t seed, const uint64_t* streams, uint64_t n, uint32_t draws, int kind,
               uint64_t* out);

/* Differentiable rendering ---------------------------------------------- */
/* No reference counterpart (the reference is not differentiable): the
 * forward/backward pair of an autograd Function over the scene's colours.
 * Parameters, 3 floats each in material order: the albedo of every
 * Lambertian and metal material and the radiance of every emissive material
 * (dielectrics have none). Setting them updates the materials and rebuilds
 * the light table as Lights.build would (light.zig:68-94). */
int zt_gpu_param_count(zt_ctx* ctx, uint32_t scene_id, uint32_t* count);
int zt_gpu_param_offsets(zt_ctx* ctx, uint32_t scene_id, int32_t* offsets); /* per material, -1: none */
int zt_gpu_get_params(zt_ctx* ctx, uint32_t scene_id, float* params);
int zt_gpu_set_params(zt_ctx* ctx, uint32_t scene_id, const float* params);
/* Forward: set `params` (if not NULL) and render (as zt_gpu_render). */
int zt_gpu_forward(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                   const float* params, void* out, int out_on_device);
/* Backward: dLoss/dparams for dLoss/dimage (width*height*3 floats, the
 * adjoint of the forward image) at `params` (if not NULL) and the same
 * options and seed. Re-traces every path from its pixel's RNG stream; keeps
 * nothing from the forward call. Requires max_depth <= 32. `grad_on_device`
 * != 0: dL_dimage is device memory. Result: num_params doubles (host). */
int zt_gpu_backward(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts, const zt_gpu_tuning* tuning,
                    const float* params, const float* dL_dimage, int grad_on_device, double* dL_dparams);

/* No counterpart in the reference: zt_gpu_backward for callers that cannot
 * allocate (code under HIP graph capture, tight training or serving loops,
 * harnesses that forbid hipMalloc). It calls no allocating routine: the
 * partial sums live in a fixed 8 MiB device array the library holds on each
 * GPU (set up by zt_gpu_create), the traversal stack entirely in LDS.
 * dL_dimage (width*height*3 floats) and dL_dparams (num_params doubles) are
 * device memory; the parameters are the scene's current ones (set them
 * beforehand with zt_gpu_set_params). Same kernels and same bits as
 * zt_gpu_backward. Rejected with ZT_ERR_INVALID_ARGUMENT: ceil(w/8)*ceil(h/8)
 * * num_params > 1048576, parameters beyond the LDS columns (126 in fp32,
 * 63 in fp64 for 8x8 tiles), a BVH needing more than 64 stack entries. All
 * calls on one GPU share the array: calls that overlap on one GPU (other
 * streams, other contexts, other threads) race. */
int zt_gpu_backward_noalloc(zt_ctx* ctx, uint32_t scene_id, const zt_render_options* opts,
                            const zt_gpu_tuning* tuning, const float* dL_dimage, double* dL_dparams);

#ifdef __cplusplus
}
#endif

#endif /* ZT_GPU_H */
