/* zt_types.h -- flat, C-ABI records shared by the host library (zt_host.h)
 * and the GPU boundary (zt_gpu.h).
 *
 * Each record is the `extern struct` mirror of a Zig type of the reference
 * (Zig unions and optionals are not C-ABI, so tags and flags are explicit).
 * All geometry and colours are f64 exactly as the reference stores them; the
 * GPU library re-lays them out internally.
 */
#ifndef ZT_TYPES_H
#define ZT_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Zig error unions of the reference map onto these. */
enum {
  ZT_OK = 0,
  ZT_ERR_INVALID_OPTIONS = -1,     /* error.InvalidOptions        render.zig:153  */
  ZT_ERR_OUT_OF_MEMORY = -2,       /* error.OutOfMemory           render.zig:154  */
  ZT_ERR_HIP = -3,                 /* a HIP runtime call failed (no reference counterpart) */
  ZT_ERR_BAD_SCENE = -4,           /* inconsistent scene view (index out of range ...)   */
  ZT_ERR_INVALID_ARGUMENT = -5,    /* null pointer / bad handle / bad enum value        */
  ZT_ERR_UNBOUNDED_SHAPE = -6,     /* error.UnboundedShape        bvh.zig:55      */
  ZT_ERR_TOO_MANY_PRIMITIVES = -7, /* error.TooManyPrimitives     bvh.zig:44      */
  ZT_ERR_SYNTAX = -8,              /* error.SyntaxError           scenefile.zig:31 */
  ZT_ERR_INTERNAL = -9,            /* device-side invariant violated (e.g. stack overflow) */
  ZT_ERR_INVALID_DIMENSIONS = -10, /* error.InvalidDimensions     png.zig:12      */
  ZT_ERR_IMAGE_TOO_LARGE = -11,    /* error.ImageTooLarge         png.zig:31      */
  ZT_ERR_BUFFER_TOO_SMALL = -12    /* output buffer shorter than the encoding (length reported) */
};

/* material.zig:15-25 `Material` (tag + payload). */
enum { ZT_LAMBERTIAN = 0, ZT_METAL = 1, ZT_DIELECTRIC = 2, ZT_EMISSIVE = 3 };
typedef struct zt_material {
  uint32_t kind;
  uint32_t _pad;
  double color[3]; /* lambertian/metal albedo, emissive colour; unused for dielectric */
  double param;    /* metal fuzz, dielectric ior; unused otherwise */
} zt_material;

/* geometry.zig:106-110 `Shape` (tag + payload). */
enum { ZT_SPHERE = 0, ZT_TRIANGLE = 1, ZT_PLANE = 2 };
typedef struct zt_shape {
  uint32_t kind;
  uint32_t material; /* MaterialId, geometry.zig:9 */
  double v[9];       /* sphere: cx cy cz radius; triangle: v0 v1 v2; plane: point normal */
} zt_shape;

/* camera.zig:8-18 `CameraConfig`; `focus_dist: ?f64` becomes value + flag. */
typedef struct zt_camera_config {
  double look_from[3];
  double look_at[3];
  double up[3];
  double vfov;       /* degrees */
  double aperture;   /* lens diameter */
  double focus_dist; /* used when has_focus_dist != 0 */
  uint32_t has_focus_dist;
  uint32_t _pad;
} zt_camera_config;

/* scene.zig:15-18 `Background`. */
enum { ZT_BG_SOLID = 0, ZT_BG_SKY = 1 };
typedef struct zt_background {
  uint32_t kind;
  uint32_t _pad;
  double a[3]; /* solid colour, or sky horizon */
  double b[3]; /* sky zenith */
} zt_background;

/* bvh.zig:22-33 `Node`: f64 bounds, leaf `first/count`, interior `first` = left child. */
typedef struct zt_bvh_node {
  double min[3];
  double max[3];
  uint32_t first;
  uint32_t count;
} zt_bvh_node;

/* light.zig:31-36 `Light`; the shape is referenced by its index in the
 * scene's bounded-shape list instead of being copied. */
typedef struct zt_light {
  uint32_t shape;
  uint32_t _pad;
  double radiance[3];
  double weight;
} zt_light;

/* render.zig:12-23 `Options`, field for field (usize -> uint64_t). */
typedef struct zt_render_options {
  uint64_t width;
  uint64_t height;
  uint64_t spp;
  uint64_t max_depth;
  uint64_t threads; /* accepted and ignored by the GPU path (CPU worker count in the reference) */
  uint64_t seed;
  uint64_t tile_size;
  uint32_t light_sampling; /* bool */
  uint32_t _pad;
} zt_render_options;

/* ray.zig:4-14 `Ray`. */
typedef struct zt_ray {
  double origin[3];
  double dir[3];
} zt_ray;

/* geometry.zig:11-30 `Hit`, plus a found flag and the primitive that was hit. */
typedef struct zt_hit {
  double t;
  double point[3];
  double normal[3];
  uint32_t front_face;
  uint32_t material;
  int32_t hit;
  int32_t prim; /* GPU results: device primitive slot hit (planes after the BVH primitives) */
} zt_hit;

/* image.zig:6-38 `Output` (and its `Tonemap` enum). Defaults: none, 2.2, 1.0
 * (the CLI's default tonemap is aces, main.zig:47). */
enum { ZT_TONEMAP_NONE = 0, ZT_TONEMAP_REINHARD = 1, ZT_TONEMAP_ACES = 2 };
typedef struct zt_output {
  uint32_t tonemap;
  uint32_t _pad;
  double gamma;    /* display gamma; 1 leaves values linear */
  double exposure; /* multiplier applied before tone mapping */
} zt_output;

#ifdef __cplusplus
}
#endif

#endif /* ZT_TYPES_H */
