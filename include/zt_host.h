/* zt_host.h -- the host side of zigtrace above the GPU boundary, in C++.
 *
 * Mirrors the reference library API (src/root.zig) for the path-tracing core:
 * Scene construction (scene.zig), Bvh.build (bvh.zig, bit-identical node
 * array), Lights.build (light.zig), the scene-file reader (scenefile.zig) and
 * render.render (render.zig:152), which here runs on the GPU through
 * zt_gpu.h. Names, argument meaning and error behaviour follow the reference;
 * Zig error unions become the ZT_ERR_* codes of zt_types.h.
 */
#ifndef ZT_HOST_H
#define ZT_HOST_H

#include <stdint.h>

#include "zt_gpu.h"
#include "zt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zt_scene zt_scene;

typedef struct zt_scene_info {
  uint64_t num_materials;
  uint64_t num_shapes; /* bounded shapes (Scene.shapes) */
  uint64_t num_planes; /* unbounded shapes (Scene.planes) */
  uint64_t num_nodes;  /* Bvh.nodes.len, 0 before build */
  uint64_t num_lights; /* Lights.items.len */
  uint64_t bvh_depth;  /* Bvh.depth(), bvh.zig:128-137 */
  uint64_t width, height, spp, max_depth; /* Settings, scene.zig:32-37 */
  uint32_t built;
  uint32_t _pad;
} zt_scene_info;

/* Scene.init / deinit (scene.zig:59-70). Defaults: sky background
 * (1,1,1)->(0.5,0.7,1), CameraConfig defaults, Settings 400x225, 32 spp, depth 16. */
zt_scene* zt_scene_create(void);
void zt_scene_destroy(zt_scene* scene);
/* Scene.addMaterial (scene.zig:66-69). */
int zt_scene_add_material(zt_scene* scene, const zt_material* m, uint32_t* id_out);
/* Scene.add (scene.zig:71-78): planes go to the unbounded list. Invalidates the build. */
int zt_scene_add(zt_scene* scene, const zt_shape* shape);
/* Scene.add for n shapes in order (one call instead of n; same result). */
int zt_scene_add_many(zt_scene* scene, const zt_shape* shapes, uint64_t n);
/* Scene.addQuad (scene.zig:82-88): corner + a*u + b*v as two triangles, front face on u x v. */
int zt_scene_add_quad(zt_scene* scene, const double corner[3], const double u[3], const double v[3],
                      uint32_t material);
/* The `camera`, `background` and `settings` fields of Scene. */
int zt_scene_set_camera(zt_scene* scene, const zt_camera_config* cam);
int zt_scene_set_background(zt_scene* scene, const zt_background* bg);
int zt_scene_set_settings(zt_scene* scene, uint64_t width, uint64_t height, uint64_t spp, uint64_t max_depth);
/* Scene.build (scene.zig:92-96): Bvh.build + Lights.build. */
int zt_scene_build(zt_scene* scene);
int zt_scene_get_info(const zt_scene* scene, zt_scene_info* out);
/* The built scene as a zt_scene_view (pointers into the scene, valid until
 * the next mutation). */
int zt_scene_get_view(const zt_scene* scene, zt_scene_view* out);

/* scenefile.parse (scenefile.zig:35-105): returns a built scene, or
 * ZT_ERR_SYNTAX with the 1-based line and the reference's message. */
int zt_scene_parse(const char* src, uint64_t len, zt_scene** out, uint64_t* diag_line, char* diag_msg,
                   uint64_t diag_cap);

/* Bvh.build (bvh.zig:41-67) over any shape list: nodes_out needs 2n-1
 * entries, indices_out n. ZT_ERR_UNBOUNDED_SHAPE for a plane,
 * ZT_ERR_TOO_MANY_PRIMITIVES beyond maxInt(u32)/2. */
int zt_bvh_build(const zt_shape* shapes, uint64_t n, zt_bvh_node* nodes_out, uint64_t* num_nodes,
                 uint32_t* indices_out);
/* Bvh.build with an explicit worker count (no reference counterpart: the
 * reference builds on one thread). threads = 1 is the sequential build,
 * 0 = up to 16 hardware threads; every count gives the same node array and
 * index permutation bit for bit (subtrees own disjoint index ranges and are
 * numbered afterwards in the reference's allocation order). zt_bvh_build and
 * zt_scene_build use 0. */
int zt_bvh_build_threads(const zt_shape* shapes, uint64_t n, uint32_t threads, zt_bvh_node* nodes_out,
                         uint64_t* num_nodes, uint32_t* indices_out);

/* Output stage (image.zig, png.zig), on the host. -------------------------- */
/* Output.encode for `count` linear values (image.zig:30-37): NaN -> 0,
 * exposure, tone map, clamp to [0,1], pow(x, 1/gamma), round half away from
 * zero to 0..255. With count = width*height*3 this is Image.toRgb8
 * (image.zig:64-72). */
int zt_encode_rgb8(const double* linear, uint64_t count, const zt_output* out, uint8_t* rgb);
/* writePpm (image.zig:75-80) and writePng (png.zig:10-57: 8-bit RGB, no
 * filtering, one IDAT of stored deflate blocks) of packed RGB8 into `dst`.
 * The encoded length goes to *len; when dst is NULL or cap < *len nothing is
 * written (ZT_ERR_BUFFER_TOO_SMALL unless dst is NULL). writePng rejects
 * zero or > 2^31-1 dimensions (ZT_ERR_INVALID_DIMENSIONS) and zlib streams
 * over 2^31-1 bytes (ZT_ERR_IMAGE_TOO_LARGE) as the reference does. */
int zt_write_ppm(uint64_t width, uint64_t height, const uint8_t* rgb, uint8_t* dst, uint64_t cap, uint64_t* len);
int zt_write_png(uint64_t width, uint64_t height, const uint8_t* rgb, uint8_t* dst, uint64_t cap, uint64_t* len);

/* render.render (render.zig:152-180) on the GPU: validates options first
 * (ZT_ERR_INVALID_OPTIONS), uploads the scene when it changed, renders and
 * widens the fp32 framebuffer to `pixels` (width*height*3 doubles, the
 * reference's []Vec3, row-major y*width+x). `tuning` may be NULL. Uses the
 * calling thread's default GPU context (device 0). */
int zt_render(zt_scene* scene, const zt_render_options* opts, const zt_gpu_tuning* tuning, double* pixels);

/* The calling thread's default GPU context and the device id of `scene`
 * in it (uploading if the scene changed since the last upload). */
int zt_default_context(zt_ctx** out);
int zt_scene_device_id(zt_scene* scene, zt_ctx* ctx, uint32_t* id_out);

#ifdef __cplusplus
}
#endif

#endif /* ZT_HOST_H */
