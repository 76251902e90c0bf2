// Host scene (scene.zig:39-52) behind the zt_scene handle.
#pragma once

#include <cstdint>
#include <vector>

#include "zt_host.h"

namespace zth {

struct HostLights {  // light.zig:59-63
  std::vector<zt_light> items;
  std::vector<double> cdf;
  double total = 0;
};

// threads: 1 = the sequential build, 0 = up to 16 hardware threads (same bits).
int buildBvh(const zt_shape* shapes, uint64_t n, std::vector<zt_bvh_node>* nodes, std::vector<uint32_t>* indices,
             unsigned threads = 0);
uint64_t bvhDepth(const std::vector<zt_bvh_node>& nodes);
void buildLights(const std::vector<zt_shape>& shapes, const std::vector<zt_material>& materials, HostLights* out);

}  // namespace zth

struct zt_scene {
  std::vector<zt_material> materials;
  std::vector<zt_shape> shapes;  // bounded, in add order
  std::vector<zt_shape> planes;  // unbounded, in add order
  zt_camera_config camera;
  zt_background background;
  uint64_t width = 400, height = 225, spp = 32, max_depth = 16;  // Settings, scene.zig:32-37
  bool built = false;
  std::vector<zt_bvh_node> nodes;
  std::vector<uint32_t> indices;
  zth::HostLights lights;
  uint64_t version = 1;  // bumped by every mutation; the GPU copy is keyed on it
  uint64_t uid = 0;      // unique for the process's lifetime (an address may be reused)

  zt_scene() {
    camera = zt_camera_config{{0, 0, 0}, {0, 0, -1}, {0, 1, 0}, 60, 0, 0, 0, 0};  // camera.zig:8-18
    background = zt_background{ZT_BG_SKY, 0, {1, 1, 1}, {0.5, 0.7, 1.0}};       // scene.zig:48
  }
  void invalidate() {  // scene.zig:98-102
    built = false;
    nodes.clear();
    indices.clear();
    lights = zth::HostLights();
    ++version;
  }
};
