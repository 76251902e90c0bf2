"""Writes tests/golden/roofline_mesh_field.json: the algorithmic bytes per path
of the 1920x1080, 64 spp, depth 16 mesh field (1,310,720 triangles), counted
by the oracle's instrumented traversal on the reference's two-child BVH.

  B_ray  = 64 N_node + 48 N_tri + 32 N_sph + 16 N_plane   (per ray)
  B_path = sum over a path's rays of B_ray + 12 / spp     (framebuffer write)

N_node counts interior-node pops (one 64-byte two-child node read each),
N_tri / N_sph / N_plane primitive tests. The sample is every 16th pixel of the
frame (row-major pixel index), all 64 samples of each. The figure depends on
the algorithm and the scene only, not on the device layout.
Run: python tests/golden/make_roofline_fixture.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from zigtrace_amd import workloads as W  # noqa: E402

STRIDE = 16


def main():
    w = W.mesh_field()
    o = w.options
    ref = O.OracleScene.from_view_arrays(w.scene.view_arrays())
    t = time.time()
    _, ctr = ref.render(o.width, o.height, o.spp, o.max_depth, seed=o.seed, pixel_stride=STRIDE)
    c = {k: int(ctr[k]) for k in ctr.dtype.names}
    paths = c["paths"]
    b = 64 * c["node_pops"] + 48 * c["tri_tests"] + 32 * c["sph_tests"] + 16 * c["plane_tests"]
    out = {
        "workload": w.description, "width": o.width, "height": o.height, "spp": o.spp, "max_depth": o.max_depth,
        "seed": o.seed, "light_sampling": True, "pixel_stride": STRIDE, "counters": c,
        "bytes_per_ray": {"node": 64, "triangle": 48, "sphere": 32, "plane": 16},
        "rays_per_path": (c["extend_rays"] + c["shadow_rays"]) / paths,
        "bytes_per_path": b / paths + 12.0 / o.spp,
        "seconds": round(time.time() - t, 2),
    }
    with open(os.path.join(HERE, "roofline_mesh_field.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
