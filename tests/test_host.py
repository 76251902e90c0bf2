"""CPU tests of the host library: bit-identical BVH build, scene-file reader,
scene API, option validation, and the C-ABI surface."""
import ctypes as C

import numpy as np
import pytest

import zigtrace_amd as zt
from zigtrace_amd import _native as N
from oracle import oracle as O

from .conftest import scene_text
from .helpers import oracle_of, random_shapes


@pytest.mark.parametrize("name", ["cornell", "spheres", "glass", "lamps"])
def test_example_scene_bvh_bit_identical_to_oracle(name):
    text = scene_text(name)
    s = zt.parse(text)
    ref = O.OracleScene.parse(text)
    a = s.view_arrays()
    nodes, idx = ref.bvh()
    assert a["nodes"].tobytes() == nodes.tobytes()
    assert a["indices"].tobytes() == idx.tobytes()
    mats, shapes, planes, cam, bg = ref.flat()
    assert a["materials"].tobytes() == mats.tobytes()
    assert a["shapes"].tobytes() == shapes.tobytes()
    assert a["planes"].tobytes() == planes.tobytes()
    assert a["camera"].tobytes() == cam.tobytes()
    assert a["background"].tobytes() == bg.tobytes()
    lights = ref.lights()
    assert np.array_equal(a["lights"]["shape"], lights["shape_index"])
    assert np.array_equal(a["lights"]["weight"], lights["weight"])
    assert np.array_equal(a["cdf"], lights["cdf"]) and a["light_total"] == lights["total"]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 50, 257, 400, 3000])
def test_random_bvh_bit_identical_to_oracle(n):
    rng = np.random.default_rng(n)
    shapes = random_shapes(rng, n)
    nodes, idx = zt.bvh_build(shapes)
    on, oi = O.bvh_build(shapes.view(O.SHAPE))
    assert nodes.tobytes() == on.tobytes() and idx.tobytes() == oi.tobytes()


@pytest.mark.parametrize("n,threads", [(70000, 2), (120000, 8), (200000, 0)])
def test_parallel_bvh_build_bit_identical(n, threads):
    """The multi-threaded build numbers nodes as the sequential one appends them:
    same node array and permutation as threads=1 and as the oracle's build."""
    rng = np.random.default_rng(n)
    shapes = random_shapes(rng, n)
    seq_nodes, seq_idx = zt.bvh_build(shapes, threads=1)
    par_nodes, par_idx = zt.bvh_build(shapes, threads=threads)
    assert par_nodes.tobytes() == seq_nodes.tobytes() and par_idx.tobytes() == seq_idx.tobytes()
    on, oi = O.bvh_build(shapes.view(O.SHAPE))
    assert seq_nodes.tobytes() == on.tobytes() and seq_idx.tobytes() == oi.tobytes()


def test_parallel_bvh_build_on_degenerate_input():
    """Many coincident centroids (median-split path) and duplicated triangles."""
    n = 80000
    shapes = np.zeros(n, dtype=N.SHAPE)
    shapes["kind"] = N.SPHERE
    shapes["material"] = 0
    shapes["v"][:, :3] = np.repeat(np.arange(n // 1000)[:, None] * 0.5, 1000, axis=0)[:, :1]
    shapes["v"][:, 3] = 0.25
    a = zt.bvh_build(shapes, threads=1)
    b = zt.bvh_build(shapes, threads=6)
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


def test_coincident_centroids_bvh_bit_identical():  # bvh.zig:356-367: median split path
    shapes = np.zeros(40, dtype=N.SHAPE)
    shapes["kind"] = N.SPHERE
    shapes["material"] = np.arange(40)
    shapes["v"][:, 3] = 1 + np.arange(40) * 0.01
    nodes, idx = zt.bvh_build(shapes)
    on, oi = O.bvh_build(shapes.view(O.SHAPE))
    assert nodes.tobytes() == on.tobytes() and idx.tobytes() == oi.tobytes()
    assert len(nodes) > 1


def test_bvh_structure_valid_permutation_enclosing_bounds():  # bvh.zig:327-354
    rng = np.random.default_rng(11)
    shapes = random_shapes(rng, 257)
    nodes, idx = zt.bvh_build(shapes)
    seen = np.zeros(len(shapes), bool)
    for nd in nodes:
        if nd["count"] == 0:
            for c in (nd["first"], nd["first"] + 1):
                cb = nodes[c]
                assert np.all(np.fmin(cb["min"], nd["min"]) == nd["min"])
                assert np.all(np.fmax(cb["max"], nd["max"]) == nd["max"])
        else:
            for i in idx[nd["first"]:nd["first"] + nd["count"]]:
                assert not seen[i]
                seen[i] = True
    assert seen.all()
    assert len(nodes) <= 2 * len(shapes) - 1


def test_bvh_rejects_unbounded_and_handles_empty():  # bvh.zig:364-372
    with pytest.raises(zt.ZTError) as e:
        zt.bvh_build(zt.plane((0, 0, 0), (0, 1, 0), 0))
    assert e.value.name == "UnboundedShape"
    assert O.bvh_build(zt.plane((0, 0, 0), (0, 1, 0), 0).view(O.SHAPE)) == "UnboundedShape"
    nodes, idx = zt.bvh_build(np.zeros(0, dtype=N.SHAPE))
    assert len(nodes) == 0 and len(idx) == 0


def test_parses_a_complete_scene():  # scenefile.zig:211-244
    src = """# a test scene
image 320 200
samples 8
depth 5
camera from 0 1 5 to 0 0 0 fov 35 aperture 0.2 focus 4.5
background color 0.1 0.2 0.3
material ground lambertian 0.5 0.5 0.5
material steel metal 0.8 0.8 0.8 0.1
material glass dielectric 1.5   # trailing comment
material lamp emissive 4 4 4

sphere 0 1 0 1 glass
plane 0 0 0  0 2 0 ground
triangle -1 0 -1  1 0 -1  0 2 -1 steel
quad -1 3 -1  2 0 0  0 0 2 lamp"""
    s = zt.parse(src)
    i = s.info()
    assert (i["width"], i["height"], i["spp"], i["max_depth"]) == (320, 200, 8, 5)
    cam = s.camera
    assert cam.vfov == 35 and cam.focus_dist == 4.5
    a = s.view_arrays()
    assert a["background"]["kind"][0] == N.BG_SOLID and tuple(a["background"]["a"][0]) == (0.1, 0.2, 0.3)
    assert i["num_materials"] == 4 and a["materials"][1]["param"] == 0.1
    assert i["num_shapes"] == 4 and i["num_planes"] == 1
    assert tuple(a["planes"][0]["v"][3:6]) == (0, 1, 0)
    assert i["built"] == 1


ERROR_CASES = [  # scenefile.zig:246-267
    ("image 10", 1, "unexpected end of line"),
    ("\n\nsphere 0 0 0 1 nope", 3, "undefined material"),
    ("material a lambertian 1 1 1\nmaterial a dielectric 1.5", 2, "material already defined"),
    ("frobnicate", 1, "unknown directive"),
    ("samples 0", 1, "expected a positive integer"),
    ("camera from 0 0 0 to 0 0 0", 1, "camera 'from' and 'to' must differ"),
    ("camera from 0 0 0", 1, "camera needs 'from' and 'to'"),
    ("material m metal 1 1 1 2", 1, "metal fuzz must be in [0, 1]"),
    ("material m lambertian 1 x 1", 1, "expected a number"),
    ("material m lambertian 1 1 1\ntriangle 0 0 0 1 1 1 2 2 2 m", 2, "degenerate triangle"),
    ("depth 3 4", 1, "unexpected extra tokens"),
    ("background gradient", 1, "background must be 'color' or 'sky'"),
    # further messages of scenefile.zig:78-203
    ("material m lambertian 1 1 1\nsphere 0 0 0 0 m", 2, "sphere radius must be non-zero"),
    ("material m lambertian 1 1 1\nplane 0 0 0 0 0 0 m", 2, "plane normal must be non-zero"),
    ("material m lambertian 1 1 1\nquad 0 0 0 1 0 0 2 0 0 m", 2, "degenerate quad"),
    ("camera from 0 0 1 to 0 0 0 fov 180", 1, "fov must be between 0 and 180 degrees"),
    ("camera from 0 0 1 to 0 0 0 aperture -1", 1, "aperture must be non-negative"),
    ("camera from 0 0 1 to 0 0 0 focus 0", 1, "focus distance must be positive"),
    ("camera from 0 0 1 to 0 0 0 zoom 2", 1, "unknown camera parameter"),
    ("camera from 0 1 0 to 0 0 0", 1, "camera 'up' is parallel to the view direction"),
    ("material m dielectric 0", 1, "index of refraction must be positive"),
    ("material m plastic 1 1 1", 1, "unknown material type"),
    ("material m lambertian 1 -1 1", 1, "colour components must be non-negative"),
    ("material m lambertian 1 inf 1", 1, "number must be finite"),
]


@pytest.mark.parametrize("src,line,msg", ERROR_CASES)
def test_reports_errors_with_line_numbers(src, line, msg):
    with pytest.raises(zt.SceneSyntaxError) as e:
        zt.parse(src)
    assert (e.value.line, e.value.message) == (line, msg)
    with pytest.raises(O.SyntaxErr) as e2:  # the oracle's restatement agrees
        O.OracleScene.parse(src)
    assert (e2.value.line, e2.value.message) == (line, msg)


@pytest.mark.parametrize("name", ["cornell", "spheres", "glass", "lamps"])
def test_example_scenes_parse(name):  # scenefile.zig:269-275
    i = zt.parse(scene_text(name)).info()
    assert i["num_shapes"] + i["num_planes"] > 0


def test_scene_routes_planes_and_quads():  # scene.zig:128-145
    s = zt.Scene()
    m = s.add_material(zt.lambertian((1, 1, 1)))
    s.add_sphere((0, 0, -5), 1, m)
    s.add_plane((0, 0, -3), (0, 0, 1), m)
    s.add_quad((-1, -1, -2), (2, 0, 0), (0, 0.5, 0), m)
    s.build()
    i = s.info()
    assert i["num_shapes"] == 3 and i["num_planes"] == 1
    ref = oracle_of(s)
    oi = ref.info()
    assert (oi.shapes, oi.planes) == (3, 1)
    h, _ = ref.intersect(O.rays([0, 0, 0], [0, 0, -1]), 1e-4, 1e9)
    assert abs(h[0]["t"] - 3) < 1e-12
    h, _ = ref.intersect(O.rays([0, -0.75, 0], [0, 0, -1]), 1e-4, 1e9)
    assert abs(h[0]["t"] - 2) < 1e-12 and h[0]["front"] == 1


def test_lights_only_bounded_emissive_shapes():  # light.zig:216-250
    s = zt.Scene()
    white = s.add_material(zt.lambertian((1, 1, 1)))
    dim = s.add_material(zt.emissive((1, 1, 1)))
    bright = s.add_material(zt.emissive((4, 4, 4)))
    dark = s.add_material(zt.emissive((0, 0, 0)))
    s.add_plane((0, 0, 0), (0, 1, 0), dim)
    s.add_sphere((0, 3, 0), 0.5, dim)
    s.add_quad((2, 2.5, -1), (1, 0, 0), (0, 0, 2), bright)
    s.add_sphere((-3, 1, 0), 1, white)
    s.add_sphere((0, 2, 0), -1, dim)
    s.add_sphere((0, 5, 0), 1, dark)
    s.build()
    a = s.view_arrays()
    assert len(a["lights"]) == 3
    sphere_w = 4 * np.pi * 0.25
    assert abs(a["lights"]["weight"][0] - sphere_w) < 1e-12
    assert abs(a["light_total"] - (sphere_w + 8.0)) < 1e-12


def test_render_validates_options_before_anything():  # render.zig:153, 282-288
    s = zt.Scene()  # unbuilt scene: validation still comes first
    for bad in (dict(width=0), dict(height=0), dict(spp=0), dict(tile_size=0)):
        o = dict(width=24, height=16, spp=8, max_depth=8, tile_size=5)
        o.update(bad)
        with pytest.raises(zt.ZTError) as e:
            zt.render(s, zt.Options(**o))
        assert e.value.name == "InvalidOptions"


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = N.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_gpu_entry_points_fail_cleanly_on_bad_handles():
    lib = N.lib()
    assert lib.zt_gpu_render(None, 0, None, None, None, 0) == N.ERR_INVALID_ARGUMENT
    assert lib.zt_gpu_create(0, None) == N.ERR_INVALID_ARGUMENT
    assert lib.zt_gpu_last_error(None) == b"null context"
