"""Bvh.build node arrays of the example scenes and the benchmark workloads against the
committed fixture (tests/golden/bvh_arrays.json, written by make_bvh_fixture.py): size
and SHA-256 of the nodes and of the index permutation, for the oracle's build and for
the product's (the one zt_scene_build runs and zt_gpu_scene_upload is handed)."""
import hashlib
import json
import os

import pytest

import zigtrace_amd as zt
from zigtrace_amd import workloads as W
from oracle import oracle as O

from .conftest import ROOT, scene_text

FIXTURE = json.load(open(os.path.join(ROOT, "tests", "golden", "bvh_arrays.json")))


def digest(nodes, indices):
    return {"nodes": int(len(nodes)), "indices": int(len(indices)),
            "nodes_sha256": hashlib.sha256(nodes.tobytes()).hexdigest(),
            "indices_sha256": hashlib.sha256(indices.tobytes()).hexdigest()}


@pytest.mark.parametrize("name", ["cornell", "spheres", "glass", "lamps"])
def test_example_scene_bvh_matches_fixture(name):
    want = FIXTURE[f"examples/{name}.txt"]
    a = zt.parse(scene_text(name)).view_arrays()
    assert digest(a["nodes"], a["indices"]) == want
    assert digest(*O.OracleScene.parse(scene_text(name)).bvh()) == want


@pytest.mark.parametrize("name", ["sphere_field", "diff_box", "mesh_field"])
def test_workload_bvh_matches_fixture(name):
    want = FIXTURE[f"workloads.{name}"]
    wl = getattr(W, name)()
    a = wl.scene.view_arrays()
    assert digest(a["nodes"], a["indices"]) == want
    assert digest(*O.OracleScene.from_view_arrays(a).bvh()) == want
