"""Writes tests/golden/bvh_arrays.json: size and SHA-256 of the Bvh.build node array and
index permutation (the flat little-endian records of include/zt_types.h: 6 f64 bounds,
u32 first, u32 count per node; u32 per index) for the example scenes and the benchmark
workloads, as built by the oracle. Regression fixtures: the hand-derived known answers in
tests/test_bvh_kat.py pin the builder's rules; these pin that neither builder drifts on
real scenes.

    python tests/golden/make_bvh_fixture.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def scenes():
    """(name, oracle scene) pairs, built from the same flat inputs the product uses."""
    from oracle import oracle as O
    from zigtrace_amd import workloads as W
    for name in ("cornell", "spheres", "glass", "lamps"):
        with open(os.path.join(ROOT, "examples", f"{name}.txt")) as f:
            yield f"examples/{name}.txt", O.OracleScene.parse(f.read())
    for name, wl in (("sphere_field", W.sphere_field()), ("diff_box", W.diff_box()), ("mesh_field", W.mesh_field())):
        yield f"workloads.{name}", O.OracleScene.from_view_arrays(wl.scene.view_arrays())


def digest(nodes, indices) -> dict:
    return {"nodes": int(len(nodes)), "indices": int(len(indices)),
            "nodes_sha256": hashlib.sha256(nodes.tobytes()).hexdigest(),
            "indices_sha256": hashlib.sha256(indices.tobytes()).hexdigest()}


def main():
    out = {}
    for name, sc in scenes():
        out[name] = digest(*sc.bvh())
        print(name, out[name]["nodes"], out[name]["nodes_sha256"][:16], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "bvh_arrays.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
