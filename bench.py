"""Benchmark: Mpaths/s of the path-tracing core at 1920x1080, 64 spp.

Workload (one GPU): the triangle-heavy mesh field -- 64 displaced icospheres
(subdivision 5, 1,310,720 triangles) over a Lambertian plane, lit by 64 small
emissive spheres and a sky; depth 16, next-event estimation with MIS, fp32.
A step is one full render of that frame (132,710,400 camera paths) into a
framebuffer that stays in HBM; scene build, upload and autotuning happen
before the timed region.

Reported beside `value` (which is the fastest method's whole-frame rate):
  * per-method Mpaths/s (plain megakernel, wavefront), each with the launch
    configuration the autotuner picked,
  * Mrays/s (extension + shadow rays, counted on the device during the timed
    renders),
  * the roofline of the megakernel: algorithmic bytes per launch (BVH node
    and primitive fetches counted by the oracle on a sample of the same
    frame, 64 B/node, 48 B/triangle, 32 B/sphere, 16 B/plane test, plus the
    12 B/pixel framebuffer write) over the launch's HIP-event duration on the
    context stream the kernel runs on,
  * batch: 256 seeded scenes of uneven size, resolution, spp and depth in one
    persistent launch (and in the wavefront renderer's shared path pool) vs
    one megakernel launch per scene,
  * adjoint: forward + path-replay backward iterations/s of Adam inverse
    rendering on the 256x256, 32 spp differentiable box,
  * the CPU baseline: the oracle (C++ f64 restatement of the reference CPU
    path, tiles + atomic counter, all hardware threads) on a strided pixel
    sample of the same frame and seed, and the image error of the GPU frame
    against those oracle pixels (relative RMSE, fraction of pixels moved by a
    diverged path, extension/shadow ray counts of both),
  * the fp64 verification path's Mpaths/s on the same frame (the reference's
    precision), and the plain megakernel's (one pixel per lane) with the
    wavefront's rate over it,
  * the two smaller configurations (the shipped Cornell box at 400x400, 64 spp;
    the sphere field at 600x340, 16 spp), GPU beside the oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--method auto|megakernel|wavefront]
                       [--no-cpu] [--no-extra] [--no-tune] [--spp S]
       python bench.py --only batch|adjoint|configs|fp64 [--adjoint-steps S]   (one side leg alone:
                                             the batch, the adjoint, the two smaller configurations, or
                                             the fp64 verification path on the benchmark frame)
       python bench.py --copy-probe BYTES   (one device-to-device copy of BYTES, for
                                             calibrating HBM traffic counters)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_NODE, BYTES_TRI, BYTES_SPH, BYTES_PLANE = 64, 48, 32, 16


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_per_path(ctr, spp: int) -> float:
    paths = max(int(ctr["paths"]), 1)
    b = (BYTES_NODE * int(ctr["node_pops"]) + BYTES_TRI * int(ctr["tri_tests"]) + BYTES_SPH * int(ctr["sph_tests"])
         + BYTES_PLANE * int(ctr["plane_tests"]))
    return b / paths + 12.0 / spp


def cpu_baseline(workload, target_seconds: float = 20.0):
    """The oracle on a strided pixel sample of the same frame, all hardware threads:
    (record, counters, rays per path, oracle image, pixel stride)."""
    from oracle import oracle as O

    o = workload.options
    t0 = time.time()
    ref = O.OracleScene.from_view_arrays(workload.scene.view_arrays())
    log(f"[bench] oracle scene built in {time.time() - t0:.1f}s")
    hw = O.hardware_threads()
    threads = hw
    npix = o.width * o.height
    stride = 4096
    for _ in range(3):
        t = time.time()
        img, ctr = ref.render(o.width, o.height, o.spp, o.max_depth, threads=threads, seed=o.seed,
                              pixel_stride=stride)
        dt = time.time() - t
        if dt > target_seconds / 2 or stride == 1:  # keep a sample of at least ~10 s
            break
        rate = int(ctr["paths"]) / max(dt, 1e-9)
        want_paths = rate * target_seconds
        stride = max(1, int(npix * o.spp / max(want_paths, 1)))
    runs = 1
    if stride == 1 and dt < target_seconds / 2:
        # The whole frame is quick: the median of at least 5 more runs (the first was the warm-up).
        times = []
        while len(times) < 5 or (sum(times) < target_seconds / 2 and len(times) < 200):
            t = time.time()
            ref.render(o.width, o.height, o.spp, o.max_depth, threads=threads, seed=o.seed)
            times.append(time.time() - t)
        dt, runs = float(sorted(times)[len(times) // 2]), len(times)
    paths = int(ctr["paths"])
    rays = int(ctr["extend_rays"]) + int(ctr["shadow_rays"])
    return {
        "value": round(paths / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "hardware_threads": hw,
        "kind": "port",
        "sample": f"{'every pixel' if stride == 1 else f'one pixel in {stride}'} of the {o.width}x{o.height} frame, "
                  f"all {o.spp} spp ({paths} paths, {dt:.3f}s" + (f", median of {runs} runs after a warm-up" if runs > 1
                                                                   else "") +
                  f"), seed {o.seed}, C++ f64 restatement of the reference CPU path",
    }, ctr, rays / max(paths, 1), img, stride


def image_error(gpu_img, ref_img, stride: int, gpu_rays=None, ref_ctr=None) -> dict:
    """The GPU frame against the oracle's pixels of the same frame and seed (common random
    numbers: the same per-pixel streams and draw order). `flip_fraction`: pixels some channel
    of which differs by more than 1e-4 relative -- fp32 rounding moves a pixel mean by ~1e-6,
    one diverged path by about a sample's value over spp -- an upper bound on the pixels in
    which a decision (hit, roulette, reflection) went the other way."""
    import numpy as np
    g = gpu_img.reshape(-1, 3).astype(np.float64)
    c = ref_img.reshape(-1, 3)
    sel = np.arange(len(c)) % stride == 0
    g, c = g[sel], c[sel]
    rmse = float(np.sqrt(np.mean((g - c) ** 2)) / np.sqrt(np.mean(c ** 2)))
    tol = 1e-4 * np.maximum(np.abs(c), 1e-3 * float(np.abs(c).mean())) + 1e-7
    flips = float((np.abs(g - c) > tol).any(axis=1).mean())
    out = {"rel_rmse": round(rmse, 7), "flip_fraction": round(flips, 6), "pixels": int(sel.sum())}
    if gpu_rays is not None and ref_ctr is not None and stride == 1:
        ext, sh = int(ref_ctr["extend_rays"]), int(ref_ctr["shadow_rays"])
        out["rays"] = {"gpu_extend": gpu_rays[0], "gpu_shadow": gpu_rays[1], "oracle_extend": ext,
                       "oracle_shadow": sh,
                       "rel_diff": round((abs(gpu_rays[0] - ext) + abs(gpu_rays[1] - sh)) / max(ext + sh, 1), 7)}
    return out


def fixture_bytes_per_path():
    """B_path of the mesh field from tests/golden/roofline_mesh_field.json (oracle counts)."""
    with open(os.path.join(ROOT, "tests", "golden", "roofline_mesh_field.json")) as f:
        fx = json.load(f)
    return fx["bytes_per_path"], fx["rays_per_path"]


def measured_traffic_per_path():
    """HBM bytes per path of the megakernel on the mesh field, from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by the same counter's reading for a copy
    kernel of known size); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "megakernel_mesh_field_1080p64_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_path:"):
            return float(line.split(":", 1)[1])
    return None


def copy_probe(nbytes: int) -> int:
    """One device-to-device copy of nbytes (plus one warm-up): the known-size kernel that HBM
    traffic counters are scaled against."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copied {nbytes} bytes")
    return 0


def bench_batch(zt, N, W, steps: int):
    """256 scenes: one persistent launch (wave-scheduled megakernel over every scene's tiles;
    also the wavefront renderer's shared path pool) vs one megakernel launch per scene."""
    t0 = time.time()
    wls = W.batch_scenes()
    log(f"[bench] batch: {len(wls)} scenes built in {time.time() - t0:.1f}s")
    ctx = zt.Context(0)
    ids = [ctx.upload(w.scene) for w in wls]
    opts = [w.options for w in wls]
    paths = sum(o.width * o.height * o.spp for o in opts)
    outs = [ctx.alloc(o.width * o.height * 12) for o in opts]

    def timed(fn):
        fn()  # warm-up
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        return (time.perf_counter() - t) / steps

    tu_m = zt.Tuning(method=N.METHOD_MEGAKERNEL)
    persistent_s = timed(lambda: ctx.render_batch_device(ids, opts, zt.Tuning(), outs))
    wavefront_s = timed(lambda: ctx.render_batch_device(ids, opts, zt.Tuning(method=N.METHOD_WAVEFRONT), outs))
    per_scene_s = timed(lambda: [ctx.render_device(sid, o, tu_m, d) for sid, o, d in zip(ids, opts, outs)])
    for d in outs:
        ctx.free(d)
    ctx.close()
    return {"scenes": len(wls), "paths": paths, "steps": steps,
            "persistent_megakernel_mpaths_s": round(paths / persistent_s / 1e6, 2),
            "persistent_wavefront_mpaths_s": round(paths / wavefront_s / 1e6, 2),
            "per_scene_megakernel_mpaths_s": round(paths / per_scene_s / 1e6, 2),
            "persistent_over_per_scene": round(per_scene_s / persistent_s, 3),
            "persistent_wavefront_over_per_scene": round(per_scene_s / wavefront_s, 3),
            "prims_total": int(sum(w.scene.info()["num_shapes"] for w in wls))}


def bench_configs(zt, N, W, steps: int = 5):
    """The smaller configurations (the shipped-style Cornell box at 400x400, 64 spp; the
    sphere field at 600x340, 16 spp): megakernel Mpaths/s (autotuned, median of `steps`) and
    the oracle's Mpaths/s on the host's cores beside it."""
    import numpy as np
    out = {}
    for name, wl in (("cornell_400x400_64spp", W.cornell()), ("sphere_field_600x340_16spp", W.sphere_field())):
        o = wl.options
        ctx = zt.Context(0)
        sid = ctx.upload(wl.scene)
        d = ctx.alloc(o.width * o.height * 12)
        tu, _ = ctx.autotune(sid, o, method=N.METHOD_MEGAKERNEL)
        ctx.render_device(sid, o, tu, d)
        kms = []
        for _ in range(steps):
            ctx.render_device(sid, o, tu, d)
            kms.append(ctx.stats()["kernel_ms"])
        st = ctx.stats()
        img = ctx.to_host(d, np.zeros((o.height, o.width, 3), np.float32))
        ctx.free(d)
        ctx.close()
        paths = o.width * o.height * o.spp
        cpu, ctr, _, ref_img, stride = cpu_baseline(wl, target_seconds=6.0)
        km = float(np.median(kms))
        gpu = paths / km / 1e3
        out[name] = {"paths": paths, "gpu_mpaths_s": round(gpu, 1), "gpu_kernel_ms_median": round(km, 3),
                     "gpu_mrays_s": round((st["extend_rays"] + st["shadow_rays"]) / km / 1e3, 1),
                     "cpu": cpu, "gpu_over_cpu": round(gpu / cpu["value"], 1),
                     "image_error": image_error(img, ref_img, stride, (st["extend_rays"], st["shadow_rays"]), ctr),
                     "tuning": {k: v for k, v in tu.__dict__.items() if v}}
        log(f"[bench] {name}: {out[name]}")
    return out


def bench_adjoint(zt, N, W, steps: int):
    """Adam inverse rendering on the differentiable box: forward + backward per iteration."""
    import numpy as np
    from zigtrace_amd.adjoint import DiffRenderer, inverse_render
    wl = W.diff_box()
    ctx = zt.Context(0)
    dr = DiffRenderer(ctx, wl.scene, wl.options)
    truth = dr.get_params()
    rng = np.random.default_rng(5)
    init = truth * rng.uniform(0.7, 1.3, truth.shape)
    o = wl.options
    target = dr.forward(truth, zt.Options(o.width, o.height, 1024, o.max_depth, seed=999)).astype(np.float64)
    inverse_render(dr, target, init, steps=2, lr=1e-2)  # warm-up
    res = inverse_render(dr, target, init, steps=steps, lr=1e-2)
    err0 = float(np.abs(init - truth).mean() / np.abs(truth).mean())
    err1 = float(np.abs(res.params - truth).mean() / np.abs(truth).mean())
    dr.release()
    ctx.close()
    return {"iters_per_s": round(res.iters_per_s, 2), "steps": steps, "width": o.width, "height": o.height,
            "spp": o.spp, "max_depth": o.max_depth, "params": int(truth.size),
            "loss_first": res.losses[0], "loss_last": res.losses[-1],
            "param_rel_err_init": round(err0, 4), "param_rel_err_final": round(err1, 4)}


def bench_fp64(zt, N, W, spp: int) -> dict:
    """The fp64 verification path (library defaults) on the benchmark frame: one
    render after a warm-up; Mpaths/s from the device's kernel time."""
    w = W.mesh_field(spp=spp)
    o = w.options
    ctx = zt.Context(0)
    sid = ctx.upload(w.scene)
    d64 = ctx.alloc(o.width * o.height * 3 * 8)
    tu = zt.Tuning(precision=N.PRECISION_FP64)
    ctx.render_device(sid, o, tu, d64)
    ctx.render_device(sid, o, tu, d64)
    ms = ctx.stats()["kernel_ms"]
    ctx.free(d64)
    ctx.close()
    return {"mpaths_s": round(o.width * o.height * o.spp / ms / 1e3, 2), "kernel_ms": round(ms, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="auto", choices=["auto", "megakernel", "wavefront"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the batch and adjoint legs")
    ap.add_argument("--no-tune", action="store_true", help="library defaults instead of the autotuner")
    ap.add_argument("--spp", type=int, default=64, help="samples per pixel (64 is the benchmark; less for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES")
    ap.add_argument("--only", choices=["batch", "adjoint", "configs", "fp64"], default=None,
                    help="run one side leg alone and print its record (for profiling its kernels); fp64: the "
                         "verification path on the benchmark frame, library defaults, one render after a warm-up")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="Adam steps of the adjoint leg")
    ap.add_argument("--schedule", type=int, default=0,
                    help="schedule with --no-tune (megakernel: 1 one pixel per lane, 0/2 wave-scheduled; wavefront: "
                         "1 one ray per lane, 2 four-launch bounce, 3 two-launch bounce, 4 one-launch bounce, "
                         "0/5 sample passes)")
    ap.add_argument("--refill", type=int, default=0, help="wave-scheduled refill threshold with --no-tune")
    ap.add_argument("--stack", type=int, default=0, help="LDS stack entries with --no-tune")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the path-tracing core is benchmarked on a single "
                                   "MI355X (images shard by pixel rows across GPUs with no exchange; not run here)"}))
        return 2

    import numpy as np

    import zigtrace_amd as zt
    from zigtrace_amd import _native as N
    from zigtrace_amd import workloads as W

    if args.only:
        if args.only == "batch":
            leg = bench_batch(zt, N, W, steps=1)
        elif args.only == "adjoint":
            leg = bench_adjoint(zt, N, W, args.adjoint_steps)
        elif args.only == "fp64":
            leg = bench_fp64(zt, N, W, args.spp)
        else:
            leg = bench_configs(zt, N, W)
        print(json.dumps({args.only: leg}), flush=True)
        return 0

    t0 = time.time()
    w = W.mesh_field(spp=args.spp)
    o = w.options
    log(f"[bench] workload {w.description}: built in {time.time() - t0:.1f}s")
    ctx = zt.Context(0)
    sid = ctx.upload(w.scene)
    upload_ms = ctx.stats()["upload_ms"]
    log(f"[bench] uploaded in {upload_ms:.0f} ms")
    frame_bytes = o.width * o.height * 3 * 4
    dout = ctx.alloc(frame_bytes)
    paths = o.width * o.height * o.spp

    methods = {"megakernel": N.METHOD_MEGAKERNEL, "wavefront": N.METHOD_WAVEFRONT}
    chosen = list(methods) if args.method == "auto" else [args.method]
    results = {}
    for name in chosen:
        tu = zt.Tuning(method=methods[name], stack_entries=args.stack)
        tu.schedule, tu.refill = args.schedule, args.refill
        if not args.no_tune:
            t = time.time()
            tu, probe_ms = ctx.autotune(sid, o, method=methods[name])
            log(f"[bench] {name} autotuned in {time.time() - t:.1f}s: {tu} ({probe_ms:.1f} ms probe)")
        for i in range(args.warmup):
            ctx.render_device(sid, o, tu, dout)
            log(f"[bench] {name} warmup {i + 1}/{args.warmup}: {ctx.stats()['kernel_ms']:.1f} ms")
        kms = []
        t = time.perf_counter()
        for i in range(args.steps):
            ctx.render_device(sid, o, tu, dout)
            kms.append(ctx.stats()["kernel_ms"])
            log(f"[bench] {name} step {i + 1}/{args.steps}: {kms[-1]:.1f} ms")
        wall = time.perf_counter() - t
        st = ctx.stats()  # the last timed render's device counters
        rays = st["extend_rays"] + st["shadow_rays"]
        results[name] = {"ms_per_step": wall * 1e3 / args.steps, "kernel_ms": float(np.mean(kms)),
                         "kernel_ms_median": float(np.median(kms)), "mpaths_s": paths * args.steps / wall / 1e6,
                         "mpaths_s_median": paths / float(np.median(kms)) / 1e3,
                         "mrays_s": rays * args.steps / wall / 1e6,
                         "extend_rays": st["extend_rays"], "shadow_rays": st["shadow_rays"],
                         "tuning": {k: v for k, v in tu.__dict__.items() if v}}
    img = ctx.to_host(dout, np.zeros((o.height, o.width, 3), np.float32))
    assert np.isfinite(img).all()
    best = max(results, key=lambda k: results[k]["mpaths_s"])
    r = results[best]

    fp64, plain = None, None
    if not args.no_extra:  # the verification path (the reference's precision) on the same frame
        d64 = ctx.alloc(o.width * o.height * 3 * 8)
        tu64 = zt.Tuning(precision=N.PRECISION_FP64)
        ctx.render_device(sid, o, tu64, d64)
        k64 = []
        for _ in range(2):
            ctx.render_device(sid, o, tu64, d64)
            k64.append(ctx.stats()["kernel_ms"])
        ctx.free(d64)
        fp64 = {"mpaths_s": round(paths / float(np.median(k64)) / 1e3, 2), "kernel_ms_median": round(float(np.median(k64)), 2),
                "steps": 2, "method": "wave-scheduled megakernel, two-child BVH walked in the reference's order"}
        log(f"[bench] fp64 verification path: {fp64}")
        # The plain megakernel (one pixel per lane, library defaults) on the same frame, beside the wavefront.
        tu1 = zt.Tuning(method=N.METHOD_MEGAKERNEL, schedule=1)
        ctx.render_device(sid, o, tu1, dout)
        k1 = []
        for _ in range(2):
            ctx.render_device(sid, o, tu1, dout)
            k1.append(ctx.stats()["kernel_ms"])
        plain = {"mpaths_s": round(paths / float(np.median(k1)) / 1e3, 2), "kernel_ms_median": round(float(np.median(k1)), 2),
                 "steps": 2, "method": "megakernel, one pixel per lane (schedule 1)"}
        if "wavefront" in results:
            plain["wavefront_over_plain"] = round(results["wavefront"]["mpaths_s"] / plain["mpaths_s"], 3)
        log(f"[bench] plain megakernel: {plain}")

    cpu, ctr, rays_per_path, img_err = None, None, None, None
    if not args.no_cpu:
        cpu, ctr, rays_per_path, ref_img, stride = cpu_baseline(w)
        log(f"[bench] cpu baseline {cpu['value']:.3f} Mpaths/s on {cpu['cores']} threads")
        rm = results.get("megakernel", r)
        img_err = image_error(img, ref_img, stride, (rm["extend_rays"], rm["shadow_rays"]), ctr)
        log(f"[bench] image error vs oracle: {img_err}")
    roofline = None
    if "megakernel" in results:
        if ctr is not None:
            bpp, source = algorithmic_bytes_per_path(ctr, o.spp), "oracle counts on this run's CPU sample"
        else:
            bpp, rays_per_path = fixture_bytes_per_path()
            source = "tests/golden/roofline_mesh_field.json"
        if args.spp != 64:
            bpp = bpp - 12.0 / 64 + 12.0 / o.spp
        km = results["megakernel"]["kernel_ms"]
        achieved = paths * bpp / (km * 1e-3) / 1e9
        tpp = measured_traffic_per_path()
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": round(tpp * paths) if tpp is not None else None,
                    "kernel": "megakernel", "kernel_ms": round(km, 3), "bytes_per_path": round(bpp, 2),
                    "bytes_per_launch": round(paths * bpp), "bytes_source": source,
                    "traffic_source": "profiles/megakernel_mesh_field_1080p64_pmc.txt" if tpp is not None else None}
        for r_m in results.values():  # each method's frame against the same algorithmic bytes
            r_m["roofline_frac"] = round(paths * bpp / (r_m["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    extra = {}
    if not args.no_extra:
        ctx.free(dout)
        dout = None
        ctx.close()
        t = time.time()
        extra["batch"] = bench_batch(zt, N, W, steps=1)
        log(f"[bench] batch leg {time.time() - t:.1f}s: {extra['batch']}")
        t = time.time()
        extra["adjoint"] = bench_adjoint(zt, N, W, steps=args.adjoint_steps)
        log(f"[bench] adjoint leg {time.time() - t:.1f}s: {extra['adjoint']}")
        if not args.no_cpu:
            t = time.time()
            extra["configs"] = bench_configs(zt, N, W)
            log(f"[bench] configs leg {time.time() - t:.1f}s")
    line = {
        "metric": "Mpaths/s @1080p 64spp (plain vs wavefront-batched); fwd+adjoint iters/s",
        "value": round(r["mpaths_s"], 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "mesh_field: 64 displaced icospheres (subdivision 5, 1310720 triangles), "
                               "64 emissive spheres r=0.05, Lambertian plane, sky",
                   "width": o.width, "height": o.height, "spp": o.spp, "max_depth": o.max_depth,
                   "light_sampling": True, "method": best},
        "methods": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in results.items()},
        "mrays_s": round(r["mrays_s"], 2),
        "mrays_source": "extension + shadow rays counted on the device during the timed renders",
        "scene_upload_ms": round(upload_ms, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "image_error": img_err,
        "fp64_verification": fp64,
        "plain_megakernel": plain,
        "library": N.loaded_library(),
        **extra,
    }
    if dout is not None:
        ctx.free(dout)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
