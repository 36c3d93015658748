#!/usr/bin/env python3
"""Headline benchmark of the MI355X ReSTIR path.

Workload (one "step" = one frame of the ReSTIR passes at 1920x1080): the seeded terrain scene
(256x256-vertex value-noise heightfield + 500 oriented boxes = 136,050 triangles) lit by 10,000
point lights with log-uniform intensity [0.01, 10]; unbiased ReSTIR with M = 32 initial
candidates, temporal reuse capped at 20*M, two unbiased spatial passes with k = 5 neighbours in a
30 px radius (each re-testing visibility), and final shading; the camera orbits 0.5 deg/frame.
G-buffers of every frame are rendered before the timed region, so inputs are resident in HBM.

Prints one JSON line: ms/frame (value), Mrays/s of the shadow-ray traversal, per-kernel times,
the roofline of the dominant kernel, and the CPU-oracle baseline on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "restir-vulkan_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "ms/frame @1080p to equal RMSE vs FP64 ref: learned vs plain RIS; Mrays/s"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per pixel: compulsory traffic of each pass (each byte read or written once)
BYTES_PER_PX = {"pass_a": 208, "pass_c": 176, "pass_b": 176, "shade": 124}
NODE_BYTES, TRI_BYTES = 80, 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--M", type=int, default=32)
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--spatial-passes", type=int, default=2)
    ap.add_argument("--tile", type=str, default="16x16")
    ap.add_argument("--trace-reps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (x3), whose hardware byte counters "
                         "calibrate the traffic counters of the other kernels")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the per-pixel passes "
              "would shard by screen bands with a 30 px halo, which is not built here.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    from restir_amd import abi, scenes
    from restir_amd.abi import Context, gbuffer_struct, make_uniforms, projection_view
    from restir_amd.pipeline import SceneOnDevice

    W, H, K, WU = args.width, args.height, args.steps, args.warmup
    npx = W * H
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    tx, ty = (int(v) for v in args.tile.split("x"))
    launch = abi.Launch(tx, ty)

    t0 = time.time()
    scene = scenes.terrain(2)
    ctx = Context(0)
    sd = SceneOnDevice(ctx, scene, dev)
    log(f"scene: {scene.n_tris} triangles, {len(scene.point_lights)} point lights, BVH depth {sd.depth} "
        f"({time.time() - t0:.1f}s)")

    nframes = WU + K
    cams = [scenes.orbit_camera(scene.camera, f) for f in range(nframes)]
    pvs = [projection_view(c, W / H) for c in cams]
    G = [torch.zeros(abi.GBUFFER_PLANES * npx, dtype=torch.float32, device=dev) for _ in range(nframes + 1)]
    for f in range(nframes):
        ctx.primary_gbuffer(cams[f], sd.d_tri_material, sd.d_materials, len(scene.materials), G[f + 1],
                            W, H, stream=stream)
    torch.cuda.synchronize()
    log(f"g-buffers: {nframes} frames at {W}x{H} rendered")
    res = [torch.zeros(16 * npx, dtype=torch.int32, device=dev) for _ in range(2)]
    tmp = torch.zeros(16 * npx, dtype=torch.int32, device=dev)
    rgb = torch.zeros(3 * npx, dtype=torch.float32, device=dev)
    flags = abi.VISIBILITY_REUSE_FLAG | abi.TEMPORAL_REUSE_FLAG
    n_sp = args.spatial_passes

    def uniforms(f):
        return make_uniforms(W, H, f + 1, cams[f]["position"], pvs[f - 1] if f > 0 else pvs[f],
                             M=args.M, temporal_mult=20, neighbors=args.k, radius=30.0, flags=flags)

    def frame(f):
        i, p = f % 2, 1 - f % 2
        g = gbuffer_struct(G[f + 1].data_ptr(), W, H)
        gp = gbuffer_struct(G[f].data_ptr(), W, H)
        u = uniforms(f)
        first = tmp if n_sp % 2 == 1 else res[i]
        ctx.initial_temporal(u, g, gp, res[p], first, launch, stream)
        src = first
        for j in range(n_sp):
            dst = res[i] if (n_sp - j) % 2 == 1 else tmp
            ctx.unbiased(u, args.k, j, g, src, dst, launch, stream)
            yield j, dst
            src = dst
        ctx.shade(cams[f]["position"], 1.0, g, res[i], rgb, npx, stream)

    def run(f):
        for _ in frame(f):
            pass

    for f in range(WU):
        run(f)
    torch.cuda.synchronize()
    # timed region: every kernel bracketed by hipExtLaunchKernel start/stop events on `stream`
    ctx.profile_read()
    ctx.profile_enable(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(K):
        run(WU + s)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    ms_per_frame = wall_ms / K
    kernels = {name: {"launches_per_frame": n / K, "ms_per_frame": tot / K, "avg_ms": tot / n}
               for name, (n, tot, _mx) in prof.items()}
    gpu_ms = sum(v["ms_per_frame"] for v in kernels.values())
    log(f"timed {K} frames: {ms_per_frame:.3f} ms/frame (kernels {gpu_ms:.3f}); " +
        ", ".join(f"{k} {v['ms_per_frame']:.3f}" for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_frame"])))

    # ---- the last timed frame again, untimed and bit-identical (same inputs and seeds; frame f
    # reads the previous frame's reservoirs, which it does not overwrite): export each Pass C's
    # queued shadow rays and count node visits / triangle tests of exactly those rays
    f_last = WU + K - 1
    cap = npx * args.k
    q1 = torch.zeros(cap, 4, dtype=torch.float32, device=dev)
    q2 = torch.zeros_like(q1)
    sh = torch.zeros(cap, dtype=torch.int32, device=dev)
    visits = torch.zeros_like(sh)
    tests = torch.zeros_like(sh)
    ray_sets = {"centre": [], "neighbour": []}
    for j, _dst in frame(f_last):
        for which, key in ((0, "centre"), (1, "neighbour")):
            n = ctx.unbiased_rays(which, q1, q2, cap, stream=stream)
            ctx.trace_shadow(q1, q2, n, sh, visits, tests, stream=stream)
            torch.cuda.synchronize()
            ray_sets[key].append({"rays": n, "V": float(visits[:n].double().mean()) if n else 0.0,
                                  "T": float(tests[:n].double().mean()) if n else 0.0,
                                  "occluded": float(sh[:n].double().mean()) if n else 0.0})
    # standalone traversal throughput on the last Pass C's neighbour rays (left in q1/q2)
    nr = ray_sets["neighbour"][-1]["rays"]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ctx.trace_shadow(q1, q2, nr, sh, stream=stream)
    e0.record(stream)
    for _ in range(args.trace_reps):
        ctx.trace_shadow(q1, q2, nr, sh, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    trace_ms = e0.elapsed_time(e1) / args.trace_reps
    Vn, Tn = ray_sets["neighbour"][-1]["V"], ray_sets["neighbour"][-1]["T"]
    rays_per_s = nr / (trace_ms * 1e-3)
    roof_rays = min(HBM_PEAK_GBS * 1e9 / (NODE_BYTES * Vn + TRI_BYTES * Tn),
                    39.3e12 / (56 * Vn + 45 * Tn))
    log(f"traversal: {nr} neighbour rays, V={Vn:.1f} T={Tn:.2f}: {trace_ms:.3f} ms, "
        f"{rays_per_s / 1e6:.0f} Mrays/s, roofline {roof_rays / 1e6:.0f} Mrays/s")

    # ---- roofline of the dominant kernel: algorithmic bytes per launch / average launch time
    def ray_bytes(key):
        sets = ray_sets[key]
        return float(np.mean([r["rays"] * (NODE_BYTES * r["V"] + TRI_BYTES * r["T"]) for r in sets]))
    algo_bytes = {"k_trace_queue(neighbour)": ray_bytes("neighbour"),
                  "k_trace_queue(centre)": ray_bytes("centre"),
                  "k_shade": BYTES_PER_PX["shade"] * npx}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_frame"])
    dom_bytes = algo_bytes.get(dominant)
    achieved = dom_bytes / (kernels[dominant]["avg_ms"] * 1e-3) / 1e9 if dom_bytes else None
    for k, v in kernels.items():
        if k in algo_bytes:
            v["algorithmic_GBps"] = round(algo_bytes[k] / (v["avg_ms"] * 1e-3) / 1e9, 1)
        for key in ("launches_per_frame", "ms_per_frame", "avg_ms"):
            v[key] = round(v[key], 4)

    out = {
        "metric": METRIC,
        "value": round(ms_per_frame, 4),
        "unit": "ms/frame",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": K,
        "warmup": WU,
        "ms_per_step": round(ms_per_frame, 4),
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "unbiased ReSTIR 1920x1080, terrain 136,050 tris, 10k point lights, M=32, "
                        "temporal cap 20M, 2x unbiased spatial k=5 r=30, orbit 0.5 deg/frame",
            "resolution": f"{W}x{H}", "triangles": scene.n_tris,
            "point_lights": int(len(scene.point_lights)), "M": args.M, "k": args.k,
            "spatial_passes": n_sp, "tile": args.tile, "learned_proposal": False,
        },
        "frame_kernel_ms": round(gpu_ms, 4),
        "kernels": kernels,
        "mrays_per_s": round(rays_per_s / 1e6, 1),
        "traversal": {"rays": nr, "ray_set": "last frame's second Pass C neighbour rays",
                      "node_visits_per_ray": round(Vn, 2), "tri_tests_per_ray": round(Tn, 2),
                      "ms": round(trace_ms, 4), "roofline_mrays_per_s": round(roof_rays / 1e6, 1),
                      "frac": round(rays_per_s / roof_rays, 4),
                      "pass_c_rays": ray_sets},
        "roofline": {"kernel": dominant, "bound": "hbm",
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": None,
                     "algorithmic_bytes_per_launch": round(dom_bytes) if dom_bytes else None,
                     "avg_launch_ms": kernels[dominant]["avg_ms"]},
    }
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ctx, scene, sd, cams[-1], args)
    print(json.dumps(out), flush=True)


def cpu_baseline(ctx, scene, sd, cam, args):
    """The CPU oracle on this host's cores: one frame of the same passes at 480x270 (1/16 of the
    1080p pixels, same scene, camera and settings), scaled to 1080p; plus the oracle's traversal
    rays/s on that frame's Pass A rays. The G-buffer input is rendered by the GPU (not timed)."""
    import torch

    import oracle
    from restir_amd import abi
    from restir_amd.abi import make_uniforms

    w, h = 480, 270
    scale = (args.width * args.height) / (w * h)
    orc = oracle.Scene(sd.nodes, sd.tris, scene.point_lights, None, sd.alias)
    gb = torch.zeros(abi.GBUFFER_PLANES * w * h, dtype=torch.float32, device="cuda")
    ctx.primary_gbuffer(cam, sd.d_tri_material, sd.d_materials, len(scene.materials), gb, w, h)
    planes = gb.cpu().numpy()
    g = oracle.gbuffer(planes, w, h)
    pv = abi.projection_view(cam, w / h)
    u = make_uniforms(w, h, 5, cam["position"], pv, M=args.M, temporal_mult=20, neighbors=args.k,
                      radius=30.0, flags=abi.VISIBILITY_REUSE_FLAG | abi.TEMPORAL_REUSE_FLAG)
    prev = np.zeros(w * h, dtype=abi.RESERVOIR_DTYPE)
    t0 = time.perf_counter()
    cur = oracle.pass_a(orc, u, g, g, prev, abi.RESERVOIR_DTYPE)
    for j in range(args.spatial_passes):
        cur = oracle.pass_c(orc, u, args.k, j, g, cur)
    oracle.shade(orc, cam["position"], 1.0, g, cur)
    frame_s = time.perf_counter() - t0
    # traversal sample: this frame's Pass A rays (RIS result before the visibility test)
    P = w * h
    pos = planes[: 3 * P].reshape(3, P).T
    u.flags = 0
    ris = oracle.pass_a(orc, u, g, g, prev, abi.RESERVOIR_DTYPE)
    lit = ris["sumWeights"] != 0
    p1 = np.zeros((int(lit.sum()), 4), dtype=np.float32)
    p2 = np.zeros_like(p1)
    p1[:, :3] = pos[lit]
    p2[:, :3] = ris["pos_lum"][lit, :3]
    t1 = time.perf_counter()
    oracle.trace_shadow(orc, p1, p2, with_counts=False)
    trace_s = time.perf_counter() - t1
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(frame_s * 1e3 * scale, 1), "unit": "ms/frame @1080p (scaled from 480x270)",
            "cores": cores, "kind": "port",
            "sample": f"one frame (pass A + {args.spatial_passes}x pass C + shade) of the same scene at "
                      f"{w}x{h}: {frame_s:.2f} s, x{scale:.0f} to 1080p; traversal of its {len(p1)} "
                      f"pass-A rays: {trace_s:.2f} s",
            "mrays_per_s": round(len(p1) / trace_s / 1e6, 3)}


if __name__ == "__main__":
    main()
