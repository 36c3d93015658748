#!/usr/bin/env python3
"""Headline benchmark of the MI355X ReSTIR path.

Workload (one "step" = one frame at 1920x1080): the seeded terrain scene (256x256-vertex
value-noise heightfield + 500 oriented boxes = 136,050 triangles) lit by 10,000 point lights with
log-uniform intensity [0.01, 10] and 2,000 emissive lamp triangles floating above it (2,000 more
triangles; both kinds sampled from one alias table by power); unbiased ReSTIR with temporal reuse capped at 20*M, two
unbiased spatial passes with k = 5 neighbours in a 30 px radius (each re-testing visibility) and
final shading; the camera orbits 0.5 deg/frame. G-buffers of every frame are rendered before the
timed regions, so inputs are resident in HBM.

Two pipelines are timed: plain RIS with M = 32 candidates from the alias table (the reference's
configuration) and RIS with M = 8 candidates from the online-trained MLP proposal (inference,
Pass A..shading and one 4096-sample Adam step per frame, all inside the timed region). Their
image error against the FP64 ground truth (seeded 16,384-pixel subset, static camera,
independent frames) gives the cost to equal error: frames of an unbiased estimator average to a
target error in proportion to their per-frame relative MSE, so the time to reach any error is
proportional to ms/frame x relMSE.

A second learned setting (M = 12, half the candidates from the alias table) is trained, timed and
measured against the same ground truth and plain frames and reported beside the first
(equal_rmse_tuned); value keeps the workload's setting.

Prints one JSON line: value = the learned pipeline's ms/frame at the plain pipeline's per-frame
error (ms_learned x relMSE_learned / relMSE_plain), the measured ms/frame of both, per-kernel
times, traversal Mrays/s and its roofline, the roofline of the dominant kernel, and the CPU-oracle
baseline on this host's cores.
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
VALU_LANE_OPS = 39.3e12  # fp32 lane-ops/s (non-packed), the traversal's ALU ceiling
# Traversal work per ray: V visits of 128-byte 4-entry nodes and T triangle tests of 48 bytes
# each (V and T counted by the tracer on the same rays); ALU per visit: 4 box tests of
# 16 lane-ops (6 FMA + 6 min/max + 4 compares) + 24 of walk control; 45 per triangle test.
NODE_BYTES, TRI_BYTES = 128, 48
VISIT_OPS, TEST_OPS = 4 * 16 + 24, 45
BYTES_PER_PX_SHADE = 124
# compulsory bytes per pixel of the resampling passes (each buffer byte read or written once):
# Pass A = current G 44 + previous G 36 + previous reservoir 64 + output 64; Pass C (without its
# rays) = G 48 + reservoir in 64 + out 64
BYTES_PER_PX_PASS_A, BYTES_PER_PX_PASS_C = 208, 176
PASS_A_KERNELS = ("k_pass_a", "k_pass_a_ris", "k_pass_a_finish")
PASS_C_KERNELS = ("k_pack_surface", "k_pass_c_merge", "k_pass_c_finish", "k_pass_c")
TRUTH_TRI_SAMPLES = 64  # stratified samples per triangle light in the FP64 ground truth
# Error study: fixed views of the orbit (0.5 deg/frame), the same whatever --steps and --warmup
# are; the proposal trains over the orbit's first ERROR_ORBIT frames.
ERROR_ORBIT = 40
ERROR_VIEWS = (6, 16, 26, 36)
BOOTSTRAP = 2000


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE, each scaled by a plain copy kernel's reading) summarised in
    profiles/traffic_1080p_terrain_pmc.md; None when the summary has no line for it."""
    path = ROOT / "profiles" / "traffic_1080p_terrain_pmc.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_per_launch_bytes:"):
            name, _, val = line.split(":", 1)[1].rpartition("=")
            if name.strip() == kernel:
                return int(val)
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--M", type=int, default=32, help="candidates of the plain pipeline")
    ap.add_argument("--learned-M", type=int, default=8, help="candidates of the learned pipeline")
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--spatial-passes", type=int, default=2)
    ap.add_argument("--tile", type=str, default="16x16")
    ap.add_argument("--trace-reps", type=int, default=20)
    ap.add_argument("--pretrain", type=int, default=1024, help="untimed training frames of the MLP")
    ap.add_argument("--mix", type=float, default=0.1,
                    help="share of the learned candidates drawn from the alias table (bounds q below)")
    ap.add_argument("--train-every", type=int, default=1,
                    help="learned pipeline: one Adam step every this many frames (pretraining too)")
    ap.add_argument("--train-on", choices=("final", "first"), default="final",
                    help="learned pipeline: the reservoirs its Adam step trains on, the frame's final ones "
                         "(after the last Pass C) or the first Pass C's output, which lets the step run on a "
                         "second stream during the later Pass C")
    ap.add_argument("--tuned-M", type=int, default=12,
                    help="the second learned setting's candidates (equal_rmse_tuned)")
    ap.add_argument("--tuned-mix", type=float, default=0.5,
                    help="the second learned setting's share of alias-table candidates")
    ap.add_argument("--no-tuned", action="store_true", help="skip the second learned setting")
    ap.add_argument("--error-frames", type=int, default=256, help="independent frames per method")
    ap.add_argument("--error-pixels", type=int, default=16384)
    ap.add_argument("--error-seed", type=int, default=0,
                    help="selects another set of independent frames for the error study")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-views", action="store_true", help="skip the batched 16-view measurement")
    ap.add_argument("--no-builds", action="store_true", help="skip the scene-builder measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=8, help="1080p frames of the CPU-oracle sample")
    ap.add_argument("--workload", choices=("1080p", "stress4k"), default="1080p",
                    help="stress4k: the 3840x2160 stress workload's learned pipeline (2,333,058 "
                         "triangles, 1M point lights in 4,096 clusters, 4,096-way proposal); "
                         "prints its own JSON line")
    ap.add_argument("--stress-paths", type=int, default=4,
                    help="stress4k: camera paths of 32 learned frames captured into one graph (0: skip)")
    ap.add_argument("--weights", choices=("standard", "reference"), default="reference",
                    help="reference (default): the reference's RIS weights, whose frames converge "
                         "several percent from the FP64 truth; standard: RESTIR_STANDARD_WEIGHTS_FLAG, "
                         "unbiased frames in both pipelines")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (x3), whose hardware byte counters "
                         "calibrate the traffic counters of the other kernels")
    return ap.parse_args()


class Frames:
    """The per-frame call sequence over pre-rendered G-buffers G[f + 1] (G[0] is an empty
    previous frame for frame 0), ping-ponging reservoirs as the reference's App does."""

    def __init__(self, torch, ctx, scene, cams, G, W, H, args, launch, stream):
        from restir_amd import abi
        from restir_amd.abi import make_uniforms, projection_view
        self.torch, self.ctx, self.cams, self.G, self.W, self.H = torch, ctx, cams, G, W, H
        self.args, self.launch, self.stream = args, launch, stream
        self.side = torch.cuda.Stream()  # the learned pipeline's Adam step, when it trains early
        self.abi, self.make_uniforms = abi, make_uniforms
        self.pvs = [projection_view(c, W / H) for c in cams]
        npx = W * H
        dev = G[0].device
        self.res = [torch.zeros(16 * npx, dtype=torch.int32, device=dev) for _ in range(2)]
        self.tmp = torch.zeros(16 * npx, dtype=torch.int32, device=dev)
        self.rgb = torch.zeros(3 * npx, dtype=torch.float32, device=dev)
        self.flags = abi.VISIBILITY_REUSE_FLAG | abi.TEMPORAL_REUSE_FLAG | \
            (abi.STANDARD_WEIGHTS_FLAG if getattr(args, "weights", "reference") == "standard" else 0)
        self.learned = None
        self.launch_a = launch  # Pass A's shape and ray schedule (choose_pass_a_schedule)
        self.normal_format = abi.NORMAL_F32

    def uniforms(self, f, M, seed_frame=None, static=False):
        fr = f + 1 if seed_frame is None else seed_frame
        prev_pv = self.pvs[f - 1] if f > 0 and not static else self.pvs[f]
        return self.make_uniforms(self.W, self.H, fr, self.cams[f]["position"], prev_pv, M=M,
                                  temporal_mult=20, neighbors=self.args.k, radius=30.0, flags=self.flags)

    def gview(self, i):
        return self.abi.gbuffer_struct(self.G[i].data_ptr(), self.W, self.H, normal_format=self.normal_format)

    def frame(self, f, M, *, learned=None, gcur=None, gprev=None, seed_frame=None, train=True):
        """Enqueues frame f; yields after each Pass C (its queues then hold that pass's rays)."""
        ctx, s, n_sp = self.ctx, self.stream, self.args.spatial_passes
        i, p = f % 2, 1 - f % 2
        g = self.gview(f + 1 if gcur is None else gcur)
        gp = self.gview(f if gprev is None else gprev)
        u = self.uniforms(f, M, seed_frame, static=gprev is not None and gprev == gcur)
        first = self.tmp if n_sp % 2 == 1 else self.res[i]
        if learned is not None:
            cl, cq = learned.propose(u, g, M, s)
            ctx.initial_temporal_candidates(u, g, gp, self.res[p], first, cl, cq, self.launch_a, s)
        else:
            ctx.initial_temporal(u, g, gp, self.res[p], first, self.launch_a, s)
        src = first
        # the second Pass C reads the first's visibility-tested output: it does not retrace the
        # segments found unoccluded there (RESTIR_INPUT_VISIBLE_FLAG; the same results)
        later = self.abi.Uniforms.from_buffer_copy(u)
        later.flags |= self.abi.INPUT_VISIBLE_FLAG
        train_now = learned is not None and train and f % self.args.train_every == 0
        early = train_now and self.args.train_on == "first" and n_sp > 1
        for j in range(n_sp):
            dst = self.res[i] if (n_sp - j) % 2 == 1 else self.tmp
            ctx.unbiased(later if j > 0 else u, self.args.k, j, g, src, dst, self.launch, s)
            if early and j == 0:
                # the Adam step on the first Pass C's output, on a second stream while the later
                # passes run (it reads that output, which they only read too, and no buffer of theirs)
                self.side.wait_stream(s)
                learned.train(u, g, dst, 4096, 1e-3, self.side)
            yield j
            src = dst
        ctx.shade(self.cams[f]["position"], 1.0, g, self.res[i], self.rgb, self.W * self.H, s)
        if early:
            s.wait_stream(self.side)
        elif train_now:
            learned.train(u, g, self.res[i], 4096, 1e-3, s)

    def run(self, f, M, **kw):
        for _ in self.frame(f, M, **kw):
            pass

    def choose_pass_a_schedule(self, f, M, learned=None, reps=5):
        """Times Pass A of frame f on both ray schedules (queued: RIS, ray queue, finish; inline:
        one dispatch) and keeps the faster; both give identical reservoirs. Returns the record."""
        torch, ctx, abi = self.torch, self.ctx, self.abi
        table = {}
        for name, sched in (("queued", abi.SCHEDULE_QUEUED), ("inline", abi.SCHEDULE_INLINE)):
            self.launch_a = abi.Launch(self.launch.tile_x, self.launch.tile_y, sched)
            self.run(f, M, learned=learned, train=False)
            torch.cuda.synchronize()
            ctx.profile_read()
            ctx.profile_enable(True)
            for _ in range(reps):
                self.run(f, M, learned=learned, train=False)
            rows = ctx.profile_read()
            ctx.profile_enable(False)
            table[name] = sum(rows[k][1] for k in ("k_pass_a", "k_pass_a_ris", "k_trace_queue(pass_a)",
                                                   "k_pass_a_finish") if k in rows) / reps
        best = min(table, key=table.get)
        self.launch_a = abi.Launch(self.launch.tile_x, self.launch.tile_y,
                                   abi.SCHEDULE_QUEUED if best == "queued" else abi.SCHEDULE_INLINE)
        return {"choice": best, "ms": {k: round(v, 4) for k, v in table.items()}}

    def timed(self, first, count, M, learned=None):
        """count frames from `first`, twice from the same state: eagerly, every kernel timed with
        the library's events (the per-kernel table), then captured into one HIP graph and replayed
        once, as a renderer replays its recorded command buffers (the frame time). The two runs
        compute the same frames bit for bit and leave the reservoirs, and a learned proposal's
        parameters and Adam state, as the eager run left them. Returns (ms/frame, kernels);
        self.last_timing records both times (the eager one alone if the capture fails)."""
        torch, ctx = self.torch, self.ctx
        torch.cuda.synchronize()
        bufs = (*self.res, self.tmp)
        saved = [t.clone() for t in bufs]
        start = learned.state() if learned is not None else None
        ctx.profile_read()
        ctx.profile_enable(True)
        t0 = time.perf_counter()
        for f in range(first, first + count):
            self.run(f, M, learned=learned)
        torch.cuda.synchronize()
        eager = (time.perf_counter() - t0) * 1e3 / count
        prof = ctx.profile_read()
        ctx.profile_enable(False)
        kernels = {name: {"launches_per_frame": n / count, "ms_per_frame": tot / count, "avg_ms": tot / n}
                   for name, (n, tot, _mx) in prof.items()}
        after = [t.clone() for t in bufs]
        end = learned.state() if learned is not None else None

        def set_state(values, st):
            for t, v in zip(bufs, values):
                t.copy_(v)
            if learned is not None:
                learned.load_state(st)
            torch.cuda.synchronize()

        set_state(saved, start)
        own = self.stream
        cs = torch.cuda.Stream()
        cs.wait_stream(torch.cuda.current_stream())
        graph = torch.cuda.CUDAGraph()
        try:
            self.stream = cs
            with torch.cuda.graph(graph, stream=cs):
                for f in range(first, first + count):
                    self.run(f, M, learned=learned)
            self.stream = own
            set_state(saved, start)  # capturing ran nothing, but advanced the host-side seeds
            t0 = time.perf_counter()
            graph.replay()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / count
            same = all(torch.equal(t, v) for t, v in zip(bufs, after))
            self.last_timing = {"frame_time": "one HIP graph of the timed frames, replayed once",
                                "ms_per_frame_graph": round(ms, 4), "ms_per_frame_eager": round(eager, 4),
                                "graph_equals_eager": same}
            if not same:
                raise RuntimeError("the graph replay's reservoirs differ from the eager run's")
        except RuntimeError as e:
            self.stream = own
            torch.cuda.synchronize()
            ms = eager
            self.last_timing = {"frame_time": f"eager (graph capture failed: {str(e)[:200]})",
                                "ms_per_frame_eager": round(eager, 4)}
        del graph
        set_state(after, end)
        return ms, kernels


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the per-pixel passes "
              "would shard by screen bands with a 30 px halo, which is not built here.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.workload == "stress4k":
        return stress_4k(torch, args)
    from restir_amd import abi, scenes
    from restir_amd.abi import Context
    from restir_amd.pipeline import SceneOnDevice

    W, H, K, WU = args.width, args.height, args.steps, args.warmup
    npx = W * H
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    tx, ty = (int(v) for v in args.tile.split("x"))
    launch = abi.Launch(tx, ty)

    t0 = time.time()
    scene = scenes.config_scene("terrain")  # 136,050 + 2,000 lamp triangles, 10k point lights
    ctx = Context(0)
    sd = SceneOnDevice(ctx, scene, dev)
    log(f"scene: {scene.n_tris} triangles, {len(scene.point_lights)} point lights, BVH depth {sd.depth} "
        f"({time.time() - t0:.1f}s)")

    nframes = WU + K
    # the orbit's first ERROR_ORBIT frames also feed the proposal's training and the error study's
    # fixed views, whatever --steps and --warmup are
    n_orbit = max(nframes, ERROR_ORBIT)
    cams = [scenes.orbit_camera(scene.camera, f) for f in range(n_orbit)]
    G = [torch.zeros(abi.GBUFFER_PLANES * npx, dtype=torch.float32, device=dev) for _ in range(n_orbit + 1)]
    for f in range(n_orbit):
        ctx.primary_gbuffer(cams[f], sd.d_tri_material, sd.d_materials, len(scene.materials), G[f + 1],
                            W, H, stream=stream)
    torch.cuda.synchronize()
    log(f"g-buffers: {n_orbit} frames at {W}x{H} rendered")
    fr = Frames(torch, ctx, scene, cams, G, W, H, args, launch, stream)

    # ---- plain pipeline (the reference's configuration)
    for f in range(WU):
        fr.run(f, args.M)
    sched_plain = fr.choose_pass_a_schedule(WU - 1, args.M)
    for t in fr.res:
        t.zero_()
    for f in range(WU):
        fr.run(f, args.M)
    plain_ms, kernels = fr.timed(WU, K, args.M)
    plain_timing = fr.last_timing
    log(f"plain M={args.M}: {plain_ms:.3f} ms/frame (eager {plain_timing['ms_per_frame_eager']:.3f}); " + ", ".join(
        f"{k} {v['ms_per_frame']:.3f}" for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_frame"])))

    # ---- the last timed frame again, untimed and bit-identical (same inputs and seeds; a frame
    # does not overwrite the previous frame's reservoirs it reads): export each Pass C's queued
    # shadow rays and count node visits / triangle tests of exactly those rays
    f_last = nframes - 1
    cap = npx * args.k
    q1 = torch.zeros(cap, 4, dtype=torch.float32, device=dev)
    q2 = torch.zeros_like(q1)
    sh = torch.zeros(cap, dtype=torch.int32, device=dev)
    visits = torch.zeros_like(sh)
    tests = torch.zeros_like(sh)
    ray_sets = {"centre": [], "neighbour": []}
    for _j in fr.frame(f_last, args.M):
        for which, key in ((0, "centre"), (1, "neighbour")):
            n = ctx.unbiased_rays(which, q1, q2, cap, stream=stream)
            ctx.trace_shadow(q1, q2, n, sh, visits, tests, stream=stream)
            torch.cuda.synchronize()
            ray_sets[key].append({"rays": n, "V": float(visits[:n].double().mean()) if n else 0.0,
                                  "T": float(tests[:n].double().mean()) if n else 0.0,
                                  "occluded": float(sh[:n].double().mean()) if n else 0.0})
    torch.cuda.synchronize()
    # standalone traversal throughput on the last Pass C's neighbour rays (left in q1/q2)
    nr = ray_sets["neighbour"][-1]["rays"]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # restir_trace_shadow's default schedule: the passes' persistent queue tracer (rays handed to
    # idle lanes in queue order, triangle tests gathered per wave)
    ctx.trace_shadow(q1, q2, nr, sh, stream=stream)
    e0.record(stream)
    for _ in range(args.trace_reps):
        ctx.trace_shadow(q1, q2, nr, sh, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    trace_ms = e0.elapsed_time(e1) / args.trace_reps
    Vn, Tn = ray_sets["neighbour"][-1]["V"], ray_sets["neighbour"][-1]["T"]
    cpu_rays = (q1[:nr].cpu().numpy(), q2[:nr].cpu().numpy()) if not args.no_cpu_baseline else None
    rays_per_s = nr / (trace_ms * 1e-3)
    # Rays/s roofline: the lane-op model is the binding roof. The node and triangle fetches
    # (128V + 48T bytes per ray) come from L1/L2 and the Infinity Cache: the HBM bytes the counters
    # measure per launch are a few percent of them, so the memory roof, priced at the measured
    # bytes per ray when the profile has them, is far above the ALU one.
    valu_rays = VALU_LANE_OPS / (VISIT_OPS * Vn + TEST_OPS * Tn)
    meas = measured_traffic("k_trace_queue(pass_c)") if W * H == 1920 * 1080 else None
    hbm_rays = HBM_PEAK_GBS * 1e9 / (meas / nr) if meas else None
    roof_rays = min(valu_rays, hbm_rays) if hbm_rays else valu_rays
    log(f"traversal: {nr} neighbour rays, V={Vn:.1f} T={Tn:.2f}: {trace_ms:.3f} ms, "
        f"{rays_per_s / 1e6:.0f} Mrays/s, roofline {roof_rays / 1e6:.0f} Mrays/s")

    # a Pass C's tracer launch traces its centre rays, then the neighbour rays of pixels whose
    # centre ray found the light: per pass, both sets with their V and T weighted by ray count
    ray_sets["pass_c"] = []
    for c, nb in zip(ray_sets["centre"], ray_sets["neighbour"]):
        n = c["rays"] + nb["rays"]
        ray_sets["pass_c"].append({"rays": n, "V": (c["rays"] * c["V"] + nb["rays"] * nb["V"]) / max(n, 1),
                                   "T": (c["rays"] * c["T"] + nb["rays"] * nb["T"]) / max(n, 1)})

    # ---- roofline of the dominant kernel: algorithmic bytes per launch / average launch time
    def ray_bytes(key):
        return float(np.mean([r["rays"] * (NODE_BYTES * r["V"] + TRI_BYTES * r["T"]) for r in ray_sets[key]]))
    algo_bytes = {"k_trace_queue(pass_c)": ray_bytes("pass_c"),
                  "k_shade": BYTES_PER_PX_SHADE * npx}
    def pass_rate(names, bytes_px, passes_per_frame):
        ms = sum(kernels[k]["ms_per_frame"] for k in names if k in kernels) / passes_per_frame
        gbps = bytes_px * npx / (ms * 1e-3) / 1e9
        return {"kernels": [k for k in names if k in kernels], "ms_per_pass": round(ms, 4),
                "algorithmic_bytes_per_px": bytes_px, "achieved_GBps": round(gbps, 1),
                "frac_of_hbm_peak": round(gbps / HBM_PEAK_GBS, 4)}
    resampling = {"pass_a": pass_rate(PASS_A_KERNELS, BYTES_PER_PX_PASS_A, 1),
                  "pass_c_without_rays": pass_rate(PASS_C_KERNELS, BYTES_PER_PX_PASS_C, args.spatial_passes),
                  "shade": pass_rate(("k_shade",), BYTES_PER_PX_SHADE, 1)}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_frame"])
    dom_bytes = algo_bytes.get(dominant)
    dom_ms = kernels[dominant]["avg_ms"]
    dom_traffic = measured_traffic(dominant) if W * H == 1920 * 1080 else None
    if dominant.startswith("k_trace_queue"):
        # VALU-bound: algorithmic lane-ops per launch (88 per 4-entry visit, 45 per triangle
        # test, V and T of this kernel's own rays) over the average launch time
        key = dominant[len("k_trace_queue("):-1]
        ops = float(np.mean([r["rays"] * (VISIT_OPS * r["V"] + TEST_OPS * r["T"]) for r in ray_sets[key]]))
        rays = float(np.mean([r["rays"] for r in ray_sets[key]]))
        ach = ops / (dom_ms * 1e-3) / 1e12
        roofline = {"kernel": dominant, "bound": "valu", "achieved": round(ach, 3),
                    "peak": VALU_LANE_OPS / 1e12, "unit": "Tlane-op/s (fp32, non-packed)",
                    "frac": round(ach * 1e12 / VALU_LANE_OPS, 4), "traffic": dom_traffic,
                    "algorithmic_lane_ops_per_launch": round(ops), "rays_per_launch": round(rays),
                    "avg_launch_ms": round(dom_ms, 4),
                    "rays_per_s": round(rays / (dom_ms * 1e-3)),
                    "roofline_rays_per_s": round(VALU_LANE_OPS * rays / ops),
                    "hbm": ({"achieved_GBps": round(dom_traffic / (dom_ms * 1e-3) / 1e9, 1),
                             "frac": round(dom_traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "bytes": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated"}
                            if dom_traffic else None),
                    "note": "model: 4 box tests of 16 lane-ops (6 FMA, 6 min/max, 4 compares) + 24 of "
                            "walk control per 4-entry node visit, 45 per triangle test; node fetches "
                            "are served by L1/L2, so the HBM roof does not bind"}
    else:
        ach = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_bytes else None
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(ach, 1) if ach else None,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None, "traffic": dom_traffic,
                    "algorithmic_bytes_per_launch": round(dom_bytes) if dom_bytes else None,
                    "avg_launch_ms": round(dom_ms, 4)}

    def rounded(ks):
        for k, v in ks.items():
            if k in algo_bytes:
                v["algorithmic_GBps"] = round(algo_bytes[k] / (v["avg_ms"] * 1e-3) / 1e9, 1)
            for key in ("launches_per_frame", "ms_per_frame", "avg_ms"):
                v[key] = round(v[key], 4)
        return ks

    out = {
        "metric": METRIC,
        "value": round(plain_ms, 4),
        "unit": "ms/frame",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": K,
        "warmup": WU,
        "ms_per_step": round(plain_ms, 4),
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "unbiased ReSTIR 1920x1080, terrain 136,050 tris + 2k emissive lamp triangles, "
                        "10k point lights + the 2k emissive triangles (one alias table over both), temporal cap 20M, 2x unbiased spatial k=5 "
                        "r=30 (the second Pass C's centre ray is the final sample's visibility re-test), "
                        "orbit 0.5 deg/frame; plain RIS M=32 vs online-trained MLP proposal M=8 "
                        "(114-128-128-128-256, bf16 MFMA, Adam 1e-3, 4096 samples/frame)",
            "resolution": f"{W}x{H}", "triangles": scene.n_tris,
            "point_lights": int(len(scene.point_lights)), "triangle_lights": int(len(scene.tri_lights)),
            "M_plain": args.M, "M_learned": args.learned_M, "weights": args.weights,
            "k": args.k, "spatial_passes": args.spatial_passes, "tile": args.tile,
        },
        "plain": {"ms_per_frame": round(plain_ms, 4), "frame_timing": plain_timing, "pass_a_schedule": sched_plain,
                  "kernels": rounded(kernels), "resampling_passes": resampling},
        "mrays_per_s": round(rays_per_s / 1e6, 1),
        "traversal": {"rays": nr, "ray_set": "last frame's second Pass C neighbour rays",
                      "walk": "4-entry nodes (two tree levels per visit), persistent queue tracer, triangle tests gathered per wave",
                      "node_bytes": NODE_BYTES, "ops_per_visit": VISIT_OPS, "ops_per_test": TEST_OPS,
                      "node_visits_per_ray": round(Vn, 2), "tri_tests_per_ray": round(Tn, 2),
                      "ms": round(trace_ms, 4), "roofline_mrays_per_s": round(roof_rays / 1e6, 1),
                      "valu_roofline_mrays_per_s": round(valu_rays / 1e6, 1),
                      "hbm_roofline_mrays_per_s": round(hbm_rays / 1e6, 1) if hbm_rays else None,
                      "frac": round(rays_per_s / roof_rays, 4), "pass_c_rays": ray_sets},
        "roofline": roofline,
    }

    if not args.no_learned:
        out.update(learned_and_error(torch, ctx, scene, sd, fr, args, plain_ms))

    if not args.no_views:
        out["batched_views"] = batched_views(torch, args)

    if not args.no_builds:
        out["scene_builds"] = scene_builds(torch, scene)

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fr, scene, sd, args, cpu_rays)
    print(json.dumps(out), flush=True)


def speedup_stats(rp, rl, plain_ms, learned_ms):
    """Equal-error speedup of learned over plain from per-frame relMSE arrays [view, frame]: the
    ratio of ms/frame x mean relMSE, its bootstrap 95% CI (frames resampled within each view) and
    the same ratio of medians."""
    mp, ml = rp.mean(), rl.mean()  # every view weighs the same
    speedup = (plain_ms * mp) / (learned_ms * ml)
    brng = np.random.default_rng(2024)
    boots = np.empty(BOOTSTRAP)
    n = rp.shape[1]
    for b in range(BOOTSTRAP):
        bp = np.take_along_axis(rp, brng.integers(0, n, size=rp.shape), axis=1).mean()
        bl = np.take_along_axis(rl, brng.integers(0, n, size=rl.shape), axis=1).mean()
        boots[b] = (plain_ms * bp) / (learned_ms * bl)
    lo, hi = np.quantile(boots, [0.025, 0.975])
    med_ratio = (plain_ms * np.median(rp)) / (learned_ms * np.median(rl))
    return speedup, lo, hi, med_ratio


def learned_and_error(torch, ctx, scene, sd, fr, args, plain_ms):
    """The learned pipeline's timed frames, then both pipelines' per-frame error against the FP64
    ground truth (GPU, restir_ground_truth) at fixed views of the orbit (ERROR_VIEWS), with a
    bootstrap confidence interval of the equal-error speedup. The workload's setting (M, mix)
    gives `value`; unless --no-tuned, a second proposal trained the same way with --tuned-M
    candidates and a --tuned-mix share of source draws is timed and measured against the same
    ground truth and plain frames (equal_rmse_tuned)."""
    from restir_amd.learned import LearnedProposal

    K, WU, W, H = args.steps, args.warmup, fr.W, fr.H
    n_orbit = len(fr.cams)

    def train_and_time(M, mix):
        lp = LearnedProposal(ctx, scene, k=256, seed=7, mix=mix)
        t0 = time.time()
        for i in range(args.pretrain):  # online training over the orbit (untimed)
            fr.run(i % n_orbit, M, learned=lp)
        torch.cuda.synchronize()
        log(f"learned (M={M}, mix {mix}): {args.pretrain} training frames ({time.time() - t0:.1f}s), "
            f"loss {lp.loss.item():.3f}")
        sched = fr.choose_pass_a_schedule(WU - 1, M, learned=lp)
        for t in fr.res:
            t.zero_()
        for f in range(WU):
            fr.run(f, M, learned=lp)
        ms, kern = fr.timed(WU, K, M, learned=lp)
        timing = fr.last_timing
        log(f"learned M={M}: {ms:.3f} ms/frame (eager {timing['ms_per_frame_eager']:.3f}); " + ", ".join(
            f"{k} {v['ms_per_frame']:.3f}" for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms_per_frame"])))
        for v in kern.values():
            for key in ("launches_per_frame", "ms_per_frame", "avg_ms"):
                v[key] = round(v[key], 4)
        return {"lp": lp, "M": M, "mix": mix, "ms": ms, "kernels": kern, "timing": timing, "sched": sched}

    runs = [train_and_time(args.learned_M, args.mix)]
    if not args.no_tuned:
        runs.append(train_and_time(args.tuned_M, args.tuned_mix))

    # ---- per-frame error at static cameras (fixed views of the orbit), independent frames: the
    # history is cleared before each frame and the previous G-buffer is the current one
    views = [v for v in ERROR_VIEWS if v < n_orbit]
    per_view = max(1, args.error_frames // len(views))
    rp, rls, n_pix = [], [[] for _ in runs], 0
    t_gt = 0.0
    for vi, f_err in enumerate(views):
        gi = f_err + 1
        cam = fr.cams[f_err]
        nrm = fr.G[gi][3 * W * H: 6 * W * H].view(3, -1)
        emissive = fr.G[gi][9 * W * H: 10 * W * H]  # emissive pixels show their emission, not lighting
        surf = torch.nonzero((nrm.abs().sum(0) > 0) & (emissive < 0.5)).view(-1).cpu().numpy()
        rng = np.random.default_rng(16 + f_err)
        pix = np.sort(rng.choice(surf, min(args.error_pixels, len(surf)), replace=False)).astype(np.uint32)
        d_pix = torch.from_numpy(pix).to(fr.G[0].device)
        n = len(pix)
        n_pix += n
        gt = torch.zeros(3 * n, dtype=torch.float64, device=fr.G[0].device)
        t0 = time.time()
        ctx.ground_truth(cam["position"], fr.gview(gi), d_pix, n, gt, n, n_tri_samples=TRUTH_TRI_SAMPLES)
        torch.cuda.synchronize()
        t_gt += time.time() - t0
        gt = gt.view(3, n)
        idx = d_pix.long()

        def errors(M, learned):
            if learned is not None:
                learned.bind()  # the clusters of the proposal in use (every proposal here has the same)
            rel = []
            for e in range(per_view):
                for t in fr.res:
                    t.zero_()
                fr.run(f_err, M, learned=learned, gcur=gi, gprev=gi,
                       seed_frame=100000 + 10000 * args.error_seed + 1000 * vi + e,
                       train=False)
                x = torch.nan_to_num(fr.rgb.view(3, -1)[:, idx].double())
                rel.append((((x - gt) ** 2) / (gt ** 2 + 1e-2)).mean().item())
            return np.array(rel)

        rp.append(errors(args.M, None))
        for r, rl in zip(runs, rls):
            rl.append(errors(r["M"], r["lp"]))
    torch.cuda.synchronize()
    log(f"FP64 ground truth of {n_pix} pixels over {len(views)} views: {t_gt:.2f}s")
    rp = np.stack(rp)  # [view, frame]
    mp = rp.mean()
    protocol = (f"static cameras at orbit frames {list(views)} (fixed; independent of --steps and "
                f"--warmup), {per_view} independent frames per view and method (history cleared), "
                f"FP64 GPU ground truth on up to {args.error_pixels} seeded non-emissive surface pixels "
                f"per view (triangle lights by {TRUTH_TRI_SAMPLES} stratified samples each); relMSE = "
                "mean (x - gt)^2 / (gt^2 + 1e-2) per frame; time to a target error is proportional to "
                "ms/frame x mean relMSE; 95% CI by bootstrap over frames within each view")
    studies = []
    for r, rl in zip(runs, rls):
        rl = np.stack(rl)
        ml = rl.mean()
        speedup, lo, hi, med_ratio = speedup_stats(rp, rl, plain_ms, r["ms"])
        log(f"M={r['M']}, mix {r['mix']}: relMSE plain {mp:.3f} (median {np.median(rp):.3f}), learned {ml:.3f} "
            f"(median {np.median(rl):.3f}); equal-error speedup learned vs plain {speedup:.3f} "
            f"(95% CI {lo:.3f}-{hi:.3f}); of medians {med_ratio:.3f}")
        studies.append({
            "relmse_learned": round(float(ml), 4),
            "relmse_learned_median": round(float(np.median(rl)), 4),
            "relmse_learned_by_view": [round(float(v), 4) for v in rl.mean(axis=1)],
            "relmse_learned_p90_max": [round(float(np.quantile(rl, 0.9)), 3), round(float(rl.max()), 3)],
            "ms_learned": round(r["ms"], 4),
            "speedup_learned_vs_plain": round(float(speedup), 4),
            "speedup_ci95": [round(float(lo), 4), round(float(hi), 4)],
            "speedup_of_medians": round(float(med_ratio), 4),
            "target": ">= 2x (not met)" if speedup < 2 else ">= 2x (met)",
            "value": learned_ms_value(r["ms"], ml, mp),
        })
    main, first = runs[0], studies[0]
    out = {
        "value": first.pop("value"),
        "ms_per_step": round(main["ms"], 4),
        "learned": {"ms_per_frame": round(main["ms"], 4), "frame_timing": main["timing"],
                    "pass_a_schedule": main["sched"], "kernels": main["kernels"],
                    "train_loss": round(main["lp"].loss.item(), 4),
                    "pretrain_frames": args.pretrain, "mix": args.mix, "train_every": args.train_every,
                    "train_on": args.train_on},
        "equal_rmse": {
            "protocol": protocol,
            "views": list(views), "frames_per_view": per_view, "error_seed": args.error_seed,
            "relmse_plain": round(float(mp), 4), "relmse_plain_median": round(float(np.median(rp)), 4),
            "relmse_plain_by_view": [round(float(v), 4) for v in rp.mean(axis=1)],
            "relmse_plain_p90_max": [round(float(np.quantile(rp, 0.9)), 3), round(float(rp.max()), 3)],
            "ms_plain": round(plain_ms, 4), **first,
        },
    }
    if len(runs) > 1:
        tuned, st = runs[1], studies[1]
        out["equal_rmse_tuned"] = {
            "setting": f"learned M={tuned['M']}, mix {tuned['mix']} (not the workload's M={args.learned_M}, "
                       f"mix {args.mix}, which `value` keeps); same training, ground truth, plain frames "
                       "and protocol as equal_rmse",
            "M": tuned["M"], "mix": tuned["mix"], "frame_timing": tuned["timing"],
            "learned_ms_at_plain_error": st.pop("value"), **st,
        }
        main["lp"].bind()
    return out


def learned_ms_value(ms, ml, mp):
    """The learned pipeline's ms/frame at the plain pipeline's per-frame error."""
    return round(ms * ml / mp, 4)


def scene_builds(torch, scene):
    """Device tree build (restir_build_bvh) against the host build of the same algorithm
    (restir_build_aabb_tree: the reference's sequential code, one core), on the 1080p scene and
    on a seeded 1M-triangle terrain. Device time: host wall clock around a build (it waits for
    the stream once per tree level), best of 5 after one warm build; the device tree is compared
    byte for byte with the host one. Then the device alias table (restir_build_alias_table_device)
    against the host build over the 1080p scene's 10k point lights and over 1M lognormal lights:
    device time by events on the stream, best of 5; tables compared byte for byte."""
    from restir_amd import abi, scenes
    from restir_amd.abi import Context

    big = scenes.terrain(31, grid=708, n_boxes=500, n_point_lights=10)
    ctx = Context(0)
    rows = []
    for name, sc in (("1080p scene", scene), ("terrain 708x708 + 500 boxes", big)):
        tri = np.ascontiguousarray(sc.tri_xyz, dtype=np.float32)
        n = len(tri)
        t = time.perf_counter()
        hn, ht, hd = abi.build_aabb_tree(tri)
        host_ms = (time.perf_counter() - t) * 1e3
        x = torch.from_numpy(tri).cuda()
        nodes = torch.empty(abi.NODE_DTYPE.itemsize * (n - 1), dtype=torch.uint8, device="cuda")
        tris = torch.empty(abi.TRIANGLE_DTYPE.itemsize * n, dtype=torch.uint8, device="cuda")
        ctx.build_bvh(x, n, nodes, tris)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(5):
            t = time.perf_counter()
            depth = ctx.build_bvh(x, n, nodes, tris)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t) * 1e3)
        ctx.profile_read()
        ctx.profile_enable(True)
        ctx.build_bvh(x, n, nodes, tris)
        torch.cuda.synchronize()
        prof = ctx.profile_read()
        ctx.profile_enable(False)
        same_tree = bool(np.array_equal(nodes.cpu().numpy(), hn.view(np.uint8)) and depth == hd)

        top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:6]
        rows.append({"scene": name, "triangles": n, "depth": depth,
                     "tree_device_ms": round(best, 3), "tree_host_ms": round(host_ms, 1),
                     "tree_speedup": round(host_ms / best, 1), "tree_identical": same_tree,
                     "tree_kernels_ms": {k: round(v[1], 3) for k, v in top},
                     "tree_launches": int(sum(v[0] for v in prof.values()))})
        log(f"scene build {name}: tree {best:.2f} ms on the GPU vs {host_ms:.0f} ms host "
            f"(identical {same_tree})")
        del x, nodes, tris
    rng = np.random.default_rng(17)
    big_lights = np.zeros(1_000_000, dtype=abi.POINT_LIGHT_DTYPE)
    big_lights["color_lum"][:, :3] = 1.0
    big_lights["color_lum"][:, 3] = rng.lognormal(0.0, 2.0, len(big_lights)).astype(np.float32)
    for name, pts in (("1080p point lights", scene.point_lights), ("1M lognormal lights", big_lights)):
        pts = np.ascontiguousarray(pts, dtype=abi.POINT_LIGHT_DTYPE)
        n = len(pts)
        t = time.perf_counter()
        host = abi.build_alias_table(pts, None)
        host_ms = (time.perf_counter() - t) * 1e3
        d_pts = torch.from_numpy(pts.view(np.uint8).reshape(-1).copy()).cuda()
        cols = torch.empty(16 * n, dtype=torch.uint8, device="cuda")
        ctx.build_alias_table(d_pts, n, None, 0, cols)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ctx.build_alias_table(d_pts, n, None, 0, cols)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        same = bool(np.array_equal(cols.cpu().numpy(), host.view(np.uint8)))
        rows.append({"alias_table": name, "lights": n, "alias_device_ms": round(best, 3),
                     "alias_host_ms": round(host_ms, 1), "alias_identical": same})
        log(f"alias table {name}: {best:.2f} ms on the GPU vs {host_ms:.0f} ms host (identical {same})")
        del d_pts, cols
    ctx.close()
    return rows


def batched_views(torch, args):
    """16 views (resolutions log-uniform in [640x360, 2560x1440], seeded) over 4 seeded terrains
    (about 25k-200k triangles, 1k-50k point lights), unbiased ReSTIR M = 32 with temporal reuse
    and 2 Pass C: 8 orbiting frames of all views captured in one HIP graph (shading as one
    persistent multi-view launch; Passes A and C either per view on the queued ray schedule, each
    scene's views on a stream of their own, or as persistent multi-view launches, both timed),
    against the same frames rendered view by view (eager, one launch sequence per view, queued
    Pass C, one stream)."""
    from restir_amd import scenes
    from restir_amd.pipeline import BatchRenderer, FrameSequenceGraph, Renderer

    scs, views = scenes.batched_view_set(3)
    px = sum(w * h for _, w, h in views)
    frames = 8
    seq = [[scenes.orbit_camera(sc.camera, f) for sc, _, _ in views] for f in range(frames)]
    kw = dict(mode="unbiased", k=args.k, M=args.M, spatial_passes=args.spatial_passes)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def graph_ms_per_frame(rays):
        b = BatchRenderer(views, rays=rays, **kw)
        b.render_frame(seq[0])
        torch.cuda.synchronize()
        b.reset()
        fg = FrameSequenceGraph(b, seq)
        fg.replay()  # warm
        torch.cuda.synchronize()
        reps = 3
        e0.record()
        for _ in range(reps):
            fg.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / (reps * frames)
        del fg, b
        return ms

    graph_by_rays = {r: graph_ms_per_frame(r) for r in ("queued", "persistent")}
    graph_ms = min(graph_by_rays.values())
    rs = [Renderer(sc, w, h, **kw) for sc, w, h in views]
    for r, cam in zip(rs, seq[0]):
        r.render_frame(cam)
    torch.cuda.synchronize()
    e0.record()
    for cams in seq:
        for r, cam in zip(rs, cams):
            r.render_frame(cam)
    e1.record()
    torch.cuda.synchronize()
    eager_ms = e0.elapsed_time(e1) / frames
    del rs
    torch.cuda.empty_cache()
    log(f"batched views: 16 views, {px / 1e6:.1f} Mpx/frame: graph of 8 frames {graph_ms:.2f} ms/frame "
        f"(rays queued per view {graph_by_rays['queued']:.2f}, persistent multi-view "
        f"{graph_by_rays['persistent']:.2f}), view-by-view {eager_ms:.2f} ms/frame")
    return {"views": 16, "scenes": [int(sc.n_tris) for sc in scs],
            "lights": [int(len(sc.point_lights)) for sc in scs],
            "triangle_lights": [int(len(sc.tri_lights)) for sc in scs],
            "resolutions": [f"{w}x{h}" for _, w, h in views], "pixels_per_frame": px,
            "frames_per_graph": frames, "graph_ms_per_frame": round(graph_ms, 3),
            "graph_ms_per_frame_by_ray_schedule": {k: round(v, 3) for k, v in graph_by_rays.items()},
            "view_by_view_ms_per_frame": round(eager_ms, 3),
            "graph_mpx_per_s": round(px / (graph_ms * 1e-3) / 1e6, 1),
            "speedup": round(eager_ms / graph_ms, 3)}


def cpu_baseline(fr, scene, sd, args, rays):
    """The CPU oracle on this host's cores, on a bounded sample of the same workload: the first
    `cpu_frames` frames of the plain pipeline at full 1080p (Pass A + 2x Pass C + shading over
    the same G-buffers, history chained from an empty one), and the oracle's traversal of the
    GPU's last Pass C neighbour rays (the ray set of the Mrays/s figure), repeated to >= 3 s."""
    import oracle
    from restir_amd import abi

    W, H = fr.W, fr.H
    orc = oracle.Scene(sd.nodes, sd.tris, scene.point_lights,
                       scene.tri_lights if len(scene.tri_lights) else None, sd.alias)
    prev = np.zeros(W * H, dtype=abi.RESERVOIR_DTYPE)
    gs = [oracle.gbuffer(fr.G[i].cpu().numpy(), W, H) for i in range(args.cpu_frames + 1)]
    t0 = time.perf_counter()
    for f in range(args.cpu_frames):
        u = fr.uniforms(f, args.M)
        cur = oracle.pass_a(orc, u, gs[f + 1], gs[f], prev, abi.RESERVOIR_DTYPE)
        for j in range(args.spatial_passes):
            cur = oracle.pass_c(orc, u, args.k, j, gs[f + 1], cur)
        oracle.shade(orc, fr.cams[f]["position"], 1.0, gs[f + 1], cur)
        prev = cur
    frame_s = (time.perf_counter() - t0) / args.cpu_frames
    p1, p2 = rays
    reps, t1 = 0, time.perf_counter()
    while True:
        oracle.trace_shadow(orc, p1, p2, with_counts=False)
        reps += 1
        if time.perf_counter() - t1 >= 3.0:
            break
    trace_s = (time.perf_counter() - t1) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(frame_s * 1e3, 1), "unit": "ms/frame @1080p", "cores": cores, "kind": "port",
            "sample": f"{args.cpu_frames} frames of the plain pipeline (M={args.M}, Pass A + "
                      f"{args.spatial_passes}x Pass C + shading) at {W}x{H}: {frame_s * args.cpu_frames:.1f} s; "
                      f"traversal of {len(p1)} Pass C neighbour rays x{reps}: {trace_s * reps:.1f} s",
            "mrays_per_s": round(len(p1) / trace_s / 1e6, 2),
            "c1_ms_per_frame": cpu_c1(args),
            "reference_baseline": "lavapipe baseline not run: the reference's GLSL path needs glslc and a "
                                  "Vulkan ICD (Mesa lavapipe), neither present on this machine; the "
                                  "baseline is the project's C oracle on the same cores"}


def cpu_c1(args):
    """The CPU-runnable configuration on the oracle: the 256x256 Cornell box (34 triangles, 64
    point lights), biased ReSTIR with M = 32, temporal reuse on a static camera and one Pass B with
    k = 5 in a 30 px radius, plus shading; ms per frame over 2 frames after one."""
    import oracle
    from restir_amd import abi, scenes

    sc = scenes.cornell_box(1)
    W = H = 256
    nodes, tris, _ = abi.build_aabb_tree(sc.tri_xyz)
    alias = abi.scene_alias_table(sc.point_lights, sc.tri_lights)
    orc = oracle.Scene(nodes, tris, sc.point_lights, None, alias)
    planes = oracle.primary_gbuffer(tris, abi.camera_struct(sc.camera), sc.tri_material, sc.materials, W, H)
    g = oracle.gbuffer(planes, W, H)
    pv = abi.projection_view(sc.camera, W / H)
    prev = np.zeros(W * H, dtype=abi.RESERVOIR_DTYPE)
    times = []
    for f in range(3):
        u = abi.make_uniforms(W, H, f + 1, sc.camera["position"], pv, M=32, neighbors=5, radius=30.0)
        t0 = time.perf_counter()
        a = oracle.pass_a(orc, u, g, g, prev, abi.RESERVOIR_DTYPE)
        b = oracle.pass_b(orc, u, 0, g, a)
        oracle.shade(orc, sc.camera["position"], 1.0, g, b)
        times.append(time.perf_counter() - t0)
        prev = b
    return round(float(np.mean(times[1:])) * 1e3, 2)


def stress_4k(torch, args):
    """The 4K stress workload's learned pipeline on one GPU: a 1024x1024-vertex heightfield plus
    20,000 boxes (2,333,058 triangles, tree built on the device), 1M point lights with log-uniform
    intensity [1e-3, 10] clustered into 4,096 groups (nearest-centroid step on the GPU), 3840x2160
    frames of the orbiting camera with the 4,096-way MLP proposal (M = 8; inference, Pass A, two
    Pass C, shading and one 4096-sample Adam step per frame) and, for comparison, plain RIS with
    M = 32; then 4 camera paths x 32 learned frames in one HIP graph (stress_paths_graph). Not
    part of it here: fp16 G-buffer normals. Prints one JSON line with the build times, ms/frame
    of both pipelines, the per-kernel times and the graph's ms/frame."""
    from restir_amd import abi, scenes
    from restir_amd.abi import Context
    from restir_amd.learned import LearnedProposal
    from restir_amd.pipeline import SceneOnDevice

    W, H, K, WU = 3840, 2160, args.steps, args.warmup
    npx = W * H
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    tx, ty = (int(v) for v in args.tile.split("x"))
    launch = abi.Launch(tx, ty)
    t0 = time.time()
    scene = scenes.stress_scene()
    log(f"stress scene: {scene.n_tris} triangles, {len(scene.point_lights)} point lights ({time.time() - t0:.1f}s)")
    ctx = Context(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sd = SceneOnDevice(ctx, scene, dev, build="device")
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"device tree (depth {sd.depth}) + device alias table over 1M lights: {setup_s:.2f} s")
    t0 = time.perf_counter()
    lp = LearnedProposal(ctx, scene, k=4096, seed=7, iterations=16, device_clustering=True)
    torch.cuda.synchronize()
    cluster_s = time.perf_counter() - t0
    log(f"4,096 clusters over 1M lights (16 Lloyd steps, nearest step on the GPU) + MLP init: {cluster_s:.2f} s")
    nframes = WU + K
    cams = [scenes.orbit_camera(scene.camera, f) for f in range(nframes)]
    G = [torch.zeros(abi.GBUFFER_PLANES * npx, dtype=torch.float32, device=dev) for _ in range(nframes + 1)]
    for f in range(nframes):  # the workload's G-buffer keeps its normals as half planes
        ctx.primary_gbuffer(cams[f], sd.d_tri_material, sd.d_materials, len(scene.materials), G[f + 1],
                            W, H, stream=stream, normal_format=abi.NORMAL_F16)
    torch.cuda.synchronize()
    fr = Frames(torch, ctx, scene, cams, G, W, H, args, launch, stream)
    fr.normal_format = abi.NORMAL_F16
    out = {"workload": "stress4k", "normals": "fp16", "width": W, "height": H, "triangles": int(scene.n_tris),
           "point_lights": int(len(scene.point_lights)), "clusters": 4096, "steps": K, "warmup": WU,
           "tree_and_alias_build_s": round(setup_s, 3), "clustering_s": round(cluster_s, 3)}
    for name, M, learned in (("plain", args.M, None), ("learned", args.learned_M, lp)):
        for t in fr.res:
            t.zero_()
        for f in range(WU):
            fr.run(f, M, learned=learned)
        ms, kernels = fr.timed(WU, K, M, learned=learned)
        log(f"{name} M={M}: {ms:.2f} ms/frame (eager {fr.last_timing['ms_per_frame_eager']:.2f}); " + ", ".join(
            f"{k} {v['ms_per_frame']:.3f}" for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_frame"])[:8]))
        out[f"{name}_M"] = M
        out[f"{name}_ms_per_frame"] = round(ms, 3)
        out[f"{name}_frame_timing"] = fr.last_timing
        out[f"{name}_kernels_ms_per_frame"] = {k: round(v["ms_per_frame"], 4) for k, v in kernels.items()}
    if args.stress_paths > 0:
        out.update(stress_paths_graph(torch, args, scene, ctx, lp, launch, W, H))
    print(json.dumps(out), flush=True)


def stress_paths_graph(torch, args, scene, ctx, lp, launch, W, H):
    """The stress workload's batching: camera paths of 32 frames each (G-buffer, learned
    proposal, Pass A, two Pass C, shading and one Adam step per frame), one renderer per path on
    one context and one proposal, all captured into one HIP graph; timed against the same
    sequence run eagerly from the same starting state."""
    from restir_amd import abi, scenes
    from restir_amd.pipeline import Renderer
    paths, plen = args.stress_paths, 32
    cams = [[scenes.orbit_camera(scene.camera, 180 * p + f) for f in range(plen)] for p in range(paths)]
    rs = [Renderer(scene, W, H, mode="unbiased", k=args.k, M=args.learned_M, spatial_passes=args.spatial_passes,
                   ctx=ctx, learned=lp, launch=launch, scene_build="device", normal_format=abi.NORMAL_F16)
          for _ in range(paths)]
    for r, c in zip(rs, cams):
        r.render_frame(c[0])  # sizes the library's scratch outside the capture
    torch.cuda.synchronize()
    start = lp.state()

    def restart():
        for r in rs:
            r.reset()
        lp.load_state(start)
        torch.cuda.synchronize()

    restart()
    t0 = time.perf_counter()
    for r, c in zip(rs, cams):
        for cam in c:
            r.render_frame(cam)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / (paths * plen)
    restart()
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    stream.wait_stream(torch.cuda.current_stream())
    t0 = time.perf_counter()
    with torch.cuda.graph(graph, stream=stream):
        for r, c in zip(rs, cams):
            for cam in c:
                r.render_frame(cam, stream=stream)
    capture_s = time.perf_counter() - t0
    restart()
    t0 = time.perf_counter()
    graph.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / (paths * plen)
    log(f"{paths} paths x {plen} frames (learned, G-buffers rendered in-frame): graph {graph_ms:.2f} ms/frame, "
        f"eager {eager_ms:.2f} ms/frame, capture {capture_s:.1f} s")
    return {"graph_paths": paths, "graph_frames_per_path": plen, "graph_ms_per_frame": round(graph_ms, 3),
            "eager_ms_per_frame_same_sequence": round(eager_ms, 3), "graph_capture_s": round(capture_s, 2)}


if __name__ == "__main__":
    main()
