"""Does the order of the shadow-ray queue matter to the tracer? (A measurement, not part of the
library.)

Renders 6 frames of the terrain scene at 1920x1080 (unbiased ReSTIR, k = 5, two spatial passes),
copies the last Pass C's queued centre and neighbour rays (restir_unbiased_rays), and times
restir_trace_shadow (the persistent queue tracer) over each set in queue order and after a stable
host sort by several keys (the sort itself is not timed). Keys quantise the ray's origin and end
point over the rays' bounding box: a Morton code interleaving all six coordinates (3-5 bits per
axis), the end point or the origin alone (8 bits per axis), end-major / origin-minor, and the
direction's octant before the origin. Every order must give the same verdicts.

    python restir-vulkan_amd/probe/ray_order.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
for p in (ROOT, ROOT / "restir-vulkan_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import torch  # noqa: E402

from restir_amd import scenes  # noqa: E402
from restir_amd.pipeline import Renderer  # noqa: E402


def interleave(cols, bits):
    key = np.zeros(len(cols[0]), np.int64)
    for bit in range(bits - 1, -1, -1):
        for c in cols:
            key = (key << 1) | ((c >> bit) & 1)
    return key


def main():
    assert torch.cuda.is_available(), "needs a HIP device"
    sc = scenes.config_scene("terrain")
    W, H = 1920, 1080
    r = Renderer(sc, W, H, mode="unbiased", k=5, spatial_passes=2)
    for f in range(6):
        r.render_frame(scenes.orbit_camera(sc.camera, f))
    torch.cuda.synchronize()
    cap = W * H * 5
    q1 = torch.zeros(cap, 4, device="cuda")
    q2 = torch.zeros_like(q1)
    for which, label in ((0, "centre"), (1, "neighbour")):
        n = r.ctx.unbiased_rays(which, q1, q2, cap)
        p1, p2 = q1[:n].clone(), q2[:n].clone()
        sh = torch.zeros(n, dtype=torch.int32, device="cuda")

        def timed(a, b, reps=20):
            r.ctx.trace_shadow(a, b, n, sh)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                r.ctx.trace_shadow(a, b, n, sh)
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        base = timed(p1, p2)
        ref = sh.clone()
        a = p1[:, :3].cpu().numpy().astype(np.float64)
        b = p2[:, :3].cpu().numpy().astype(np.float64)
        lo = np.minimum(a.min(0), b.min(0))
        hi = np.maximum(a.max(0), b.max(0))

        def quant(x, bits):
            return np.clip(((x - lo) / (hi - lo) * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)

        print(f"{label} rays: {n}; queue order {base:.3f} ms", flush=True)

        def run(name, key):
            perm = torch.from_numpy(np.argsort(key, kind="stable")).cuda()
            ms = timed(p1[perm].contiguous(), p2[perm].contiguous())
            same = bool(torch.equal(sh, ref[perm]))
            print(f"  {name:34s} {ms:.3f} ms  {ms / base:.3f} x queue order  verdicts equal: {same}", flush=True)
            assert same

        for bits in (3, 4, 5):
            qa, qb = quant(a, bits), quant(b, bits)
            run(f"Morton(origin, end) {bits} bits/axis",
                interleave([qa[:, 0], qa[:, 1], qa[:, 2], qb[:, 0], qb[:, 1], qb[:, 2]], bits))
        qa, qb = quant(a, 8), quant(b, 8)
        run("Morton(end) 8 bits/axis", interleave([qb[:, 0], qb[:, 1], qb[:, 2]], 8))
        run("Morton(origin) 8 bits/axis", interleave([qa[:, 0], qa[:, 1], qa[:, 2]], 8))
        qa4, qb4 = quant(a, 4), quant(b, 4)
        run("Morton(end) 4, then Morton(origin) 4",
            (interleave([qb4[:, 0], qb4[:, 1], qb4[:, 2]], 4) << 12) | interleave([qa4[:, 0], qa4[:, 1], qa4[:, 2]], 4))
        d = b - a
        octant = ((d[:, 0] > 0).astype(np.int64) << 2) | ((d[:, 1] > 0).astype(np.int64) << 1) | (d[:, 2] > 0)
        qa7 = quant(a, 7)
        run("octant of d, then Morton(origin) 7", (octant << 21) | interleave([qa7[:, 0], qa7[:, 1], qa7[:, 2]], 7))


if __name__ == "__main__":
    main()
