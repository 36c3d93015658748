"""Workgroup-tile autotuner for the per-pixel passes.

Every tile shape compiled into the library (8x8, 16x8, 16x16, 32x8, 64x1: one, two or four
wave64s per workgroup) computes bit-identical reservoirs (each pixel's work and random stream are
independent of the launch shape), so the choice is purely a matter of speed: cache locality of
the neighbour gathers and traversal divergence inside a workgroup. The tuner times each pass of
a renderer's current frame with every tile through the library's per-kernel event timing and
returns the fastest shape per pass together with the measured table, so the choice is recorded.
"""
from __future__ import annotations

from . import abi

TILES = ((8, 8), (16, 8), (16, 16), (32, 8), (64, 1))
# kernels whose time depends on the tile shape, per pass
_PASS_KERNELS = {
    "a": ("k_pass_a", "k_pass_a_ris", "k_trace_queue(pass_a)", "k_pass_a_finish"),
    "b": ("k_pass_b",),
    "c": ("k_pack_surface", "k_pass_c_merge",
          "k_trace_queue(pass_c)", "k_pass_c_finish", "k_pass_c"),
}


def _time_ms(ctx, run, names, reps):
    torch = __import__("torch")
    run()  # warm (scratch growth, caches)
    torch.cuda.synchronize()
    ctx.profile_read()
    ctx.profile_enable(True)
    try:
        for _ in range(reps):
            run()
        rows = ctx.profile_read()
    finally:
        ctx.profile_enable(False)
    return sum(rows[n][1] for n in names if n in rows) / reps


def autotune(r, reps: int = 5, tiles=TILES, schedule: int = abi.SCHEDULE_QUEUED) -> dict:
    """Times the passes of renderer ``r``'s current frame (call after at least one render_frame)
    for each tile; sets ``r.launches`` to the fastest shapes and returns
    {"a"|"b"|"c": {"choice": "TXxTY", "ms": {"TXxTY": ms per call}}} for the passes ``r`` runs."""
    torch = __import__("torch")
    u, _ = r.uniforms(r.cam)
    i, p = r.cur, 1 - r.cur
    g, gp = r.gstruct(i), r.gstruct(p)
    scratch = torch.zeros_like(r.tmp)
    passes = ["a"] + (["c"] if r.mode == "unbiased" else ["b"])
    out = {}
    for ps in passes:
        table = {}
        for tx, ty in tiles:
            la = abi.Launch(tx, ty, schedule)
            if ps == "a":
                def run(la=la):
                    r.ctx.initial_temporal(u, g, gp, r.res[p], scratch, la)
            elif ps == "b":
                def run(la=la):
                    r.ctx.spatial(u, 0, g, r.res[i], scratch, la)
            else:
                def run(la=la):
                    r.ctx.unbiased(u, r.k, 0, g, r.res[i], scratch, la)
            table[f"{tx}x{ty}"] = _time_ms(r.ctx, run, _PASS_KERNELS[ps], reps)
        best = min(table, key=table.get)
        tx, ty = (int(v) for v in best.split("x"))
        r.launches[ps] = abi.Launch(tx, ty, schedule)
        out[ps] = {"choice": best, "ms": {k: round(v, 4) for k, v in table.items()}}
    return out
