"""restir_initial_temporal_noalloc (Pass A whose queued schedule uses the library's static device
scratch array instead of allocating) against restir_initial_temporal on the same inputs: bit for
bit, on two calls in a row, the second with a smaller frame (the array keeps the first call's
contents), on the queued and inline schedules; frames above RESTIR_NOALLOC_MAX_PIXELS are
rejected before any launch."""
import numpy as np
import pytest

from restir_amd import abi, scenes
from restir_amd.pipeline import Renderer
from rt_helpers import reservoir_mismatch

pytestmark = pytest.mark.gpu

MAX_PIXELS = 2287792  # RESTIR_NOALLOC_MAX_PIXELS


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def _frame_two_ways(r, cam, launch):
    """One frame's Pass A through both entry points; returns both outputs (host copies)."""
    r.begin_frame(cam)
    u, pv_now = r.uniforms(cam)
    i, p = r.cur, 1 - r.cur
    g, gp = r.gstruct(i), r.gstruct(p)
    print(f"[noalloc] {r.W}x{r.H} frame {r.frame} schedule {launch.schedule}", flush=True)
    r.ctx.initial_temporal(u, g, gp, r.res[p], r.tmp, launch)
    ref = r.tmp.cpu().numpy().view(abi.RESERVOIR_DTYPE).copy()
    r.ctx.initial_temporal_noalloc(u, g, gp, r.res[p], r.res[i], launch)
    got = r.res[i].cpu().numpy().view(abi.RESERVOIR_DTYPE).copy()
    r.prev_pv = pv_now
    return got, ref


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_QUEUED, abi.SCHEDULE_INLINE])
def test_noalloc_pass_a_equals_allocating_pass_a(torch, schedule):
    sc = scenes.config_scene("terrain")
    ctx = abi.Context(0)
    launch = abi.Launch(16, 16, schedule)
    for W, H in ((640, 360), (320, 180)):  # the second call is smaller
        r = Renderer(sc, W, H, mode="unbiased", k=5, spatial_passes=1, ctx=ctx)
        for f in range(2):  # the second frame reuses the first one's reservoirs (temporal)
            got, ref = _frame_two_ways(r, scenes.orbit_camera(sc.camera, f), launch)
            bad = reservoir_mismatch(got, ref)
            assert not bad.any(), f"{W}x{H} frame {f}: {int(bad.sum())} reservoirs differ"
            assert (ref["sumWeights"] > 0).mean() > 0.1
        assert (ref["M"] > r.M).mean() > 0.3


def test_noalloc_rejects_frames_above_the_static_array(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 16, 16)
    r.render_gbuffer(sc.camera, 0)
    g = r.gstruct(0)
    for W, H in ((MAX_PIXELS + 1, 1), (2560, 1440)):
        u = abi.make_uniforms(W, H, 1, sc.camera["position"])
        with pytest.raises(abi.RestirError) as e:
            r.ctx.initial_temporal_noalloc(u, g, g, r.res[1], r.res[0], abi.Launch(16, 16, abi.SCHEDULE_QUEUED))
        assert e.value.code == -1 and "static array" in str(e.value)  # RESTIR_E_INVALID_ARGUMENT
    # the largest frame it takes is accepted up to the G-buffer check (this G-buffer is 16x16)
    u = abi.make_uniforms(MAX_PIXELS, 1, 1, sc.camera["position"])
    with pytest.raises(abi.RestirError) as e:
        r.ctx.initial_temporal_noalloc(u, g, g, r.res[1], r.res[0], abi.Launch(16, 16, abi.SCHEDULE_QUEUED))
    assert "static array" not in str(e.value)
