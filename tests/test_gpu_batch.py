"""GPU checks of batching: a HIP-graph-captured 8-frame sequence against the same frames run
eagerly, the tile autotuner (its choice is recorded and does not change results), and the
persistent multi-view passes against per-view launches (bit-identical)."""
import numpy as np
import pytest

from restir_amd import abi, scenes
from restir_amd.pipeline import BatchRenderer, FrameSequenceGraph, Renderer
from restir_amd.tuning import autotune

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def bits(t):
    """A device tensor's words as int32 (bitwise comparison, NaN payloads included)."""
    import torch as T
    return t.contiguous().view(-1).view(T.int32)


@pytest.mark.parametrize("mode", ["unbiased", "biased"])
def test_graph_of_8_frames_matches_eager(torch, mode):
    sc = scenes.cornell_box(1)
    W, H = 64, 48
    cams = [scenes.orbit_camera(sc.camera, f) for f in range(8)]
    kw = dict(mode=mode, k=3, M=8, spatial_passes=2)
    eager = Renderer(sc, W, H, **kw)
    imgs, res = [], []
    for cam in cams:
        eager.render_frame(cam)
        imgs.append(eager.rgb.clone())
        res.append(eager.res[eager.cur].clone())
    torch.cuda.synchronize()

    r = Renderer(sc, W, H, **kw)
    r.render_frame(cams[0])  # allocates the library's scratch outside the capture
    torch.cuda.synchronize()
    r.reset()
    torch.cuda.synchronize()
    fg = FrameSequenceGraph(r, cams)
    assert r.frame == 8
    for rep in range(3):
        # state before the sequence: no temporal history
        for t in (*r.gbuf, *r.res, r.tmp):
            t.zero_()
        fg.replay()
        torch.cuda.synchronize()
        for f in range(8):
            assert torch.equal(bits(fg.images[f]), bits(imgs[f])), (rep, f)
        assert torch.equal(bits(r.res[r.cur]), bits(res[-1])), rep
        # eager work between replays (another renderer's frame on the default stream)
        eager.render_frame(cams[rep])
        torch.cuda.synchronize()


def test_graph_rejects_learned_renderer_with_host_step_count(torch):
    from restir_amd.learned import LearnedProposal
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 32, 32)
    r.learned = LearnedProposal(r.ctx, sc, k=64, seed=1)
    r.learned.mlp.adam_step_device = None
    with pytest.raises(ValueError):
        FrameSequenceGraph(r, [sc.camera])


def test_graph_of_learned_frames_matches_eager(torch):
    """Proposal, passes, shading and one Adam step per frame, 6 frames in one graph: each replay
    from the same starting state (parameters, moments, step count, reservoirs) gives the eager
    sequence's images and trained parameters bit for bit."""
    from restir_amd.learned import LearnedProposal
    sc = scenes.terrain(11, grid=32, n_boxes=20, n_point_lights=500, extent=8.0)
    W, H = 64, 48
    cams = [scenes.orbit_camera(sc.camera, f) for f in range(6)]
    kw = dict(mode="unbiased", k=3, M=8, spatial_passes=2)

    def renderer():
        r = Renderer(sc, W, H, **kw)
        r.learned = LearnedProposal(r.ctx, sc, k=64, seed=3)
        r.train_samples = 1024
        return r

    eager = renderer()
    start = eager.learned.state()
    imgs = []
    for cam in cams:
        eager.render_frame(cam)
        imgs.append(eager.rgb.clone())
    torch.cuda.synchronize()
    end = eager.learned.state()
    assert end["adam_step"] == 6 and not np.array_equal(end["w"], start["w"])

    r = renderer()
    r.render_frame(cams[0])  # sizes the library's scratch outside the capture
    torch.cuda.synchronize()
    r.reset()
    r.learned.load_state(start)
    torch.cuda.synchronize()
    fg = FrameSequenceGraph(r, cams)
    for rep in range(2):
        for t in (*r.gbuf, *r.res, r.tmp):
            t.zero_()
        r.learned.load_state(start)
        torch.cuda.synchronize()
        fg.replay()
        torch.cuda.synchronize()
        for f in range(len(cams)):
            assert torch.equal(bits(fg.images[f]), bits(imgs[f])), (rep, f)
        got = r.learned.state()
        for key in ("w", "b", "m", "v", "w_bf16"):
            assert np.array_equal(got[key].view(np.uint8), end[key].view(np.uint8)), (rep, key)
        assert got["adam_step"] == 6
        eager.render_frame(cams[rep])  # eager work between replays
        torch.cuda.synchronize()


def test_autotune_records_choice_and_keeps_results(torch):
    sc = scenes.cornell_box(1)
    W, H = 96, 64
    ref = Renderer(sc, W, H, mode="unbiased", k=3, M=8, spatial_passes=2)
    for f in range(3):
        ref.render_frame(scenes.orbit_camera(sc.camera, f))
    r = Renderer(sc, W, H, mode="unbiased", k=3, M=8, spatial_passes=2)
    r.render_frame(scenes.orbit_camera(sc.camera, 0))
    table = autotune(r, reps=2)
    assert set(table) == {"a", "c"}
    for ps, rec in table.items():
        assert rec["choice"] in rec["ms"] and len(rec["ms"]) == 5
        assert all(v > 0 for v in rec["ms"].values())
        la = r.launches[ps]
        assert f"{la.tile_x}x{la.tile_y}" == rec["choice"]
    for f in (1, 2):
        r.render_frame(scenes.orbit_camera(sc.camera, f))
    torch.cuda.synchronize()
    assert torch.equal(bits(r.rgb), bits(ref.rgb))
    assert torch.equal(bits(r.res[r.cur]), bits(ref.res[ref.cur]))


def batch_views():
    """Three views over two scenes at different sizes (one of them not a multiple of any tile)."""
    a = scenes.cornell_box(1)
    b = scenes.terrain(11, grid=48, n_boxes=60, n_point_lights=2000, extent=16.0)
    return [(a, 64, 48), (b, 80, 45), (a, 33, 70)]


@pytest.mark.parametrize("mode,rays,concurrent", [("unbiased", "queued", True), ("unbiased", "queued", False),
                                                  ("unbiased", "persistent", True), ("biased", "queued", True),
                                                  ("biased", "persistent", True)])
def test_batched_views_match_single_view_renderers(torch, mode, rays, concurrent):
    """Each view of a batch equals a Renderer of that view alone; on the queued schedule the two
    scenes' views run on streams of their own (concurrent) or one after another on one stream."""
    views = batch_views()
    kw = dict(mode=mode, k=3, M=8, spatial_passes=2)
    batch = BatchRenderer(views, rays=rays, concurrent_scenes=concurrent, **kw)
    assert batch.concurrent == (concurrent and rays == "queued")
    singles = [Renderer(sc, w, h, launch=abi.Launch(16, 16, abi.SCHEDULE_INLINE), **kw) for sc, w, h in views]
    for f in range(3):
        cams = [scenes.orbit_camera(sc.camera, f) for sc, _, _ in views]
        batch.render_frame(cams)
        for r, cam in zip(singles, cams):
            r.render_frame(cam)
        torch.cuda.synchronize()
        for v, r in enumerate(singles):
            assert torch.equal(bits(batch.images[v]), bits(r.rgb)), (f, v)
            assert torch.equal(bits(batch.views[v]["res"][batch.cur]), bits(r.res[r.cur])), (f, v)


@pytest.mark.parametrize("rays", ["queued", "persistent"])
def test_batched_views_in_a_graph_of_8_frames(torch, rays):
    views = batch_views()
    kw = dict(mode="unbiased", k=3, M=8, spatial_passes=2, rays=rays)
    seq = [[scenes.orbit_camera(sc.camera, f) for sc, _, _ in views] for f in range(8)]
    eager = BatchRenderer(views, **kw)
    imgs = []
    for cams in seq:
        eager.render_frame(cams)
        imgs.append(eager.rgb.clone())
    torch.cuda.synchronize()
    b = BatchRenderer(views, **kw)
    b.render_frame(seq[0])  # allocates the view table outside the capture
    torch.cuda.synchronize()
    b.reset()
    torch.cuda.synchronize()
    fg = FrameSequenceGraph(b, seq)
    for rep in range(3):
        for v in b.views:
            for t in (*v["gbuf"], *v["res"], v["tmp"]):
                t.zero_()
        fg.replay()
        torch.cuda.synchronize()
        for f in range(8):
            assert torch.equal(bits(fg.images[f]), bits(imgs[f])), (rep, f)
        eager.render_frame(seq[rep])
        torch.cuda.synchronize()


def test_batched_views_reject_overlapping_outputs(torch):
    views = batch_views()[:2]
    b = BatchRenderer(views, mode="unbiased", k=3, M=8)
    b.render_frame([sc.camera for sc, _, _ in views])
    v0, v1 = b.views
    u = abi.make_uniforms(v0["W"], v0["H"], 1, views[0][0].camera["position"],
                          abi.projection_view(views[0][0].camera, v0["W"] / v0["H"]), M=8)
    g0 = b._g(v0, 0)
    import ctypes as C
    rec = [abi.View(v0["ctx"].h.value, C.pointer(u), C.pointer(g0), C.pointer(g0), abi.as_ptr(v0["res"][0]),
                    abi.as_ptr(v0["res"][1]), None, 0)] * 2  # both views write the same buffer
    with pytest.raises(abi.RestirError):
        b.ctx.initial_temporal_views(rec)
