"""GPU checks of batching: a HIP-graph-captured 8-frame sequence against the same frames run
eagerly, and the tile autotuner (its choice is recorded and does not change results)."""
import pytest

from restir_amd import scenes
from restir_amd.pipeline import FrameSequenceGraph, Renderer
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


def test_graph_rejects_learned_renderer(torch):
    from restir_amd.learned import LearnedProposal
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 32, 32)
    r.learned = LearnedProposal(r.ctx, sc, k=64, seed=1)
    with pytest.raises(ValueError):
        FrameSequenceGraph(r, [sc.camera])


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
