"""GPU FP64 ground truth (restir_ground_truth) against the oracle's FP64 ground truth on the same
G-buffer: same integrand, same FP64 visibility; only the summation order over lights differs
(light chunks, then chunks in order), so the two agree to a few units of 1e-16 relative."""
import dataclasses

import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.pipeline import Renderer
from rt_helpers import HostScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def with_ceiling_lights(sc):
    """The Cornell box plus its two ceiling triangles as emissive triangle lights."""
    tri = sc.tri_xyz.reshape(-1, 3, 3)
    ceiling = np.flatnonzero(np.all(np.abs(tri[:, :, 1] - tri[:, :, 1].max()) < 1e-6, axis=1))[:2]
    assert len(ceiling) == 2
    # lower them a little so they light the room from inside
    t = tri[ceiling].copy()
    t[:, :, 1] -= 0.01
    tl = scenes.make_tri_lights(t.reshape(-1, 9), np.array([[2.0, 1.8, 1.5], [2.0, 1.8, 1.5]]))
    return dataclasses.replace(sc, tri_lights=tl)


@pytest.mark.parametrize("case", ["cornell", "cornell_tri_lights", "terrain_subset"])
def test_ground_truth_matches_oracle(torch, case):
    if case == "terrain_subset":
        sc = scenes.terrain(11, grid=48, n_boxes=60, n_point_lights=2000, extent=16.0)
    else:
        sc = scenes.cornell_box(1)
        if case == "cornell_tri_lights":
            sc = with_ceiling_lights(sc)
    W, H = 64, 48
    r = Renderer(sc, W, H)
    r.render_gbuffer(sc.camera, 0)
    planes = r.gbuf[0].cpu().numpy()
    g = oracle.gbuffer(planes, W, H)
    hs = HostScene(sc)
    rng = np.random.default_rng(3)
    pix = np.arange(W * H) if case != "terrain_subset" else np.sort(rng.choice(W * H, 700, replace=False))
    mask = np.zeros(W * H, np.uint8)
    mask[pix] = 1
    ref = oracle.ground_truth_f64(hs.oracle, sc.camera["position"], g, use_bvh=True, n_tri_samples=64,
                                  mask=mask).reshape(3, -1)[:, pix]
    n = len(pix)
    d_pix = torch.from_numpy(pix.astype(np.uint32)).cuda()
    out = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    r.ctx.ground_truth(sc.camera["position"], r.gstruct(0), d_pix, n, out, n, n_tri_samples=64)
    got = out.view(3, n).cpu().numpy()
    assert (ref > 0).mean() > 0.3
    scale = np.maximum(np.abs(ref), 1e-300)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300), np.abs(got - ref).max()


def test_ground_truth_rejects_bad_pixels(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 16, 16)
    r.render_gbuffer(sc.camera, 0)
    bad = torch.tensor([0, 256], dtype=torch.int32, device="cuda")
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    with pytest.raises(abi.RestirError):
        r.ctx.ground_truth(sc.camera["position"], r.gstruct(0), bad, 2, out, 2)
