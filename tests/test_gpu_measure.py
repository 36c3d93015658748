"""GPU checks of the measurement entry points: per-kernel timing (restir_profile_*) and the ray
export of the queued Pass C (restir_unbiased_rays), whose rays the benchmark counts and re-traces."""
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


def test_profile_records_each_kernel(torch):
    r = Renderer(scenes.cornell_box(1), 64, 48, mode="unbiased", k=3, M=8, spatial_passes=2)
    r.render_frame(r.scene.camera)
    torch.cuda.synchronize()
    r.ctx.profile_enable(True)
    r.render_frame(r.scene.camera)
    rows = r.ctx.profile_read()
    r.ctx.profile_enable(False)
    # queued schedule (the default): Pass A is RIS, its ray queue, and the finish
    assert rows["k_pass_a_ris"][0] == 1 and rows["k_pass_a_finish"][0] == 1
    assert rows["k_trace_queue(pass_a)"][0] == 1
    assert rows["k_pass_c_merge"][0] == 2 and rows["k_trace_queue(pass_c)"][0] == 2
    # the second, flagged Pass C reuses the surface records the first packed from the same planes
    assert rows["k_pack_surface"][0] == 1
    assert rows["k_shade"][0] == 1
    for name, (n, total, mx) in rows.items():
        assert n >= 1 and total > 0.0 and 0.0 < mx <= total + 1e-6, (name, n, total, mx)
    assert r.ctx.profile_read() == {}  # reading clears
    r.render_frame(r.scene.camera)      # disabled: nothing recorded
    assert r.ctx.profile_read() == {}


def test_unbiased_rays_are_the_queued_rays(torch):
    sc = scenes.cornell_box(1)
    W, H, k = 64, 48, 3
    r = Renderer(sc, W, H, mode="unbiased", k=k, M=8, spatial_passes=1)
    r.render_frame(sc.camera)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(r.cur)
    out = torch.zeros_like(r.tmp)
    r.ctx.unbiased(u, k, 0, g, r.res[r.cur], out)
    npx = W * H
    ref_out = out.clone()
    r.ctx.set_ray_export(True)  # kept for ABI 14 callers; it changes nothing
    r.ctx.unbiased(u, k, 0, g, r.res[r.cur], out)
    assert torch.equal(out, ref_out)
    p1 = torch.zeros(npx * k, 4, dtype=torch.float32, device="cuda")
    p2 = torch.zeros_like(p1)
    nc = r.ctx.unbiased_rays(0, p1, p2, npx * k)
    planes = r.gbuf[r.cur].cpu().numpy()
    pos = planes[: 3 * npx].reshape(3, npx).T
    nrm = planes[3 * npx: 6 * npx].reshape(3, npx).T
    surface = np.flatnonzero(np.abs(nrm).sum(1) > 0)
    # one centre ray per surface pixel, from its position to the merged sample's light position
    assert nc == len(surface)
    c1, c2 = p1[:nc].cpu().numpy(), p2[:nc].cpu().numpy()
    tag = c1[:, 3].view(np.uint32)
    assert np.array_equal(np.sort(tag), surface)
    assert np.array_equal(c1[:, :3], pos[tag])
    res = out.cpu().numpy().view(abi.RESERVOIR_DTYPE)
    assert np.array_equal(c2[:, :3], res["pos_lum"][tag, :3])
    # the centre rays' hit/miss through restir_trace_shadow matches the oracle's traversal
    sh = torch.zeros(nc, dtype=torch.int32, device="cuda")
    r.ctx.trace_shadow(p1[:nc].contiguous(), p2[:nc].contiguous(), nc, sh)
    hs = HostScene(sc)
    ref = oracle.trace_shadow(hs.oracle, c1.copy(), c2.copy(), with_counts=False)[0]
    assert np.array_equal(sh.cpu().numpy().astype(bool), ref.astype(bool))
    # neighbour rays: at most k per pixel, tagged with a pixel and a slot < k
    nn = r.ctx.unbiased_rays(1, p1, p2, npx * k)
    assert 0 < nn <= npx * k
    t = p1[:nn, 3].cpu().numpy().view(np.uint32)
    assert (t & 0x07ffffff).max() < npx and (t >> 27).max() < k
    # ... exactly the traced ones: rays of pixels whose centre ray found the light, each
    # (pixel, slot) once, towards the pixel's merged light, from a neighbour pixel's position (a
    # background neighbour has a zero normal, which the above-horizon test lets through)
    pix, slot = t & 0x07ffffff, t >> 27
    lit = tag[ref == 0]
    assert np.isin(pix, lit).all() and len(np.unique(pix)) > 0.5 * len(lit)
    assert len(np.unique(pix.astype(np.int64) * 16 + slot)) == nn
    n1, n2 = p1[:nn].cpu().numpy(), p2[:nn].cpu().numpy()
    assert np.array_equal(n2[:, :3], res["pos_lum"][pix, :3])
    pixel_pos = {tuple(v) for v in pos}
    assert all(tuple(v) in pixel_pos for v in n1[:, :3])
    # capacity smaller than the queue: copies a prefix, still reports the full length
    q1 = torch.zeros(5, 4, dtype=torch.float32, device="cuda")
    q2 = torch.zeros_like(q1)
    assert r.ctx.unbiased_rays(1, q1, q2, 5) == nn
    assert torch.equal(q1, p1[:5])


def test_unbiased_rays_needs_a_queued_pass(torch):
    r = Renderer(scenes.cornell_box(1), 32, 32, mode="unbiased", k=3, M=4, spatial_passes=1)
    p = torch.zeros(16, 4, dtype=torch.float32, device="cuda")
    with pytest.raises(abi.RestirError):
        r.ctx.unbiased_rays(0, p, p, 16)
    with pytest.raises(abi.RestirError):
        r.ctx.unbiased_rays(2, p, p, 16)
