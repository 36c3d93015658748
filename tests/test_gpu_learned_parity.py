"""GPU parity of the learned proposal's remaining pieces against the oracle:

- one Adam step (restir_mlp_adam, the step restir_mlp_train ends with) against oracle.adam in
  FP64, from a fixed state, at K = 256 and 4,096, at steps t = 1 and 1,000, with the step count on
  the host and on the device;
- Pass A over the proposal's candidates (restir_initial_temporal_candidates) against
  oracle.pass_a_candidates, bit for bit, on the 1080p workload's scene (point and triangle lights
  in one table) at 480x270 over three frames of the orbiting camera, both ray schedules;
- binding new lights unbinds the light clusters (the proposal refuses to run on stale tables).
"""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.abi import MlpStruct, mlp_sizes
from restir_amd.learned import LearnedProposal
from restir_amd.pipeline import Renderer
from rt_helpers import reservoir_mismatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("K", [256, 4096])
@pytest.mark.parametrize("t0", [0, 999])
@pytest.mark.parametrize("device_step", [False, True])
def test_adam_step_matches_oracle(torch, K, t0, device_step):
    ctx = abi.Context(0)
    nW, nB = mlp_sizes(K)
    w_bf = torch.zeros(nW, dtype=torch.int16, device="cuda")
    w = torch.zeros(nW, dtype=torch.float32, device="cuda")
    b = torch.zeros(nB, dtype=torch.float32, device="cuda")
    m = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
    v = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    mlp = MlpStruct(K, 0, w_bf.data_ptr(), w.data_ptr(), b.data_ptr(), m.data_ptr(), v.data_ptr(),
                    step.data_ptr() if device_step else None)
    prior = torch.linspace(-3, 0, K, device="cuda")
    ctx.mlp_init(mlp, 11, prior)
    # a fixed mid-training state: biases, moments and step count
    g = torch.Generator().manual_seed(K + t0)
    b.copy_(torch.randn(nB, generator=g) * 0.05)
    m.copy_(torch.randn(nW + nB, generator=g) * 1e-3)
    v.copy_(torch.rand(nW + nB, generator=g) * 1e-5 + 1e-9)
    mlp.adam_step = t0
    step.fill_(t0)
    # gradients of the cross-entropy on a fixed batch
    n = 256
    X = (torch.rand(n, 128, generator=g) * 2 - 1)
    X[:, 114:] = 0
    X16 = X.to(torch.bfloat16).view(torch.int16).cuda()
    tgt = torch.randint(0, K, (n,), generator=g, dtype=torch.int32).cuda()
    mask = torch.ones(n, dtype=torch.int32).cuda()
    gw = torch.zeros(nW, device="cuda")
    gb = torch.zeros(nB, device="cuda")
    loss = torch.zeros(1, device="cuda")
    ctx.mlp_loss_grad(mlp, X16, tgt, mask, n, gw, gb, loss)
    torch.cuda.synchronize()
    p0 = np.concatenate([w.cpu().numpy(), b.cpu().numpy()]).astype(np.float64)
    m0, v0 = m.cpu().numpy().astype(np.float64), v.cpu().numpy().astype(np.float64)
    grad = np.concatenate([gw.cpu().numpy(), gb.cpu().numpy()]).astype(np.float64)
    assert np.abs(grad).max() > 0
    lr = 1e-3
    ctx.mlp_adam(mlp, gw, gb, lr)
    torch.cuda.synchronize()
    assert mlp.adam_step == t0 + 1
    if device_step:
        assert int(step.item()) == t0 + 1
    p_ref, m_ref, v_ref = p0.copy(), m0.copy(), v0.copy()
    oracle.adam(p_ref, grad, m_ref, v_ref, lr, 0.9, 0.999, 1e-8, t0 + 1)
    p_gpu = np.concatenate([w.cpu().numpy(), b.cpu().numpy()]).astype(np.float64)
    # The moments are updated in fp32 with fp32 betas (1 - 0.9f is 2.4e-7 above 0.1 relative,
    # 1 - 0.999f 1.3e-5 below 0.001); the oracle keeps FP64 betas. Bounds per element, relative
    # to the magnitude of the two terms summed (m and v may cancel to near zero):
    sm = 0.9 * np.abs(m0) + 0.1 * np.abs(grad)
    sv = 0.999 * v0 + 0.001 * grad * grad
    em = np.abs(m.cpu().numpy() - m_ref)
    ev = np.abs(v.cpu().numpy() - v_ref)
    assert (em <= 1e-6 * sm + 1e-30).all(), (em / (1e-6 * sm + 1e-30)).max()
    assert (ev <= 3e-5 * sv + 1e-30).all(), (ev / (3e-5 * sv + 1e-30)).max()
    # parameters: the step's error (from m and v as above, then fp32 division and square root)
    # plus half an ulp of the fp32 weight it is added to
    t = t0 + 1
    c1, c2 = 1 - 0.9 ** t, 1 - 0.999 ** t
    d_ref = p_ref - p0
    err = np.abs(p_gpu - p_ref)
    tol = (2.0 ** -23 * np.abs(p_ref) + lr * (1e-6 * sm / c1) / (np.sqrt(v_ref / c2) + 1e-8)
           + 3e-5 * np.abs(d_ref) + 1e-12)
    assert (err <= tol).all(), (err / tol).max()
    assert np.abs(d_ref).max() > 1e-4  # the step moved the parameters
    # the bf16 working copy is the rounded fp32 master copy
    got_bf = w_bf.view(torch.bfloat16).float().cpu().numpy()
    np.testing.assert_array_equal(got_bf, bf16_round(w.cpu().numpy()))


def test_pass_a_over_candidates_bit_exact_1080p_scene(torch):
    sc = scenes.config_scene("terrain")
    W, H = 480, 270
    r = Renderer(sc, W, H, mode="unbiased", k=5, spatial_passes=2, M=8)
    lp = LearnedProposal(r.ctx, sc, k=256, seed=3)
    r.learned = lp
    orc = oracle.Scene(r.dev.nodes, r.dev.tris, sc.point_lights, sc.tri_lights, r.dev.alias)
    assert r.dev.union
    for f in range(3):
        cam = scenes.orbit_camera(sc.camera, f)
        r.begin_frame(cam)
        u, pv_now = r.uniforms(cam)
        i, p = r.cur, 1 - r.cur
        g, gp = r.gstruct(i), r.gstruct(p)
        cl, cq = lp.propose(u, g, r.M)
        planes, prev_planes = r.gbuf[i].cpu().numpy(), r.gbuf[p].cpu().numpy()
        prev_res = r.res[p].cpu().numpy().view(abi.RESERVOIR_DTYPE)
        ref = oracle.pass_a_candidates(orc, u, oracle.gbuffer(planes, W, H), oracle.gbuffer(prev_planes, W, H),
                                       prev_res, cl.cpu().numpy(), cq.cpu().numpy(), abi.RESERVOIR_DTYPE)
        for schedule in (abi.SCHEDULE_QUEUED, abi.SCHEDULE_INLINE):
            r.ctx.initial_temporal_candidates(u, g, gp, r.res[p], r.tmp, cl, cq, abi.Launch(16, 16, schedule))
            got = r.tmp.cpu().numpy().view(abi.RESERVOIR_DTYPE)
            bad = reservoir_mismatch(got, ref)
            assert not bad.any(), f"frame {f} schedule {schedule}: {int(bad.sum())} reservoirs differ"
        li = ref["lightIndex"][ref["sumWeights"] > 0]
        assert (li < 0).any() and (li >= 0).any()  # both kinds drawn from the proposal
        if f > 0:
            assert (ref["M"] > r.M).mean() > 0.3  # temporal reuse took place
        # the rest of the frame (Pass A, two Pass C), so the next frame has a history
        r.ctx.initial_temporal_candidates(u, g, gp, r.res[p], r.res[i], cl, cq)
        r.ctx.unbiased(u, r.k, 0, g, r.res[i], r.tmp)
        r.ctx.unbiased(u, r.k, 1, g, r.tmp, r.res[i])
        r.prev_pv = pv_now


def test_set_lights_unbinds_clusters(torch):
    sc = scenes.cornell_box(1)
    W, H = 32, 32
    r = Renderer(sc, W, H, M=4)
    lp = LearnedProposal(r.ctx, sc, k=64, seed=3)
    r.render_gbuffer(sc.camera, 0)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(0)
    lp.propose(u, g, 4)  # bound: runs
    d = r.dev
    r.ctx.set_lights(d.d_points, d.d_points.numel(), d.d_trilights, d.d_trilights.numel(), d.d_alias,
                     d.d_alias.numel())
    with pytest.raises(abi.RestirError) as e:
        lp.propose(u, g, 4)
    assert e.value.code == -3  # RESTIR_E_NOT_READY
    with pytest.raises(abi.RestirError) as e:
        lp.train(u, g, r.res[0], n=64)
    assert e.value.code == -3  # RESTIR_E_NOT_READY
    # binding the clusters again restores the proposal
    r.ctx.set_light_clusters(lp.k, lp.d_light_cluster, lp.d_light_pwithin, lp.d_begin, lp.d_count,
                             lp.d_lights, lp.d_alias)
    lp.propose(u, g, 4)
