"""The learned proposal over more than 256 clusters (k_propose_wide: the output layer streamed
through LDS in column blocks of 128 clusters, up to the 4,096-way output of the 4K stress
workload), and training at that width.

- Exact: a K = 256 network embedded in a K = 512 or 4,096 network (its output rows at a cluster
  offset, every other row masked by a -30000 bias, the light tables shifted by the same offset)
  must draw the same lights with bit-identical q as k_propose does for the K = 256 network: the
  masked numerators are exactly zero and the running sums of the streamed scan take the same
  values in the same order.
- Density and distribution at K = 4,096: q of every candidate against the softmax of the GEMM
  path's logits, and a chi-square test of one pixel's draws over many frames.
- Gradients at K = 4,096 against the oracle's MLP, as for K = 256.
"""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.abi import MlpStruct, mlp_sizes
from restir_amd.learned import LearnedProposal
from restir_amd.pipeline import Renderer
from test_gpu_mlp import _expected_q, bf16_to_f32, quad_representative, rand_features

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def many_light_terrain(n_lights):
    return scenes.terrain(11, grid=48, n_boxes=60, n_point_lights=n_lights, extent=16.0)


class Embedded:
    """Device buffers of a K-way network and light tables holding lp's 256 clusters at `offset`."""

    def __init__(self, torch, lp, K, offset):
        nW, nB = mlp_sizes(K)
        w = lp.w_bf16.clone()
        n3 = 3 * 128 * 128
        self.w_bf16 = torch.zeros(nW, dtype=torch.int16, device="cuda")
        self.w_bf16[:n3] = w[:n3]
        self.w_bf16[n3 + offset * 128: n3 + (offset + 256) * 128] = w[n3:]
        self.b = torch.full((nB,), -30000.0, dtype=torch.float32, device="cuda")
        self.b[:384] = lp.b[:384]
        self.b[384 + offset: 384 + offset + 256] = lp.b[384:]
        self.w = torch.zeros(nW, dtype=torch.float32, device="cuda")
        self.m = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
        self.v = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
        self.mlp = MlpStruct(K, 0, self.w_bf16.data_ptr(), self.w.data_ptr(), self.b.data_ptr(),
                             self.m.data_ptr(), self.v.data_ptr())
        cl = lp.clusters
        begin = np.zeros(K, np.int32)
        count = np.zeros(K, np.int32)
        begin[offset: offset + 256] = cl.begin
        count[offset: offset + 256] = cl.count
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self.tables = (t(cl.light_cluster + offset), t(cl.light_pwithin), t(begin), t(count),
                       t(cl.lights), t(cl.alias.view(np.uint8)))


@pytest.mark.parametrize("K,offset", [(512, 0), (512, 256), (4096, 3840)])
def test_wide_proposal_equals_k256_when_embedded(torch, K, offset):
    sc = many_light_terrain(5000)
    W, H, M = 64, 48, 8
    r = Renderer(sc, W, H)
    lp = LearnedProposal(r.ctx, r.scene, k=256, seed=9)
    r.begin_frame(sc.camera)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(r.cur)
    cl, cq = lp.propose(u, g, M)
    ref_l, ref_q = cl.cpu().numpy().copy(), cq.cpu().numpy().copy()
    e = Embedded(torch, lp, K, offset)
    r.ctx.set_light_clusters(K, *e.tables)
    got_l = torch.full((M * W * H,), -7, dtype=torch.int32, device="cuda")
    got_q = torch.full((M * W * H,), -7.0, dtype=torch.float32, device="cuda")
    r.ctx.learned_candidates(u, g, e.mlp, lp.lo, lp.hi, M, lp.mix, got_l, got_q)
    gl, gq = got_l.cpu().numpy(), got_q.cpu().numpy()
    normal = r.gbuf[r.cur].cpu().numpy()[3 * W * H: 6 * W * H].reshape(3, -1)
    surf = np.abs(normal).sum(axis=0) > 0
    on = np.tile(surf, M)
    assert on.sum() > 1000
    np.testing.assert_array_equal(gl[on], ref_l[on])
    np.testing.assert_array_equal(gq[on].view(np.uint32), ref_q[on].view(np.uint32))
    assert np.all(gl[~on] == -7)  # background pixels get no candidates
    assert (ref_l[on] >= 0).mean() > 0.99


def test_wide_proposal_density_and_distribution(torch):
    K = 4096
    sc = many_light_terrain(20000)
    W, H, M = 64, 48, 8
    r = Renderer(sc, W, H)
    lp = LearnedProposal(r.ctx, r.scene, k=K, seed=9, iterations=6)
    r.begin_frame(sc.camera)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(r.cur)
    cl, cq = lp.propose(u, g, M)
    cand = cl.cpu().numpy().reshape(M, W * H)
    q = cq.cpu().numpy().reshape(M, W * H)
    normal = r.gbuf[r.cur].cpu().numpy()[3 * W * H: 6 * W * H].reshape(3, -1)
    surf = np.flatnonzero(np.abs(normal).sum(axis=0) > 0)[::7][:256]
    n = 64 * (len(surf) // 64)
    assert n >= 64
    surf = surf[:n]
    reps = np.array([quad_representative(normal, W, H, p) for p in surf], dtype=np.uint32)
    X = torch.zeros(n * 128, dtype=torch.int16, device="cuda")
    r.ctx.mlp_features(g, sc.camera["position"], lp.lo, lp.hi, torch.from_numpy(reps).cuda(), n, X)
    logits = torch.zeros(n, K, device="cuda")
    r.ctx.mlp_forward(lp.mlp, X, n, logits)
    l32 = logits.cpu().numpy()
    src_p = r.dev.alias["oriProb"].astype(np.float64)
    pw = lp.clusters.light_pwithin.astype(np.float64)
    got_l = cand[:, surf].T
    got_q = q[:, surf].T
    valid = got_l >= 0
    assert valid.mean() > 0.99
    exp_q = np.zeros_like(got_q, dtype=np.float64)
    for j in range(M):
        exp_q[:, j] = _expected_q(lp, l32, np.where(valid[:, j], got_l[:, j], 0), pw, src_p)
    rel = np.abs(got_q[valid] - exp_q[valid]) / exp_q[valid]
    # fused kernel and GEMM path round hidden activations to bf16 at slightly different sums
    assert np.median(rel) < 1e-2 and np.mean(rel < 5e-2) > 0.95
    # one pixel's candidates over many frames follow q (chi-square over clusters, bins with
    # expected counts below 5 pooled)
    pix = int(surf[0])
    counts = np.zeros(len(src_p))
    frames = 600
    for f in range(frames):
        u.frame = 1000 + f
        cl, cq = lp.propose(u, g, M)
        ls = cl.view(M, W * H)[:, pix].cpu().numpy()
        np.add.at(counts, ls[ls >= 0], 1)
    all_l = np.arange(len(src_p))
    qa = _expected_q(lp, np.repeat(l32[0:1], len(all_l), axis=0), all_l, pw, src_p)
    assert abs(qa.sum() - 1.0) < 1e-3
    c_of = lp.clusters.light_cluster
    exp_c = np.bincount(c_of, weights=qa, minlength=K) * frames * M
    obs_c = np.bincount(c_of, weights=counts, minlength=K)
    big = exp_c >= 5
    chi2 = (((obs_c[big] - exp_c[big]) ** 2) / exp_c[big]).sum() + \
        ((obs_c[~big].sum() - exp_c[~big].sum()) ** 2) / max(exp_c[~big].sum(), 1.0)
    dof = int(big.sum())
    assert dof >= 20
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def test_wide_proposal_rejects_more_than_8_candidates(torch):
    sc = many_light_terrain(5000)
    r = Renderer(sc, 32, 32)
    lp = LearnedProposal(r.ctx, r.scene, k=256, seed=9)
    e = Embedded(torch, lp, 512, 0)
    r.ctx.set_light_clusters(512, *e.tables)
    r.begin_frame(sc.camera)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(r.cur)
    cl = torch.zeros(9 * 32 * 32, dtype=torch.int32, device="cuda")
    cq = torch.zeros(9 * 32 * 32, dtype=torch.float32, device="cuda")
    with pytest.raises(abi.RestirError):
        r.ctx.learned_candidates(u, g, e.mlp, lp.lo, lp.hi, 9, lp.mix, cl, cq)


def test_gradients_match_oracle_4096(torch):
    K, n = 4096, 256
    r = Renderer(many_light_terrain(5000), 32, 32)
    nW, nB = mlp_sizes(K)
    w_bf16 = torch.zeros(nW, dtype=torch.int16, device="cuda")
    w = torch.zeros(nW, dtype=torch.float32, device="cuda")
    b = torch.zeros(nB, dtype=torch.float32, device="cuda")
    m = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
    v = torch.zeros(nW + nB, dtype=torch.float32, device="cuda")
    mlp = MlpStruct(K, 0, w_bf16.data_ptr(), w.data_ptr(), b.data_ptr(), m.data_ptr(), v.data_ptr())
    prior = torch.from_numpy(np.random.default_rng(4).normal(0, 1, K).astype(np.float32)).cuda()
    r.ctx.mlp_init(mlp, 5, prior)
    X16, Xf = rand_features(torch, n, 1)
    rng = np.random.default_rng(2)
    tgt = rng.integers(0, K, n).astype(np.int32)
    mask = (rng.uniform(size=n) < 0.8).astype(np.uint32)
    gw = torch.zeros(nW, device="cuda")
    gb = torch.zeros(nB, device="cuda")
    loss = torch.zeros(1, device="cuda")
    r.ctx.mlp_loss_grad(mlp, X16, torch.from_numpy(tgt).cuda(), torch.from_numpy(mask).cuda(), n,
                        gw, gb, loss)
    Wf, Bf = bf16_to_f32(w_bf16), b.cpu().numpy()
    ref_loss, rgW, rgB = oracle.mlp_loss_grad(Wf, Bf, K, Xf, tgt, mask.astype(np.uint8), emulate_bf16=True)
    assert abs(loss.item() - ref_loss) < 1e-3 * abs(ref_loss)
    gW, gB = gw.cpu().numpy().astype(np.float64), gb.cpu().numpy().astype(np.float64)
    for layer in range(4):
        sl = slice(layer * 128 * 128, layer * 128 * 128 + (128 if layer < 3 else K) * 128)
        bl = slice(layer * 128, layer * 128 + (128 if layer < 3 else K))
        ew = np.linalg.norm(gW[sl] - rgW[sl]) / np.linalg.norm(rgW[sl])
        eb = np.linalg.norm(gB[bl] - rgB[bl]) / np.linalg.norm(rgB[bl])
        assert ew < 2e-2 and eb < 2e-2, (layer, ew, eb)
    # the forward pass at this width as well
    logits = torch.zeros(n, K, dtype=torch.float32, device="cuda")
    r.ctx.mlp_forward(mlp, X16, n, logits)
    ref = oracle.mlp_forward(Wf, Bf, K, Xf, emulate_bf16=True)
    got = logits.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-3
