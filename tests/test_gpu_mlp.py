"""GPU checks of the learned light proposal: MFMA forward and backward against the oracle's MLP
(which the CPU suite pins by finite differences), features, the proposal density q against the
logits, the sampled distribution against q, unbiasedness of RIS over the proposal, and what
online training gains over the plain source distribution."""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.learned import LearnedProposal
from restir_amd.pipeline import Renderer
from rt_helpers import HostScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def small_terrain():
    return scenes.terrain(11, grid=48, n_boxes=60, n_point_lights=2000, extent=16.0)


def bf16_to_f32(t16):
    import torch
    return t16.view(torch.bfloat16).float().cpu().numpy()


def params(lp):
    return bf16_to_f32(lp.w_bf16), lp.b.cpu().numpy()


def rand_features(torch, n, seed):
    g = torch.Generator().manual_seed(seed)
    X = (torch.rand(n, 128, generator=g) * 2 - 1)
    X[:, 114:] = 0
    Xb = X.to(torch.bfloat16)
    return Xb.view(torch.int16).cuda(), Xb.float().numpy()


def test_forward_matches_oracle(torch):
    r = Renderer(small_terrain(), 64, 64)
    lp = LearnedProposal(r.ctx, r.scene, k=256, seed=3)
    X16, Xf = rand_features(torch, 512, 0)
    logits = torch.zeros(512, 256, dtype=torch.float32, device="cuda")
    r.ctx.mlp_forward(lp.mlp, X16, 512, logits)
    W, B = params(lp)
    ref = oracle.mlp_forward(W, B, 256, Xf, emulate_bf16=True)
    got = logits.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    # fp32 MFMA accumulation vs FP64, with the same bf16 rounding of hidden activations
    assert rel < 2e-3, rel
    assert np.abs(got - ref).max() < 2e-2 * np.abs(ref).max()


def test_gradients_match_oracle(torch):
    r = Renderer(small_terrain(), 64, 64)
    lp = LearnedProposal(r.ctx, r.scene, k=256, seed=5)
    n = 256
    X16, Xf = rand_features(torch, n, 1)
    rng = np.random.default_rng(2)
    tgt = rng.integers(0, 256, n).astype(np.int32)
    mask = (rng.uniform(size=n) < 0.8).astype(np.uint32)
    nW, nB = abi.mlp_sizes(256)
    gw = torch.zeros(nW, device="cuda")
    gb = torch.zeros(nB, device="cuda")
    loss = torch.zeros(1, device="cuda")
    r.ctx.mlp_loss_grad(lp.mlp, X16, torch.from_numpy(tgt).cuda(), torch.from_numpy(mask).cuda(), n,
                        gw, gb, loss)
    W, B = params(lp)
    ref_loss, rgW, rgB = oracle.mlp_loss_grad(W, B, 256, Xf, tgt, mask.astype(np.uint8), emulate_bf16=True)
    assert abs(loss.item() - ref_loss) < 1e-3 * abs(ref_loss)
    gW, gB = gw.cpu().numpy().astype(np.float64), gb.cpu().numpy().astype(np.float64)
    for layer in range(4):
        sl = slice(layer * 128 * 128, layer * 128 * 128 + (128 if layer < 3 else 256) * 128)
        bl = slice(layer * 128, layer * 128 + (128 if layer < 3 else 256))
        ew = np.linalg.norm(gW[sl] - rgW[sl]) / np.linalg.norm(rgW[sl])
        eb = np.linalg.norm(gB[bl] - rgB[bl]) / np.linalg.norm(rgB[bl])
        # bf16 dlogits / dhidden as on the GPU; residual = fp32 accumulation order
        assert ew < 2e-2 and eb < 2e-2, (layer, ew, eb)


def test_features_match_oracle(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 64, 64)
    r.render_gbuffer(sc.camera, 0)
    lo, hi = sc.aabb()
    pix = np.arange(64 * 64, dtype=np.uint32)
    out = torch.zeros(len(pix) * 128, dtype=torch.int16, device="cuda")
    r.ctx.mlp_features(r.gstruct(0), sc.camera["position"], lo, hi, torch.from_numpy(pix).cuda(),
                       len(pix), out)
    got = bf16_to_f32(out).reshape(-1, 128)
    ref = oracle.mlp_features(oracle.gbuffer(r.gbuf[0].cpu().numpy(), 64, 64), sc.camera["position"],
                              lo, hi, pix)
    # bf16 rounding of values that agree to ~1e-5 before rounding: at most one bf16 step apart
    step = np.maximum(np.abs(ref), 1e-30) * 2.0 ** -7
    assert np.all(np.abs(got - ref) <= step + 1e-5)
    assert np.mean(np.abs(got - ref) <= np.abs(ref) * 2.0 ** -8 + 1e-6) > 0.97


def _expected_q(lp, logits, lights, pwithin, src):
    """q of each candidate from the logits; the kernel keeps the softmax numerators
    exp(z - max z) in fp16 and normalises by their sum."""
    z = logits.astype(np.float64)
    e = np.exp(z - z.max(axis=1, keepdims=True)).astype(np.float16).astype(np.float64)
    pi = e / e.sum(axis=1, keepdims=True)
    c = lp.clusters.light_cluster[lights]
    rows = np.arange(len(lights))[:, None] if lights.ndim == 2 else np.arange(len(lights))
    return (1 - lp.mix) * pi[rows, c] * pwithin[lights] + lp.mix * src[lights]


def quad_representative(normal, W, H, p):
    """The pixel whose features the proposal uses for pixel p: the first surface pixel of p's
    2x2 quad in the order (0,0), (1,0), (0,1), (1,1), else the quad's top-left pixel."""
    x0, y0 = (p % W) & ~1, (p // W) & ~1
    for j in range(4):
        x, y = x0 + (j & 1), y0 + (j >> 1)
        if x < W and y < H and np.any(normal[:, y * W + x] != 0):
            return y * W + x
    return y0 * W + x0


def test_proposal_density_and_distribution(torch):
    sc = small_terrain()
    W, H = 64, 48
    r = Renderer(sc, W, H)
    lp = LearnedProposal(r.ctx, r.scene, k=256, seed=9)
    r.begin_frame(sc.camera)
    u, _ = r.uniforms(sc.camera)
    g = r.gstruct(r.cur)
    M = 8
    cl, cq = lp.propose(u, g, M)
    cand = cl.cpu().numpy().reshape(M, W * H)
    q = cq.cpu().numpy().reshape(M, W * H)
    normal = r.gbuf[r.cur].cpu().numpy()[3 * W * H: 6 * W * H].reshape(3, -1)
    surf = np.flatnonzero(np.abs(normal).sum(axis=0) > 0)[::7][:256]
    assert len(surf) >= 64
    # logits through the GEMM path at the pixel that stands for each pixel's 2x2 quad
    n = 256 if len(surf) >= 256 else 64 * (len(surf) // 64)
    surf = surf[:n]
    reps = np.array([quad_representative(normal, W, H, p) for p in surf], dtype=np.uint32)
    X = torch.zeros(n * 128, dtype=torch.int16, device="cuda")
    r.ctx.mlp_features(g, sc.camera["position"], lp.lo, lp.hi, torch.from_numpy(reps).cuda(), n, X)
    logits = torch.zeros(n, 256, device="cuda")
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
    # the empirical distribution of one pixel's candidates over many frames follows q
    pix = int(surf[0])
    counts = {}
    frames = 600
    for f in range(frames):
        u.frame = 1000 + f
        cl, cq = lp.propose(u, g, M)
        ls = cl.view(M, W * H)[:, pix].cpu().numpy()
        for l in ls:
            counts[int(l)] = counts.get(int(l), 0) + 1
    x16 = l32[0:1]
    all_l = np.arange(len(sc.point_lights))
    qa = _expected_q(lp, np.repeat(x16, len(all_l), axis=0), all_l, pw, src_p)
    assert abs(qa.sum() - 1.0) < 1e-3
    # chi-square over clusters (aggregating lights) with expected counts >= 5
    c_of = lp.clusters.light_cluster
    exp_c = np.bincount(c_of, weights=qa, minlength=256) * frames * M
    obs_c = np.zeros(256)
    for l, v in counts.items():
        obs_c[c_of[l]] += v
    big = exp_c >= 5
    chi2 = (((obs_c[big] - exp_c[big]) ** 2) / exp_c[big]).sum() + \
        ((obs_c[~big].sum() - exp_c[~big].sum()) ** 2) / max(exp_c[~big].sum(), 1.0)
    dof = int(big.sum())
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def independent_frame(r):
    """One frame with a static camera whose temporal reuse meets an empty history. The temporal
    combine still runs (the previous G-buffer matches), which replaces Pass A's selection-time
    W (reservoir.glsl:33) by sumWeights / (M * pHat); frames are then independent estimates."""
    r.res[r.cur].zero_()  # the buffer the next frame reads as history
    r.render_frame(r.scene.camera)


def independent_frames_mean(r, frames):
    """Mean shaded image over independent frames (the first frame, whose previous G-buffer is
    still empty, is rendered but not counted)."""
    import torch
    acc = torch.zeros(3 * r.W * r.H, dtype=torch.float64, device="cuda")
    independent_frame(r)
    for f in range(frames):
        independent_frame(r)
        acc += torch.nan_to_num(r.rgb.double(), nan=0.0)
    return (acc / frames).view(3, r.H, r.W).cpu().numpy()


def frames_mean_and_var(r, frames):
    """Per-pixel mean over independent frames and the variance of that mean."""
    import torch
    acc = torch.zeros(3 * r.W * r.H, dtype=torch.float64, device="cuda")
    acc2 = torch.zeros_like(acc)
    independent_frame(r)
    for f in range(frames):
        independent_frame(r)
        x = torch.nan_to_num(r.rgb.double(), nan=0.0)
        acc += x
        acc2 += x * x
    mean = acc / frames
    var_mean = (acc2 / frames - mean * mean).clamp_min(0) / (frames - 1)
    return (mean.view(3, r.H, r.W).cpu().numpy(), var_mean.view(3, r.H, r.W).cpu().numpy())


def cornell_with_ceiling_lights():
    """The Cornell box (lights 0.25 from the walls) plus its two ceiling triangles, shrunk to half
    size about the ceiling's centre and lowered by 0.3, as emissive triangle lights: both kinds in
    one alias table (the union). (Lights just under the ceiling would light it with a near-singular
    1/d^2 that a stratified ground truth cannot integrate to 1%.)"""
    import dataclasses
    sc = scenes.cornell_box(1, light_margin=0.25)
    tri = sc.tri_xyz.reshape(-1, 3, 3)
    ceiling = np.flatnonzero(np.all(np.abs(tri[:, :, 1] - tri[:, :, 1].max()) < 1e-6, axis=1))[:2]
    t = tri[ceiling].copy()
    centre = t.reshape(-1, 3).mean(axis=0)
    t = centre + 0.5 * (t - centre)
    t[:, :, 1] -= 0.3
    tl = scenes.make_tri_lights(t.reshape(-1, 9), np.array([[6.0, 5.0, 4.0], [4.0, 5.0, 6.0]]))
    return dataclasses.replace(sc, tri_lights=tl)


@pytest.mark.parametrize("lights", ["points", "points+triangles"])
@pytest.mark.parametrize("learned", [False, True])
def test_ris_converges_to_fp64_ground_truth(torch, learned, lights):
    """Independent frames of unbiased ReSTIR (plain, or over the online-trained proposal)
    average to the FP64 oracle's direct lighting of the Cornell box: the squared error of the
    mean matches the variance the frames themselves show, and the image mean agrees to 1%; the per-pixel
    residual (about 6% relative RMSE after 1500 frames) is the frames' own noise. With triangle
    lights the alias table and the proposal's clusters cover both kinds, and the ground truth
    integrates the triangles with 1024 stratified samples each."""
    sc = scenes.cornell_box(1, light_margin=0.25) if lights == "points" else cornell_with_ceiling_lights()
    W = H = 96
    hs = HostScene(sc)
    r = Renderer(sc, W, H, mode="unbiased", k=3, M=16)
    assert r.dev.union == (lights != "points")
    if learned:
        r.learned = LearnedProposal(r.ctx, sc, k=64, seed=4)
        r.train_samples = 1024
    r.render_gbuffer(sc.camera, 0)
    gt = oracle.ground_truth_f64(hs.oracle, sc.camera["position"], oracle.gbuffer(r.gbuf[0].cpu().numpy(), W, H),
                                 n_tri_samples=1024 if lights != "points" else 0)
    mean, var_mean = frames_mean_and_var(r, 1500)
    surf = gt.sum(axis=0) > 0
    err2 = ((mean - gt)[:, surf] ** 2).sum()
    t = err2 / var_mean[:, surf].sum()
    # unbiased: E[err^2] = sum of variances (t ~ 1); heavy tails make the sample variance small,
    # so allow up to 2; a bias of a few percent gives t >> 10 here
    assert t < 2.0, t
    assert abs(mean[:, surf].sum() - gt[:, surf].sum()) < 0.01 * gt[:, surf].sum()
    rel_rmse = np.sqrt(((mean - gt)[:, surf] ** 2).mean()) / np.sqrt((gt[:, surf] ** 2).mean())
    assert rel_rmse < 0.1, rel_rmse


def occluded_fraction_of_picks(r, learned):
    """Fraction of this frame's Pass-A RIS picks (before any visibility test) that are shadowed."""
    import torch as T
    u, _ = r.uniforms(r.cam)
    u.flags = 0
    g = r.gstruct(r.cur)
    out = T.zeros_like(r.tmp)
    if learned:
        cl, cq = r.learned.propose(u, g, r.M)
        r.ctx.initial_temporal_candidates(u, g, g, r.res[0], out, cl, cq)
    else:
        r.ctx.initial_temporal(u, g, g, r.res[0], out)
    rv = out.view(-1, 16)
    lit = rv[:, 10].view(T.float32) != 0
    pos = r.gbuf[r.cur][: 3 * r.W * r.H].view(3, -1).t()
    p1 = T.zeros(int(lit.sum()), 4, device="cuda")
    p2 = T.zeros_like(p1)
    p1[:, :3] = pos[lit]
    p2[:, :3] = rv[lit, 0:3].view(T.float32)
    sh = T.zeros(p1.shape[0], dtype=T.int32, device="cuda")
    r.ctx.trace_shadow(p1, p2, p1.shape[0], sh)
    return sh.float().mean().item()


def test_training_lowers_loss_and_learns_visibility(torch):
    """Online training on the terrain scene (10k lights among 500 boxes) lowers the
    cross-entropy, and the trained proposal learns where lights are visible: far fewer of its
    RIS picks are shadowed than of the alias table's, which samples by power alone."""
    sc = scenes.terrain(2)
    W, H = 240, 136
    r = Renderer(sc, W, H, mode="unbiased", k=5, M=8, spatial_passes=2)
    r.learned = LearnedProposal(r.ctx, sc, k=256, seed=6)
    losses = []
    for f in range(80):
        r.render_frame(sc.camera)
        losses.append(r.learned.loss.item())
    assert np.mean(losses[-10:]) < 0.6 * np.mean(losses[:3]), losses
    plain = occluded_fraction_of_picks(r, learned=False)
    learned = occluded_fraction_of_picks(r, learned=True)
    assert plain > 0.5
    assert learned < 0.75 * plain, (plain, learned)
