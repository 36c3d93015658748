"""The batched-views and 4K stress workloads at their own sizes.

Batched views (16 views of 640x360..2560x1440 over 4 seeded terrains of 25k-200k triangles and
1k-50k point lights, unbiased ReSTIR M = 32, temporal reuse, 2 Pass C with k = 5): the 8-frame
sequence captured as one HIP graph equals the 16 views rendered one by one, bit for bit, at full
size; at a reduced size one view of each scene equals the oracle pass by pass, and the batched
views equal the single-view renderers there too.

4K stress workload (3840x2160, 2,333,058 triangles, 1M point lights in 4,096 clusters, FP16
G-buffer normals, learned proposal M = 8): the device tree equals the host tree byte for byte;
the shadow walk answers 20,000 random segments as the oracle does; two frames' Pass A over the
proposal's candidates, both Pass C and the shading equal the oracle over every pixel; the half
normals the primary-ray kernel writes are the oracle's float normals rounded to nearest.
"""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.pipeline import BatchRenderer, FrameSequenceGraph, Renderer, float_planes
from rt_helpers import HostScene, floats_equal_nan, random_segments, reservoir_mismatch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def bits(t):
    import torch as T
    return t.contiguous().view(-1).view(T.int32)


C3_KW = dict(mode="unbiased", k=5, M=32, spatial_passes=2)


# ----------------------------------------------------------------------------- batched views
@pytest.mark.parametrize("rays", ["queued", "persistent"])
def test_batched_16_views_graph_equals_view_by_view(torch, rays):
    scs, views = scenes.batched_view_set(3)
    seq = [[scenes.orbit_camera(sc.camera, f) for sc, _, _ in views] for f in range(8)]
    print(f"[c3] {len(views)} views, {sum(w * h for _, w, h in views)} px, rays {rays}", flush=True)
    singles = [Renderer(sc, w, h, **C3_KW) for sc, w, h in views]
    ref_imgs, ref_res = [], None
    for cams in seq:
        for r, cam in zip(singles, cams):
            r.render_frame(cam)
        ref_imgs.append(torch.cat([r.rgb for r in singles]).clone())
    ref_res = [r.res[r.cur].clone() for r in singles]
    torch.cuda.synchronize()
    del singles
    b = BatchRenderer(views, rays=rays, **C3_KW)
    b.render_frame(seq[0])  # sizes scratch and view tables outside the capture
    torch.cuda.synchronize()
    b.reset()
    torch.cuda.synchronize()
    fg = FrameSequenceGraph(b, seq)
    for rep in range(2):
        for v in b.views:
            for t in (*v["gbuf"], *v["res"], v["tmp"]):
                t.zero_()
        fg.replay()
        torch.cuda.synchronize()
        for f in range(8):
            assert torch.equal(bits(fg.images[f]), bits(ref_imgs[f])), (rep, f)
        for v, want in enumerate(ref_res):
            assert torch.equal(bits(b.views[v]["res"][b.cur]), bits(want)), (rep, v)


def test_batched_scenes_equal_oracle_at_reduced_size(torch):
    from test_gpu_parity import run_frame_and_compare
    scs, _ = scenes.batched_view_set(3)
    W, H = 192, 108
    views = [(sc, W, H) for sc in scs]
    batch = BatchRenderer(views, rays="queued", **C3_KW)
    singles = [Renderer(sc, W, H, **C3_KW) for sc in scs]
    for f in range(2):
        cams = [scenes.orbit_camera(sc.camera, f) for sc in scs]
        for i, (r, cam) in enumerate(zip(singles, cams)):
            print(f"[c3 oracle] scene {i} ({scs[i].n_tris} tris) frame {f}", flush=True)
            run_frame_and_compare(r, cam)  # every pass against the oracle
        batch.render_frame(cams)
        torch.cuda.synchronize()
        for v, r in enumerate(singles):
            assert torch.equal(bits(batch.images[v]), bits(r.rgb)), (f, v)
            assert torch.equal(bits(batch.views[v]["res"][batch.cur]), bits(r.res[r.cur])), (f, v)


# ------------------------------------------------------------------------ 4K stress workload
@pytest.fixture(scope="module")
def stress():
    sc = scenes.stress_scene()
    assert sc.n_tris == 2_333_058 and len(sc.point_lights) == 1_000_000
    nodes, tris, depth = abi.build_aabb_tree(sc.tri_xyz)
    return sc, nodes, tris, depth


def test_stress_device_tree_equals_host_tree(torch, stress):
    sc, nodes, tris, depth = stress
    ctx = abi.Context(0)
    n = sc.n_tris
    x = torch.from_numpy(np.ascontiguousarray(sc.tri_xyz, dtype=np.float32)).cuda()
    d_nodes = torch.full((abi.NODE_DTYPE.itemsize * (n - 1),), 0xAB, dtype=torch.uint8, device="cuda")
    d_tris = torch.full((abi.TRIANGLE_DTYPE.itemsize * n,), 0xAB, dtype=torch.uint8, device="cuda")
    dd = ctx.build_bvh(x, n, d_nodes, d_tris)
    torch.cuda.synchronize()
    assert dd == depth
    assert np.array_equal(d_nodes.cpu().numpy(), nodes.view(np.uint8))
    assert np.array_equal(d_tris.cpu().numpy(), tris.view(np.uint8))


def test_stress_wide_walk_matches_oracle(torch, stress):
    sc, nodes, tris, _ = stress
    ctx = abi.Context(0)
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).cuda()
    d_tris = torch.from_numpy(tris.view(np.uint8).copy()).cuda()
    ctx.set_bvh(d_nodes, len(nodes), d_tris, len(tris))
    lo, hi = sc.aabb()
    p1, p2 = random_segments(np.random.default_rng(21), lo, hi, 20000)
    ref_sh, _, _ = oracle.trace_shadow(oracle.Scene(nodes, tris), p1, p2)
    n = len(p1)
    sh = torch.zeros(n, dtype=torch.int32, device="cuda")
    for kw in ({}, {"one_ray_per_lane": True}):
        ctx.trace_shadow(torch.from_numpy(p1).cuda(), torch.from_numpy(p2).cuda(), n, sh, **kw)
        torch.cuda.synchronize()
        got = sh.cpu().numpy().astype(np.uint32)
        assert np.array_equal(got, ref_sh), f"{kw}: {int((got != ref_sh).sum())} verdicts differ"
    assert 0.05 < ref_sh.mean() < 0.99


def test_stress_half_normals_are_rounded_float_normals(torch, stress):
    sc, nodes, tris, _ = stress
    W, H = 480, 270
    r = Renderer(sc, W, H, normal_format=abi.NORMAL_F16)
    r.render_gbuffer(sc.camera, 0)
    got = r.gbuf[0].cpu().numpy()
    ref = oracle.primary_gbuffer(tris, abi.camera_struct(sc.camera), sc.tri_material, sc.materials, W, H)
    P = W * H
    halves = got[3 * P: 6 * P].view(np.float16)[: 3 * P]
    np.testing.assert_array_equal(halves, ref[3 * P: 6 * P].astype(np.float16))
    for lo, hi in ((0, 3 * P), (6 * P, 13 * P)):  # every other plane bit for bit
        assert floats_equal_nan(got[lo:hi], ref[lo:hi]).all()


def test_stress_4k_frames_equal_oracle(torch, stress):
    from restir_amd.learned import LearnedProposal
    sc, nodes, tris, _ = stress
    W, H, M = 3840, 2160, 8
    r = Renderer(sc, W, H, mode="unbiased", k=5, M=M, spatial_passes=2, normal_format=abi.NORMAL_F16)
    lp = LearnedProposal(r.ctx, sc, k=4096, seed=7, iterations=4, device_clustering=True)
    orc = oracle.Scene(r.dev.nodes, r.dev.tris, sc.point_lights, None, r.dev.alias)
    for f in range(2):
        cam = scenes.orbit_camera(sc.camera, f)
        print(f"[c5] 4K frame {f}", flush=True)
        r.begin_frame(cam)
        u, pv_now = r.uniforms(cam)
        i, p = r.cur, 1 - r.cur
        g, gp = r.gstruct(i), r.gstruct(p)
        cl, cq = lp.propose(u, g, M)
        r.ctx.initial_temporal_candidates(u, g, gp, r.res[p], r.res[i], cl, cq)
        got_a = r.res[i].cpu().numpy().view(abi.RESERVOIR_DTYPE)
        r.ctx.unbiased(u, r.k, 0, g, r.res[i], r.tmp)
        r.ctx.unbiased(u, r.k, 1, g, r.tmp, r.res[i])
        r.shade()
        torch.cuda.synchronize()
        og = oracle.gbuffer(r.gbuffer_host(i), W, H)
        ogp = oracle.gbuffer(r.gbuffer_host(p), W, H)
        prev = r.res[p].cpu().numpy().view(abi.RESERVOIR_DTYPE)
        a = oracle.pass_a_candidates(orc, u, og, ogp, prev, cl.cpu().numpy(), cq.cpu().numpy(),
                                     abi.RESERVOIR_DTYPE)
        bad = reservoir_mismatch(got_a, a)
        assert not bad.any(), f"frame {f}: {int(bad.sum())} of {W * H} Pass A reservoirs differ"
        c0 = oracle.pass_c(orc, u, r.k, 0, og, a)
        c1 = oracle.pass_c(orc, u, r.k, 1, og, c0)
        got = r.res[i].cpu().numpy().view(abi.RESERVOIR_DTYPE)
        bad = reservoir_mismatch(got, c1)
        assert not bad.any(), f"frame {f}: {int(bad.sum())} of {W * H} final reservoirs differ"
        img = oracle.shade(orc, cam["position"], 1.0, og, c1)
        assert floats_equal_nan(r.image(), img).all(), f"frame {f}: shading differs"
        if f == 1:
            assert (a["M"] > M).mean() > 0.3  # temporal reuse took place
        r.prev_pv = pv_now
