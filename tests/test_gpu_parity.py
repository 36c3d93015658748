"""GPU parity: the HIP path, called through the C ABI, against the CPU oracle on the same inputs.

Bar: bit-exact reservoirs (all 16 words; two NaNs in a float field count as equal), bit-exact
shadow hit/miss and node-visit counts, bit-exact G-buffer and shaded image.
"""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from restir_amd.pipeline import Renderer
from rt_helpers import HostScene, floats_equal_nan, random_segments, reservoir_mismatch, segments_from_gbuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def host_planes(r, i):
    return r.gbuf[i].cpu().numpy()


def host_res(t):
    return t.cpu().numpy().view(abi.RESERVOIR_DTYPE)


def oracle_for(r):
    sc = r.scene
    return oracle.Scene(r.dev.nodes, r.dev.tris, sc.point_lights if len(sc.point_lights) else None,
                        sc.tri_lights if len(sc.tri_lights) else None, r.dev.alias)


def assert_res_equal(gpu, ref, what):
    bad = reservoir_mismatch(gpu, ref)
    assert not bad.any(), f"{what}: {int(bad.sum())} of {len(bad)} reservoirs differ; first {np.flatnonzero(bad)[:5]}"


def run_frame_and_compare(r, cam, *, check_shade=True):
    """One frame through the C ABI; every pass compared with the oracle on the same inputs."""
    orc = oracle_for(r)
    r.begin_frame(cam)
    u, pv_now = r.uniforms(cam)
    i, p = r.cur, 1 - r.cur
    W, H = r.W, r.H
    g_cur, g_prev = host_planes(r, i), host_planes(r, p)
    og, ogp = oracle.gbuffer(g_cur, W, H), oracle.gbuffer(g_prev, W, H)
    prev_res = host_res(r.res[p])
    if r.mode == "unbiased":
        n = r.spatial_passes
        first = r.tmp if n % 2 == 1 else r.res[i]
        r.ctx.initial_temporal(u, r.gstruct(i), r.gstruct(p), r.res[p], first)
        a_ref = oracle.pass_a(orc, u, og, ogp, prev_res, abi.RESERVOIR_DTYPE)
        assert_res_equal(host_res(first), a_ref, f"pass A frame {r.frame}")
        src, src_ref = first, a_ref
        # as the renderer runs them: a later Pass C with RESTIR_INPUT_VISIBLE_FLAG (it does not
        # retrace the segments the previous pass found unoccluded); the oracle traces every ray
        later = u
        if r.reuse_verdicts and (u.flags & abi.VISIBILITY_REUSE_FLAG):
            later = abi.Uniforms.from_buffer_copy(u)
            later.flags |= abi.INPUT_VISIBLE_FLAG
        for j in range(n):
            dst = r.res[i] if (n - j) % 2 == 1 else r.tmp
            r.ctx.unbiased(later if j > 0 else u, r.k, j, r.gstruct(i), src, dst)
            c_ref = oracle.pass_c(orc, u, r.k, j, og, src_ref)
            assert_res_equal(host_res(dst), c_ref, f"pass C{j} frame {r.frame}")
            src, src_ref = dst, c_ref
        final_ref = src_ref
    else:
        r.ctx.initial_temporal(u, r.gstruct(i), r.gstruct(p), r.res[p], r.res[i])
        a_ref = oracle.pass_a(orc, u, og, ogp, prev_res, abi.RESERVOIR_DTYPE)
        assert_res_equal(host_res(r.res[i]), a_ref, f"pass A frame {r.frame}")
        cur_ref = a_ref
        for j in range(r.spatial_passes):
            r.ctx.spatial(u, 2 * j, r.gstruct(i), r.res[i], r.res[p])
            b0 = oracle.pass_b(orc, u, 2 * j, og, cur_ref)
            assert_res_equal(host_res(r.res[p]), b0, f"pass B{2 * j} frame {r.frame}")
            r.ctx.spatial(u, 2 * j + 1, r.gstruct(i), r.res[p], r.res[i])
            b1 = oracle.pass_b(orc, u, 2 * j + 1, og, b0)
            assert_res_equal(host_res(r.res[i]), b1, f"pass B{2 * j + 1} frame {r.frame}")
            cur_ref = b1
        final_ref = cur_ref
    if check_shade:
        r.shade()
        img = r.image()
        ref_img = oracle.shade(orc, cam["position"], 1.0, og, final_ref)
        eq = floats_equal_nan(img, ref_img)
        assert eq.all(), f"shade: {int((~eq).sum())} values differ"
    r.prev_pv = pv_now
    return a_ref, final_ref


def test_primary_gbuffer_matches_oracle(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 256, 256)
    r.render_gbuffer(sc.camera, 0)
    gpu = host_planes(r, 0)
    hs = HostScene(sc)
    ref = hs.oracle_gbuffer(256, 256)
    eq = floats_equal_nan(gpu, ref)
    assert eq.all(), f"{int((~eq).sum())} G-buffer values differ"
    normal = gpu[3 * 65536: 6 * 65536].reshape(3, -1)
    assert (np.abs(normal).sum(axis=0) > 0).mean() > 0.6  # the room fills most of the view


@pytest.mark.parametrize("tile", [(16, 16), (64, 1), (8, 8)])
def test_biased_pipeline_cornell_bit_exact(torch, tile):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 256, 256, mode="biased", spatial_passes=1, spatial_neighbors=5,
                 launch=abi.Launch(*tile))
    a1, _ = run_frame_and_compare(r, sc.camera)
    assert (a1["sumWeights"] > 0).mean() > 0.5
    # static camera: the second frame reprojects every pixel onto itself and reuses it
    a2, _ = run_frame_and_compare(r, sc.camera)
    assert (a2["M"] > 32).mean() > 0.5


@pytest.mark.parametrize("k,schedule", [(3, abi.SCHEDULE_QUEUED), (5, abi.SCHEDULE_QUEUED),
                                        (3, abi.SCHEDULE_INLINE), (5, abi.SCHEDULE_INLINE)])
def test_unbiased_pipeline_cornell_bit_exact(torch, k, schedule):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 256, 256, mode="unbiased", k=k, spatial_passes=1,
                 launch=abi.Launch(16, 16, schedule))
    run_frame_and_compare(r, sc.camera)
    run_frame_and_compare(r, sc.camera)


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_QUEUED, abi.SCHEDULE_INLINE])
def test_unbiased_without_visibility_bit_exact(torch, schedule):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 200, 150, mode="unbiased", k=4, spatial_passes=2, visibility=False,
                 launch=abi.Launch(8, 8, schedule))
    run_frame_and_compare(r, sc.camera)
    run_frame_and_compare(r, sc.camera)


def test_unbiased_two_passes_moving_camera_terrain(torch):
    sc = scenes.terrain(2)
    r = Renderer(sc, 480, 270, mode="unbiased", k=5, spatial_passes=2)
    for f in range(3):
        run_frame_and_compare(r, scenes.orbit_camera(sc.camera, f))


def test_unbiased_terrain_with_point_and_triangle_lights(torch):
    """The 1080p workload's scene: 10k point lights and 2k emissive box triangles in one alias
    table (the union); every pass and the shading equal the oracle's on the same inputs."""
    sc = scenes.config_scene("terrain")
    r = Renderer(sc, 480, 270, mode="unbiased", k=5, spatial_passes=2)
    assert r.dev.union and len(r.dev.alias) == len(sc.point_lights) + len(sc.tri_lights) == 12000
    for f in range(3):
        a, _ = run_frame_and_compare(r, scenes.orbit_camera(sc.camera, f))
    li = a["lightIndex"][a["sumWeights"] > 0]
    assert (li < 0).any() and (li >= 0).any()  # both kinds were drawn


def test_input_visible_flag_same_reservoirs_fewer_rays(torch):
    """RESTIR_INPUT_VISIBLE_FLAG on the second Pass C of a frame (its input is the first's
    visibility-tested output): the same reservoirs as without it, bit for bit, and fewer rays in
    the queues: the centre rays of pixels that kept their own sample and the neighbour rays from
    the pixel that supplied the merged sample retrace segments the first pass found unoccluded."""
    sc = scenes.config_scene("terrain")
    W, H, k = 480, 270, 5
    r = Renderer(sc, W, H, mode="unbiased", k=k, spatial_passes=2)
    for f in range(3):  # temporal history
        r.render_frame(scenes.orbit_camera(sc.camera, f))
    cam = scenes.orbit_camera(sc.camera, 3)
    r.begin_frame(cam)
    u, _ = r.uniforms(cam)
    i, p = r.cur, 1 - r.cur
    g = r.gstruct(i)
    r.ctx.initial_temporal(u, g, r.gstruct(p), r.res[p], r.res[i])
    r.ctx.unbiased(u, k, 0, g, r.res[i], r.tmp)
    later = abi.Uniforms.from_buffer_copy(u)
    later.flags |= abi.INPUT_VISIBLE_FLAG
    npx = W * H
    p1 = torch.zeros(npx * k, 4, dtype=torch.float32, device="cuda")
    p2 = torch.zeros_like(p1)
    outs, rays = [], []
    for uu in (u, later):
        out = torch.zeros_like(r.tmp)
        r.ctx.unbiased(uu, k, 1, g, r.tmp, out)
        rays.append((r.ctx.unbiased_rays(0, p1, p2, npx * k), r.ctx.unbiased_rays(1, p1, p2, npx * k)))
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    (c0, n0), (c1, n1) = rays
    print(f"[input visible] centre rays {c0} -> {c1}, neighbour rays {n0} -> {n1}", flush=True)
    assert c1 < 0.9 * c0 and n1 < 0.95 * n0, rays


def test_input_visible_flag_reuses_surface_records(torch):
    """A flagged Pass C right after a queued Pass C on the same G-buffer planes, size and k reuses
    that pass's packed surface records (no k_pack_surface launch). On a copy of the planes, after
    a queued Pass A on the context, or after a Pass C with another k it packs its own. The
    reservoirs equal those of an unflagged pass in every case."""
    sc = scenes.config_scene("terrain")
    W, H, k = 160, 90, 5
    r = Renderer(sc, W, H, mode="unbiased", k=k, spatial_passes=2)
    for f in range(2):
        r.render_frame(scenes.orbit_camera(sc.camera, f))
    cam = scenes.orbit_camera(sc.camera, 2)
    r.begin_frame(cam)
    u, _ = r.uniforms(cam)
    i, p = r.cur, 1 - r.cur
    g = r.gstruct(i)
    r.ctx.initial_temporal(u, g, r.gstruct(p), r.res[p], r.res[i])
    later = abi.Uniforms.from_buffer_copy(u)
    later.flags |= abi.INPUT_VISIBLE_FLAG
    first, junk = torch.zeros_like(r.tmp), torch.zeros_like(r.tmp)
    r.ctx.unbiased(u, k, 0, g, r.res[i], first)
    want = torch.zeros_like(r.tmp)
    r.ctx.unbiased(u, k, 1, g, first, want)  # unflagged: the reference result
    planes = r.gbuf[i].clone()
    g_copy = abi.gbuffer_struct(planes.data_ptr(), W, H)

    def packs_and_out(before, g2):
        r.ctx.unbiased(u, k, 0, g, r.res[i], junk)  # a queued Pass C on g packs its records
        before()
        torch.cuda.synchronize()
        r.ctx.profile_read()
        r.ctx.profile_enable(True)
        out = torch.zeros_like(r.tmp)
        r.ctx.unbiased(later, k, 1, g2, first, out)
        torch.cuda.synchronize()
        rows = r.ctx.profile_read()
        r.ctx.profile_enable(False)
        return rows.get("k_pack_surface", (0,))[0], out

    cases = {
        "same planes": (lambda: None, g, 0),
        "copied planes": (lambda: None, g_copy, 1),
        "queued Pass A between": (lambda: r.ctx.initial_temporal(u, g, r.gstruct(p), r.res[p], torch.zeros_like(r.tmp)),
                                  g, 1),
        "other k between": (lambda: r.ctx.unbiased(u, 3, 0, g, r.res[i], torch.zeros_like(r.tmp)), g, 1),
    }
    for name, (before, g2, packs) in cases.items():
        n, out = packs_and_out(before, g2)
        assert n == packs, (name, n)
        assert torch.equal(out, want), name


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_QUEUED, abi.SCHEDULE_INLINE])
def test_standard_weights_bit_exact(torch, schedule):
    """RESTIR_STANDARD_WEIGHTS_FLAG (standard RIS W, triangle lights by area, sumW = p_hat W M
    after Pass C) on the 1080p scene's point and triangle lights: every pass equals the oracle."""
    sc = scenes.config_scene("terrain")
    r = Renderer(sc, 320, 180, mode="unbiased", k=5, spatial_passes=2, standard_weights=True,
                 launch=abi.Launch(16, 16, schedule))
    assert r.flags & abi.STANDARD_WEIGHTS_FLAG
    for f in range(3):
        a, c = run_frame_and_compare(r, scenes.orbit_camera(sc.camera, f))
    lit = c["w"] > 0
    # the invariant a later pass relies on
    np.testing.assert_array_equal(c["sumWeights"][lit],
                                  (c["w"][lit] * c["M"][lit].astype(np.float32) * c["pHat"][lit]).astype(np.float32))
    assert not np.isnan(c["w"]).any()


@pytest.mark.parametrize("schedule", [abi.SCHEDULE_QUEUED, abi.SCHEDULE_INLINE])
def test_biased_cornell_with_triangle_lights_bit_exact(torch, schedule):
    import dataclasses
    sc = scenes.cornell_box(1)
    tri = sc.tri_xyz.reshape(-1, 3, 3)
    ceiling = np.flatnonzero(np.all(np.abs(tri[:, :, 1] - tri[:, :, 1].max()) < 1e-6, axis=1))[:2]
    t = tri[ceiling].copy()
    t[:, :, 1] -= 0.01
    sc = dataclasses.replace(sc, tri_lights=scenes.make_tri_lights(t.reshape(-1, 9), np.full((2, 3), 3.0)))
    r = Renderer(sc, 128, 96, mode="biased", spatial_passes=1, spatial_neighbors=5,
                 launch=abi.Launch(16, 16, schedule))
    assert r.dev.union
    run_frame_and_compare(r, sc.camera)
    run_frame_and_compare(r, sc.camera)


def test_shade_visibility_retest_bit_exact(torch):
    """The final shading's visibility re-test (restir_shade_ex with RESTIR_SHADE_VISIBILITY_FLAG):
    after biased Pass B frames on the terrain it equals the oracle's re-tested shading bit for bit
    and darkens the pixels whose reused sample is shadowed; after an unbiased Pass C it changes
    no value (a shadowed centre already has W = 0)."""
    sc = scenes.config_scene("terrain")
    for mode, kw in (("biased", dict(spatial_passes=1, spatial_neighbors=5)), ("unbiased", dict(k=5, spatial_passes=2))):
        r = Renderer(sc, 320, 180, mode=mode, **kw)
        for f in range(2):
            cam = scenes.orbit_camera(sc.camera, f)
            _, final_ref = run_frame_and_compare(r, cam)
            plain = r.image().copy()
            r.shade_visibility = True
            r.shade()
            vis = r.image()
            r.shade_visibility = False
            og = oracle.gbuffer(host_planes(r, r.cur), r.W, r.H)
            ref = oracle.shade(oracle_for(r), cam["position"], 1.0, og, final_ref, visibility=True)
            eq = floats_equal_nan(vis, ref)
            assert eq.all(), f"{mode} frame {f}: {int((~eq).sum())} values differ"
            changed = int((~floats_equal_nan(vis, plain)).sum())
            print(f"[shade visibility] {mode} frame {f}: {changed} values darkened", flush=True)
            if mode == "unbiased":
                assert changed == 0
            else:
                assert changed > 0 and np.all((vis <= plain) | np.isnan(plain))


def test_trace_shadow_matches_oracle(torch):
    for sc in (scenes.cornell_box(1), scenes.terrain(2)):
        hs = HostScene(sc)
        r = Renderer(sc, 128, 128, ctx=None)
        lo, hi = sc.aabb()
        p1, p2 = random_segments(np.random.default_rng(5), lo, hi, 50000)
        n = len(p1)
        d1 = torch.from_numpy(p1).cuda()
        d2 = torch.from_numpy(p2).cuda()
        sh = torch.zeros(n, dtype=torch.int32, device="cuda")
        vis = torch.zeros(n, dtype=torch.int32, device="cuda")
        tst = torch.zeros(n, dtype=torch.int32, device="cuda")
        r.ctx.trace_shadow(d1, d2, n, sh, vis, tst, reference_loop=True)
        ref_sh, ref_v, _ = oracle.trace_shadow(hs.oracle, p1, p2)
        np.testing.assert_array_equal(sh.cpu().numpy().astype(np.uint32), ref_sh)
        np.testing.assert_array_equal(vis.cpu().numpy().astype(np.uint32), ref_v)
        # production loop (first hit ends the ray): same hit/miss
        sh2 = torch.zeros(n, dtype=torch.int32, device="cuda")
        r.ctx.trace_shadow(d1, d2, n, sh2)
        np.testing.assert_array_equal(sh2.cpu().numpy().astype(np.uint32), ref_sh)
        assert 0.05 < ref_sh.mean() < 0.95


def test_trace_shadow_on_pass_a_rays(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 256, 256, temporal=False)
    r.render_frame(sc.camera)
    planes = host_planes(r, r.cur)
    # Pass A output before visibility is not exposed; use the final reservoirs' light positions
    res = r.reservoirs()
    p1, p2 = segments_from_gbuffer(planes, res, 256, 256)
    hs = HostScene(sc)
    n = len(p1)
    d1, d2 = torch.from_numpy(p1).cuda(), torch.from_numpy(p2).cuda()
    sh = torch.zeros(n, dtype=torch.int32, device="cuda")
    vis = torch.zeros(n, dtype=torch.int32, device="cuda")
    r.ctx.trace_shadow(d1, d2, n, sh, vis, reference_loop=True)
    ref_sh, ref_v, _ = oracle.trace_shadow(hs.oracle, p1, p2)
    np.testing.assert_array_equal(sh.cpu().numpy().astype(np.uint32), ref_sh)
    np.testing.assert_array_equal(vis.cpu().numpy().astype(np.uint32), ref_v)


def test_abi_errors_on_device(torch):
    sc = scenes.cornell_box(1)
    r = Renderer(sc, 64, 64)
    g = r.gstruct(0)
    u, _ = r.uniforms(sc.camera)
    with pytest.raises(abi.RestirError) as e:
        r.ctx.spatial(u, 0, g, r.res[0], r.res[0])
    assert e.value.code == -6
    u.screenSize[0] = 65
    with pytest.raises(abi.RestirError) as e:
        r.ctx.initial_temporal(u, g, g, r.res[1], r.res[0])
    assert e.value.code == -2
    u.screenSize[0] = 64
    with pytest.raises(abi.RestirError) as e:
        r.ctx.unbiased(u, 17, 0, g, r.res[0], r.res[1])
    assert e.value.code == -1
    bad = r.dev.nodes.copy()
    bad[0]["left"] = 10**6
    d_bad = torch.from_numpy(bad.view(np.uint8).copy()).cuda()
    with pytest.raises(abi.RestirError) as e:
        r.ctx.set_bvh(d_bad, len(bad), r.dev.d_tris, len(r.dev.tris))
    assert e.value.code == -5
    with pytest.raises(abi.RestirError) as e:
        r.ctx.set_bvh(r.dev.d_nodes, len(bad) - 1, r.dev.d_tris, len(r.dev.tris))
    assert e.value.code == -5
