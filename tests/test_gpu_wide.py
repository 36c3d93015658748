"""The production shadow walk over 4-entry nodes against the reference's binary walk.

restir_trace_shadow's first-hit loop (and every pass that traces a shadow ray) walks 4-entry nodes
built from the bound tree, two levels per visit; the reference loop walks the binary nodes
(softwareRaytracing.glsl:42-85). The verdicts must equal the oracle's binary walk
(oracle_trace_shadow) on every ray. The cases aim at the places where skipping the boxes between
visited levels could change the answer:

- rays with a zero direction component whose origin lies exactly on box planes: the slab
  distance 0/0 is NaN there, the reference's min/max chains then drop that slab, and the walk
  must test every box of such rays;
- a tree whose stored box for an internal node is shrunk so that it no longer contains the boxes
  below it (a tree the library did not build): the reference culls the whole subtree for rays
  that miss the shrunk box, the walk must detect the tree as not nested and test every box;
- the 1080p workload's scene on random segments, where the fast path decides almost every ray.
"""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from rt_helpers import HostScene, random_segments

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def _bind(torch, ctx, nodes, tris):
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).cuda()
    d_tris = torch.from_numpy(tris.view(np.uint8).copy()).cuda()
    ctx.set_bvh(d_nodes, len(nodes), d_tris, len(tris))
    return d_nodes, d_tris  # keep alive while the context uses them


def _trace(torch, ctx, p1, p2, **kw):
    n = len(p1)
    d1, d2 = torch.from_numpy(p1).cuda(), torch.from_numpy(p2).cuda()
    sh = torch.zeros(n, dtype=torch.int32, device="cuda")
    vis = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.trace_shadow(d1, d2, n, sh, vis, **kw)
    torch.cuda.synchronize()
    return sh.cpu().numpy().astype(np.uint32), vis.cpu().numpy()


def _check_all_loops(torch, ctx, osc, p1, p2, name):
    ref_sh, ref_v, _ = oracle.trace_shadow(osc, p1, p2)
    for kw in ({}, {"one_ray_per_lane": True}, {"reference_loop": True}):
        sh, vis = _trace(torch, ctx, p1, p2, **kw)
        bad = np.flatnonzero(sh != ref_sh)
        assert len(bad) == 0, f"{name} {kw}: {len(bad)} of {len(p1)} verdicts differ, first ray {bad[:4]}"
        if kw.get("reference_loop"):
            np.testing.assert_array_equal(vis.astype(np.uint32), ref_v)
    return ref_sh, ref_v


def _plane_segments(rng, nodes, lo, hi, n):
    """Segments with one or two direction components exactly zero whose fixed coordinates are
    box planes of the tree (so 0/0 occurs in the slab tests of those boxes)."""
    planes = [np.concatenate([nodes[k][:, a] for k in ("lmin", "lmax", "rmin", "rmax")]) for a in range(3)]
    p1, p2 = random_segments(rng, lo, hi, n)
    for i in range(n):
        axes = rng.choice(3, size=1 + (i % 2), replace=False)
        for a in axes:
            v = planes[a][rng.integers(len(planes[a]))]
            p1[i, a] = p2[i, a] = v
    return p1, p2


@pytest.mark.parametrize("name", ["cornell", "terrain_small"])
def test_wide_walk_random_and_plane_rays(torch, name):
    sc = scenes.cornell_box(1) if name == "cornell" else scenes.terrain(5, grid=64, n_boxes=100, n_point_lights=10)
    hs = HostScene(sc)
    ctx = abi.Context(0)
    keep = _bind(torch, ctx, hs.nodes, hs.tris)
    lo, hi = sc.aabb()
    rng = np.random.default_rng(11)
    p1, p2 = random_segments(rng, lo, hi, 40000)
    ref_sh, ref_v = _check_all_loops(torch, ctx, hs.oracle, p1, p2, f"{name} random")
    assert 0.05 < ref_sh.mean() < 0.95
    _, wide_v = _trace(torch, ctx, p1, p2)
    miss = ref_sh == 0
    # a miss walks every reachable box: two levels per visit take fewer visits than the binary walk
    assert wide_v[miss].mean() < 0.75 * ref_v[miss].mean()
    q1, q2 = _plane_segments(rng, hs.nodes, lo, hi, 40000)
    ref_sh, _ = _check_all_loops(torch, ctx, hs.oracle, q1, q2, f"{name} plane")
    assert 0.02 < ref_sh.mean() < 0.98
    del keep


def test_wide_walk_on_a_tree_that_is_not_nested(torch):
    """Shrink the box the tree stores for one internal subtree: the reference culls that subtree
    for rays missing the shrunk box even where they cross its triangles. Rays aimed through the
    subtree's original box must get the reference's (culled) answer."""
    sc = scenes.terrain(5, grid=64, n_boxes=100, n_point_lights=10)
    hs = HostScene(sc)
    nodes = hs.nodes.copy()
    # an internal child with an internal grandchild on both sides, a few levels down
    x = next(i for i in range(len(nodes)) if nodes[i]["left"] > 0 and nodes[nodes[i]["left"]]["left"] > 0
             and nodes[nodes[i]["left"]]["right"] > 0 and i > 6)
    c = int(nodes[x]["left"])
    lo0, hi0 = nodes[x]["lmin"][:3].copy(), nodes[x]["lmax"][:3].copy()
    ax = int(np.argmax(hi0 - lo0))
    nodes[x]["lmax"][ax] = lo0[ax] + np.float32(0.3) * (hi0[ax] - lo0[ax])  # keeps 30% of the extent
    osc = oracle.Scene(nodes, hs.tris, None, None, None)
    ctx = abi.Context(0)
    keep = _bind(torch, ctx, nodes, hs.tris)
    rng = np.random.default_rng(12)
    n = 30000
    p1 = np.zeros((n, 4), np.float32)
    p2 = np.zeros((n, 4), np.float32)
    slo, shi = sc.aabb()
    p1[:, :3] = rng.uniform(lo0, hi0, size=(n, 3))  # inside the subtree's original box
    p2[:, :3] = rng.uniform(slo, shi, size=(n, 3))
    ref_sh, _ = _check_all_loops(torch, ctx, osc, p1, p2, "shrunk box")
    # the shrink matters: the unmodified tree answers some of these rays differently
    orig_sh, _, _ = oracle.trace_shadow(hs.oracle, p1, p2)
    assert (orig_sh != ref_sh).sum() > 50, f"subtree {c}: shrinking changed too few rays"
    del keep


def test_wide_walk_full_size_terrain(torch):
    """The 1080p workload's scene (138,050 triangles), bound through the device build."""
    sc = scenes.config_scene("terrain")
    hs = HostScene(sc)
    ctx = abi.Context(0)
    n = len(hs.tris)
    x = torch.from_numpy(np.ascontiguousarray(sc.tri_xyz, dtype=np.float32)).cuda()
    d_nodes = torch.empty(abi.NODE_DTYPE.itemsize * (n - 1), dtype=torch.uint8, device="cuda")
    d_tris = torch.empty(abi.TRIANGLE_DTYPE.itemsize * n, dtype=torch.uint8, device="cuda")
    ctx.build_bvh(x, n, d_nodes, d_tris)
    lo, hi = sc.aabb()
    p1, p2 = random_segments(np.random.default_rng(13), lo, hi, 20000)
    ref_sh, _ = _check_all_loops(torch, ctx, hs.oracle, p1, p2, "terrain 1080p scene")
    assert 0.05 < ref_sh.mean() < 0.99
