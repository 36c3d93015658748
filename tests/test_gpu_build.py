"""Device scene builders against the host builders of the reference's algorithms.

restir_build_bvh must produce the tree of restir_build_aabb_tree (AabbTree::build,
aabbTreeBuilder.cpp:52-214; itself checked against the oracle's restatement) byte for byte: same
breadth-first node ids, boxes, leaf order inside each node and depth. The cases cover the places
where a parallel build could drift from the reference's sequential one: ties between equal
coordinates (the fold keeps the last one in primitive order, and +0/-0 differ in bits),
duplicated and coincident triangles (the sweep puts everything on one side and the range is
halved), ranges that split off one primitive at a time (deep trees), and the smallest inputs."""
import numpy as np
import pytest

from restir_amd import abi, scenes
from restir_amd.pipeline import Renderer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


@pytest.fixture(scope="module")
def ctx(torch):
    return abi.Context(0)


def _signed_zero_plane(rng):
    g = rng.integers(-4, 5, size=(800, 9)).astype(np.float32)
    g[:, 2::3] = np.where(rng.random((800, 3)) < 0.5, -0.0, 0.0)
    return g


def _signed_zero_heavy(rng):
    g = (rng.integers(-2, 3, size=(500, 9)) * 0.0).astype(np.float32)
    g *= np.where(rng.random((500, 9)) < 0.5, -1, 1).astype(np.float32)
    g[:, 0] += np.arange(500, dtype=np.float32) % 7
    return g


def _exponential_chain():
    x = (2.0 ** (np.arange(60) / 2)).astype(np.float32)
    t = np.zeros((60, 9), np.float32)
    t[:, 0], t[:, 3], t[:, 6], t[:, 7] = x, x + 1, x, 1
    return t


def _hand_worked_swap_partition():
    """The 6-triangle tree of tests/golden/bvh_swap_partition_hand.json (worked by hand; its leaf
    order shows the reference's unstable swap partition)."""
    import json
    from pathlib import Path
    g = json.loads((Path(__file__).resolve().parent / "golden" / "bvh_swap_partition_hand.json").read_text())
    return np.array([[cx - 0.5, 0, 0, cx + 0.5, 1, 0, cx, 0, 1] for cx in g["centroid_x"]], dtype=np.float32)


CASES = {
    "hand_worked_swap_partition": lambda rng: _hand_worked_swap_partition(),
    "two": lambda rng: rng.standard_normal((2, 9)).astype(np.float32),
    "three": lambda rng: rng.standard_normal((3, 9)).astype(np.float32),
    "soup_17": lambda rng: rng.standard_normal((17, 9)).astype(np.float32),
    "soup_5000": lambda rng: rng.standard_normal((5000, 9)).astype(np.float32),
    "duplicates": lambda rng: np.repeat(rng.standard_normal((20, 9)).astype(np.float32), 50, axis=0),
    "coincident": lambda rng: np.tile(rng.standard_normal((1, 9)).astype(np.float32), (64, 1)),
    "signed_zero_plane": _signed_zero_plane,
    "signed_zero_heavy": _signed_zero_heavy,
    "exponential_chain": lambda rng: _exponential_chain(),
    "cornell": lambda rng: scenes.cornell_box(1).tri_xyz,
    "terrain_small": lambda rng: scenes.terrain(5, grid=64, n_boxes=100, n_point_lights=10).tri_xyz,
}


def _device_build(torch, ctx, tri):
    n = len(tri)
    x = torch.from_numpy(np.ascontiguousarray(tri, dtype=np.float32)).cuda()
    nodes = torch.full((abi.NODE_DTYPE.itemsize * (n - 1),), 0xAB, dtype=torch.uint8, device="cuda")
    tris = torch.full((abi.TRIANGLE_DTYPE.itemsize * n,), 0xAB, dtype=torch.uint8, device="cuda")
    depth = ctx.build_bvh(x, n, nodes, tris)
    torch.cuda.synchronize()
    return nodes.cpu().numpy(), tris.cpu().numpy(), depth


def _assert_same_tree(torch, ctx, tri, name):
    print(f"[build] {name}: {len(tri)} triangles", flush=True)
    hn, ht, hd = abi.build_aabb_tree(tri)
    dn, dt, dd = _device_build(torch, ctx, tri)
    assert dd == hd, f"{name}: depth {dd} != host {hd}"
    assert np.array_equal(dt, ht.view(np.uint8)), f"{name}: triangle buffer differs"
    if not np.array_equal(dn, hn.view(np.uint8)):
        d = dn.view(abi.NODE_DTYPE)
        bad = np.flatnonzero(np.any(d.view(np.uint8).reshape(len(d), -1) != hn.view(np.uint8).reshape(len(hn), -1), axis=1))
        i = int(bad[0])
        raise AssertionError(f"{name}: {len(bad)} of {len(hn)} nodes differ; first {i}: device {d[i]} host {hn[i]}")


@pytest.mark.parametrize("case", list(CASES))
def test_device_tree_matches_host_build(torch, ctx, case):
    _assert_same_tree(torch, ctx, CASES[case](np.random.default_rng(3)), case)


def test_device_tree_matches_host_build_full_size_terrain(torch, ctx):
    """The 1080p workload's scene (terrain 256x256 + 500 boxes + 2,000 lamps, 138,050 triangles)."""
    _assert_same_tree(torch, ctx, scenes.config_scene("terrain").tri_xyz, "terrain")


def test_device_tree_large_soup_and_reuse(torch, ctx):
    """A 400k-triangle soup (several scan groups per level, long pointer-jumping chains), then a
    smaller build on the same context: the scratch of the first build must not leak into it."""
    rng = np.random.default_rng(9)
    big = (rng.random((400_000, 9), dtype=np.float32) * 100.0).astype(np.float32)
    big[:, 3:] = big[:, :3].repeat(2, axis=0).reshape(-1, 6) + rng.random((400_000, 6), dtype=np.float32)
    _assert_same_tree(torch, ctx, big, "soup_400k")
    _assert_same_tree(torch, ctx, CASES["soup_5000"](np.random.default_rng(4)), "soup_5000_after")


def test_device_tree_rejects_bad_input(torch, ctx):
    tri = np.random.default_rng(1).standard_normal((10, 9)).astype(np.float32)
    tri[4, 5] = np.inf
    with pytest.raises(abi.RestirError):
        _device_build(torch, ctx, tri)
    x = torch.zeros(9, dtype=torch.float32, device="cuda")
    with pytest.raises(abi.RestirError):
        ctx.build_bvh(x, 1, torch.empty(80, dtype=torch.uint8, device="cuda"),
                      torch.empty(48, dtype=torch.uint8, device="cuda"))


def test_frame_on_device_built_scene_equals_host_built(torch):
    """Renderer with the tree built in HBM: every reservoir of two unbiased frames equals the
    host-built scene's, bit for bit."""
    sc = scenes.terrain(11, grid=48, n_boxes=60, n_point_lights=2000, extent=16.0)
    out = []
    for build in ("host", "device"):
        r = Renderer(sc, 160, 96, mode="unbiased", k=3, spatial_passes=1, scene_build=build)
        for f in range(2):
            r.render_frame(scenes.orbit_camera(sc.camera, f))
        torch.cuda.synchronize()
        out.append(r.res[r.cur].cpu().numpy().copy())
    assert np.array_equal(out[0], out[1])


# ---- alias table: restir_build_alias_table_device against createAliasTable's host restatement.
# The "dominant" case (one light paired with every small in turn) once faulted the GPU: the kernel
# lacked a barrier between building its queues and the first pairing step, so lane 0 could read
# a queue entry before its writer stored it (found by the kernel body's CPU build under
# ThreadSanitizer and AddressSanitizer, tests/test_alias_kernel_cpu.py).
def _points(lum):
    pts = np.zeros(len(lum), dtype=abi.POINT_LIGHT_DTYPE)
    pts["color_lum"][:, 3] = np.asarray(lum, dtype=np.float32)
    pts["color_lum"][:, :3] = 1.0
    return pts


ALIAS_CASES = {
    "one": lambda rng: _points([2.5]),
    "two": lambda rng: _points([1.0, 3.0]),
    "equal": lambda rng: _points(np.ones(1000)),
    "uniform_10k": lambda rng: _points(rng.uniform(0.1, 2.0, 10_000)),
    "dominant": lambda rng: _points(np.r_[np.ones(5000), 4.0e4]),
    "with_zeros": lambda rng: _points(np.where(rng.random(3000) < 0.3, 0.0, rng.exponential(1.0, 3000))),
    "lognormal_100k": lambda rng: _points(rng.lognormal(0.0, 2.0, 100_000)),
    "terrain_lights": lambda rng: scenes.config_scene("terrain").point_lights,
}


def _device_alias(torch, ctx, points=None, tri_lights=None):
    pts = None if points is None else torch.from_numpy(np.ascontiguousarray(points).view(np.uint8).copy()).cuda()
    tls = None if tri_lights is None else torch.from_numpy(np.ascontiguousarray(tri_lights).view(np.uint8).copy()).cuda()
    n = len(points) if points is not None else len(tri_lights)
    cols = torch.full((16 * n,), 0xCD, dtype=torch.uint8, device="cuda")
    ctx.build_alias_table(pts, 0 if points is None else n, tls, 0 if tri_lights is None else n, cols)
    torch.cuda.synchronize()
    return cols.cpu().numpy()


@pytest.mark.parametrize("case", list(ALIAS_CASES))
def test_device_alias_table_matches_host_build(torch, ctx, case):
    pts = ALIAS_CASES[case](np.random.default_rng(5))
    print(f"[alias] {case}: {len(pts)} point lights", flush=True)
    host = abi.build_alias_table(pts, None)
    dev = _device_alias(torch, ctx, points=pts)
    if not np.array_equal(dev, host.view(np.uint8)):
        dw, hw = dev.view(np.uint32).reshape(-1, 4), host.view(np.uint32).reshape(-1, 4)
        bad = np.flatnonzero(np.any(dw != hw, axis=1))
        i = int(bad[0])
        raise AssertionError(f"{case}: {len(bad)} of {len(host)} columns differ; first {i}: "
                             f"device {dev.view(abi.ALIAS_DTYPE)[i]} host {host[i]}")


def test_device_alias_table_triangle_lights(torch, ctx):
    rng = np.random.default_rng(8)
    tri = rng.standard_normal((700, 9)).astype(np.float32)
    tls = scenes.make_tri_lights(tri, rng.uniform(0.0, 3.0, (700, 3)))
    host = abi.build_alias_table(None, tls)
    assert np.array_equal(_device_alias(torch, ctx, tri_lights=tls), host.view(np.uint8))


def test_device_alias_build_rejects_bad_arguments(torch, ctx):
    """NULL buffers and columns that overlap the light records are refused before any launch."""
    pts = torch.zeros(32 * 4, dtype=torch.uint8, device="cuda")
    cols = torch.zeros(16 * 4, dtype=torch.uint8, device="cuda")
    with pytest.raises(abi.RestirError):
        ctx.build_alias_table(None, 4, None, 0, cols)
    with pytest.raises(abi.RestirError):
        ctx.build_alias_table(pts, 4, None, 0, pts)