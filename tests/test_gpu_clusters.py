"""restir_build_light_clusters_device (the k-means light clustering of the learned proposal with
its nearest-centroid step on the GPU) against the host build: byte-identical outputs for the same
arguments, including duplicated light positions (ties go to the lower centroid) and cluster
counts that are not a multiple of the kernel's LDS chunk. At the 4K stress workload's size (1M
lights, 4,096 clusters), where the host build's nearest step would take minutes, the final
assignment is checked on a sample against an fp32 brute-force argmin over the returned
centroids, and the grouping and per-cluster tables for consistency."""
import time

import numpy as np
import pytest

from restir_amd import abi, scenes
from restir_amd.abi import LightClusters

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def random_lights(n, seed, dup=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-20, 20, size=(n, 3)).astype(np.float32)
    if dup:
        idx = rng.integers(0, n, size=dup)
        pos[idx] = pos[rng.integers(0, n, size=dup)]
    rgb = np.exp(rng.uniform(np.log(1e-3), np.log(10), size=(n, 3))).astype(np.float32)
    return scenes.make_point_lights(pos, rgb)


FIELDS = ("light_cluster", "light_pwithin", "begin", "count", "lights", "centroids")


@pytest.mark.parametrize("n,k,iters,dup", [(30000, 1000, 5, 0), (7777, 4096, 3, 0), (20000, 1500, 4, 5000),
                                           (5000, 1, 2, 0)])
def test_device_clustering_equals_host(torch, n, k, iters, dup):
    ctx = abi.Context(0)
    pts = random_lights(n, 17 + k, dup)
    h = LightClusters(pts, k, seed=5, iterations=iters)
    d = LightClusters(pts, k, seed=5, iterations=iters, ctx=ctx)
    for f in FIELDS:
        a, b = getattr(h, f), getattr(d, f)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f
    assert np.array_equal(h.alias.view(np.uint8), d.alias.view(np.uint8))


def test_device_clustering_1m_lights_4096_clusters(torch):
    n, k = 1_000_000, 4096
    ctx = abi.Context(0)
    pts = random_lights(n, 3)
    t0 = time.perf_counter()
    d = LightClusters(pts, k, seed=7, iterations=16, ctx=ctx)
    dt = time.perf_counter() - t0
    print(f"device clustering, {n} lights into {k} clusters, 16 iterations: {dt:.2f} s", flush=True)
    rng = np.random.default_rng(1)
    sample = rng.choice(n, size=4096, replace=False)
    p = pts["pos"][sample, :3]
    c = d.centroids
    dx = p[:, None, 0] - c[None, :, 0]
    dy = p[:, None, 1] - c[None, :, 1]
    dz = p[:, None, 2] - c[None, :, 2]
    dist = (dx * dx + dy * dy) + dz * dz  # fp32 elementwise, as the kernel and the host build
    assert dist.dtype == np.float32
    np.testing.assert_array_equal(d.light_cluster[sample], np.argmin(dist, axis=1))
    # grouping: every light once, in its cluster's range, index order within a cluster
    assert d.count.sum() == n and (d.count > 0).mean() > 0.99
    assert np.array_equal(np.sort(d.lights), np.arange(n))
    assert np.array_equal(np.repeat(np.arange(k), d.count), d.light_cluster[d.lights])
    starts = np.cumsum(np.concatenate([[0], d.count[:-1]]))
    assert np.array_equal(starts, d.begin)
    within = np.diff(d.lights)[np.repeat(np.arange(k), d.count)[1:] == np.repeat(np.arange(k), d.count)[:-1]]
    assert (within > 0).all()
    # p_within sums to one in every non-empty cluster
    s = np.bincount(d.light_cluster, weights=d.light_pwithin.astype(np.float64), minlength=k)
    assert np.allclose(s[d.count > 0], 1.0, atol=1e-4)


def test_device_clustering_rejects_bad_sizes(torch):
    ctx = abi.Context(0)
    pts = random_lights(100, 1)
    with pytest.raises(abi.RestirError):
        LightClusters(pts, 101, ctx=ctx)  # more clusters than lights
    with pytest.raises(abi.RestirError):
        LightClusters(pts[:0], 1, ctx=ctx)  # no lights


@pytest.mark.parametrize("K", [300, 4224])
def test_mlp_rejects_unsupported_cluster_counts(torch, K):
    """K must be 64, 128, 256 or a multiple of 128 up to 4,096."""
    from restir_amd.abi import MlpStruct, mlp_sizes
    ctx = abi.Context(0)
    nW, nB = mlp_sizes(K)
    bufs = [torch.zeros(n, dtype=dt, device="cuda") for n, dt in
            ((nW, torch.int16), (nW, torch.float32), (nB, torch.float32), (nW + nB, torch.float32),
             (nW + nB, torch.float32))]
    mlp = MlpStruct(K, 0, *[b.data_ptr() for b in bufs])
    prior = torch.zeros(K, dtype=torch.float32, device="cuda")
    with pytest.raises(abi.RestirError):
        ctx.mlp_init(mlp, 1, prior)
