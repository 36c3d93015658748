"""Light clusters of the learned proposal: the product's host builder against the oracle (bit for
bit), and k-means / alias-table invariants that pin the oracle."""
import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes


def _lights(n, seed):
    rng = np.random.default_rng(seed)
    return scenes.make_point_lights(rng.uniform(-10, 10, (n, 3)),
                                    np.exp(rng.uniform(np.log(0.01), np.log(10), (n, 3))))


@pytest.mark.parametrize("n,k", [(64, 8), (10000, 256), (3000, 3000)])
def test_clusters_match_oracle(n, k):
    pts = _lights(n, n)
    a = abi.LightClusters(pts, k, seed=7, iterations=10)
    b = oracle.build_light_clusters(pts, k, 7, 10, abi.ALIAS_DTYPE)
    np.testing.assert_array_equal(a.light_cluster, b["light_cluster"])
    np.testing.assert_array_equal(a.begin, b["begin"])
    np.testing.assert_array_equal(a.count, b["count"])
    np.testing.assert_array_equal(a.lights, b["lights"])
    assert a.alias.tobytes() == b["alias"].tobytes()
    assert a.light_pwithin.tobytes() == b["light_pwithin"].tobytes()
    assert a.centroids.tobytes() == b["centroids"].tobytes()


def test_kmeans_invariants():
    pts = _lights(5000, 1)
    k = 64
    o = oracle.build_light_clusters(pts, k, 3, 12, abi.ALIAS_DTYPE)
    pos = pts["pos"][:, :3].astype(np.float64)
    c = o["centroids"].astype(np.float64)
    # every light sits in its nearest centroid's cluster
    d = ((pos[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    best = d.min(axis=1)
    mine = d[np.arange(len(pos)), o["light_cluster"]]
    assert np.all(mine <= best * (1 + 1e-5) + 1e-9)
    # the grouping is a partition listed by cluster in index order
    assert o["count"].sum() == len(pts)
    assert np.all(o["begin"] == np.concatenate([[0], np.cumsum(o["count"])[:-1]]))
    for j in range(k):
        members = o["lights"][o["begin"][j]: o["begin"][j] + o["count"][j]]
        assert np.all(np.diff(members) > 0)
        assert np.all(o["light_cluster"][members] == j)
        if len(members):
            lum = pts["color_lum"][members, 3].astype(np.float64)
            np.testing.assert_allclose(o["light_pwithin"][members], lum / lum.sum(), rtol=1e-4)
            al = o["alias"][o["begin"][j]: o["begin"][j] + o["count"][j]]
            assert np.all((al["alias"] >= 0) & (al["alias"] < len(members)))
            recon = al["prob"].astype(np.float64).copy()
            np.add.at(recon, al["alias"], 1.0 - al["prob"])
            recon -= np.where(al["alias"] == np.arange(len(members)), 1.0 - al["prob"], 0.0)
            np.testing.assert_allclose(recon / len(members), al["oriProb"], rtol=2e-3, atol=1e-7)
    # Lloyd steps do not increase the k-means objective
    prev = None
    for it in (0, 1, 4, 12):
        oo = oracle.build_light_clusters(pts, k, 3, it, abi.ALIAS_DTYPE)
        cc = oo["centroids"].astype(np.float64)
        obj = ((pos - cc[oo["light_cluster"]]) ** 2).sum()
        if prev is not None:
            assert obj <= prev * (1 + 1e-6)
        prev = obj
