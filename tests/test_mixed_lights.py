"""The union light space: one alias table over point lights followed by triangle lights.

The reference's createAliasTable covers point lights if there are any, else triangle lights
(misc.cpp:426-442, restirOmni.glsl:117); a scene with both kinds (the 1080p workload's 10k point
lights and 2k emissive triangles) needs one table over both. These checks pin that extension:
the library's host builder equals the oracle's, it reduces to the reference's table when one kind
is absent, it reconstructs the union distribution, and the oracle's Pass A decodes a union index
into the right light (a triangle candidate's sample lies on the triangle its index names)."""
import math

import numpy as np
import pytest

import oracle
from restir_amd import abi, scenes
from rt_helpers import HostScene


def _lights(n_point, n_tri, seed):
    rng = np.random.default_rng(seed)
    rgb = np.exp(rng.uniform(math.log(0.01), math.log(10.0), size=(n_point, 3)))
    pts = scenes.make_point_lights(rng.uniform(-1, 1, (n_point, 3)), rgb)
    tri = rng.uniform(-1, 1, (n_tri, 9)).astype(np.float32)
    tls = scenes.make_tri_lights(tri, np.exp(rng.uniform(math.log(0.01), math.log(10.0), size=(n_tri, 3))))
    return pts, tls


@pytest.mark.parametrize("n_point,n_tri,seed", [(1, 1, 0), (64, 2, 1), (1000, 300, 2), (10000, 2000, 3)])
def test_union_table_equals_oracle(n_point, n_tri, seed):
    pts, tls = _lights(n_point, n_tri, seed)
    t = abi.build_alias_table_mixed(pts, tls)
    ref = oracle.build_alias_table_mixed(pts, tls, abi.ALIAS_DTYPE)
    assert np.array_equal(t.view(np.uint8), ref.view(np.uint8))


def test_union_table_reduces_to_the_reference_table():
    pts, tls = _lights(500, 200, 4)
    only_points = abi.build_alias_table_mixed(pts, tls[:0])
    assert np.array_equal(only_points.view(np.uint8), abi.build_alias_table(pts, None).view(np.uint8))
    only_tris = abi.build_alias_table_mixed(pts[:0], tls)
    assert np.array_equal(only_tris.view(np.uint8), abi.build_alias_table(None, tls).view(np.uint8))


def test_union_table_reconstructs_union_distribution():
    pts, tls = _lights(3000, 1000, 5)
    t = abi.build_alias_table_mixed(pts, tls)
    n = len(t)
    recon = t["prob"].astype(np.float64).copy()
    np.add.at(recon, t["alias"], 1.0 - t["prob"].astype(np.float64))
    recon -= np.where(t["alias"] == np.arange(n), 1.0 - t["prob"], 0.0)
    w = np.concatenate([pts["color_lum"][:, 3].astype(np.float64),
                        tls["emission_lum"][:, 3].astype(np.float64) * tls["normal_area"][:, 3]])
    np.testing.assert_allclose(recon / n, w / w.sum(), rtol=3e-3, atol=1e-8)
    np.testing.assert_allclose(t["oriProb"], w / w.sum(), rtol=1e-4)
    assert np.all(t["aliasOriProb"] == t["oriProb"][t["alias"]])


def test_scene_alias_table_rule():
    pts, tls = _lights(10, 4, 6)
    assert abi.uses_union(pts, tls) and len(abi.scene_alias_table(pts, tls)) == 14
    assert not abi.uses_union(pts, tls, union=False) and len(abi.scene_alias_table(pts, tls, union=False)) == 10
    assert not abi.uses_union(pts, tls[:0]) and len(abi.scene_alias_table(pts, tls[:0])) == 10
    assert not abi.uses_union(pts[:0], tls) and len(abi.scene_alias_table(pts[:0], tls)) == 4


def _cornell_with_triangle_lights(point_scale):
    """The Cornell box's 64 point lights (scaled) plus its two ceiling triangles, lowered a little,
    as emissive triangle lights."""
    import dataclasses
    sc = scenes.cornell_box(1)
    tri = sc.tri_xyz.reshape(-1, 3, 3)
    ceiling = np.flatnonzero(np.all(np.abs(tri[:, :, 1] - tri[:, :, 1].max()) < 1e-6, axis=1))[:2]
    t = tri[ceiling].copy()
    t[:, :, 1] -= 0.01
    tl = scenes.make_tri_lights(t.reshape(-1, 9), np.array([[2.0, 1.8, 1.5], [1.5, 1.8, 2.0]]))
    pts = sc.point_lights.copy()
    pts["color_lum"] *= np.float32(point_scale)
    return dataclasses.replace(sc, point_lights=pts, tri_lights=tl)


def _oracle_pass_a(sc, W=48, H=40, M=8):
    hs = HostScene(sc)
    planes = hs.oracle_gbuffer(W, H)
    g = oracle.gbuffer(planes, W, H)
    u = abi.make_uniforms(W, H, 1, sc.camera["position"], M=M, flags=0)
    return oracle.pass_a(hs.oracle, u, g, None, None, abi.RESERVOIR_DTYPE), planes, hs


def test_oracle_pass_a_decodes_union_indices():
    """Almost all power on the two triangle lights: nearly every selected sample is a triangle
    light, and each lies on the triangle its index names (an index left in the union space, or
    shifted the wrong way, names a point light or the other triangle)."""
    sc = _cornell_with_triangle_lights(1e-6)
    res, planes, hs = _oracle_pass_a(sc)
    assert len(hs.alias) == len(sc.point_lights) + len(sc.tri_lights)
    sel = res["sumWeights"] > 0
    li = res["lightIndex"][sel]
    assert sel.sum() > 500 and (li < 0).mean() > 0.99
    for r in res[sel][li < 0]:
        t = sc.tri_lights[-1 - r["lightIndex"]]
        a, b, c = (t[k][:3].astype(np.float64) for k in ("p1", "p2", "p3"))
        p = r["pos_lum"][:3].astype(np.float64)
        # barycentric coordinates of the sample in the named triangle
        m = np.stack([b - a, c - a], axis=1)
        x, *_ = np.linalg.lstsq(m, p - a, rcond=None)
        assert np.abs(m @ x - (p - a)).max() < 1e-5
        assert x.min() > -1e-5 and x.sum() < 1 + 1e-5
        assert np.allclose(r["normal"][:3], t["normal_area"][:3]) and r["normal"][3] == 1.0


def test_oracle_pass_a_union_samples_both_kinds():
    """Both kinds at their own power (the two triangles hold about 5% of it): both get selected,
    and point lights keep their own indices."""
    sc = _cornell_with_triangle_lights(1.0)
    res, _, _ = _oracle_pass_a(sc)
    li = res["lightIndex"][res["sumWeights"] > 0]
    assert (li < 0).mean() > 0.005 and (li >= 0).mean() > 0.5
    assert li[li >= 0].max() < len(sc.point_lights)
    assert (-1 - li[li < 0]).max() < len(sc.tri_lights)
