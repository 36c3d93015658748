"""Shared helpers for the ReSTIR tests: scene preparation on the host, oracle wiring, and
NaN-aware bitwise comparison of reservoir buffers."""
from __future__ import annotations

import numpy as np

import oracle
from restir_amd import abi, scenes

# word layout of one 64-byte reservoir (restirStructs.glsl:19-34)
INT_WORDS = (8, 12, 13, 14, 15)  # lightIndex, numStreamSamples, padding


def res_words(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint32).reshape(-1, 16)


def reservoir_mismatch(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-pixel mask of reservoirs that differ bitwise, treating two NaNs in a float field as
    equal (x86 and gfx950 produce different NaN payloads for 0/0)."""
    wa, wb = res_words(a), res_words(b)
    neq = wa != wb
    fa, fb = wa.view(np.float32), wb.view(np.float32)
    both_nan = np.isnan(fa) & np.isnan(fb)
    both_nan[:, list(INT_WORDS)] = False
    return (neq & ~both_nan).any(axis=1)


def floats_equal_nan(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


class HostScene:
    """A procedural scene with its BVH and alias table built by the product's host builders,
    plus the matching oracle scene."""

    def __init__(self, scene: scenes.Scene):
        self.scene = scene
        self.nodes, self.tris, self.depth = abi.build_aabb_tree(scene.tri_xyz)
        self.alias = abi.scene_alias_table(scene.point_lights, scene.tri_lights)
        pts = scene.point_lights if len(scene.point_lights) else None
        tls = scene.tri_lights if len(scene.tri_lights) else None
        self.oracle = oracle.Scene(self.nodes, self.tris, pts, tls, self.alias)

    def camera(self, cam=None):
        return abi.camera_struct(cam or self.scene.camera)

    def oracle_gbuffer(self, width, height, cam=None):
        return oracle.primary_gbuffer(self.tris, self.camera(cam), self.scene.tri_material,
                                      self.scene.materials, width, height)


def segments_from_gbuffer(planes: np.ndarray, res: np.ndarray, width: int, height: int):
    """Shadow segments surface point -> selected light of every lit pixel (float4 each)."""
    P = width * height
    pos = planes[: 3 * P].reshape(3, P).T
    lit = res["sumWeights"] != 0
    p1 = np.zeros((int(lit.sum()), 4), dtype=np.float32)
    p2 = np.zeros_like(p1)
    p1[:, :3] = pos[lit]
    p2[:, :3] = res["pos_lum"][lit, :3]
    return p1, p2


def random_segments(rng, lo, hi, n):
    p1 = np.zeros((n, 4), dtype=np.float32)
    p2 = np.zeros((n, 4), dtype=np.float32)
    p1[:, :3] = rng.uniform(lo, hi, size=(n, 3))
    p2[:, :3] = rng.uniform(lo, hi, size=(n, 3))
    return p1, p2
