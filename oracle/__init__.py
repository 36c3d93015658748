"""Python binding of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this module, and
only as the checker. It reuses the wire-format dtypes of the product binding so inputs can be
shared byte for byte, but calls nothing in the product library.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"

_vp, _u32, _u64, _i32, _f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float


class OracleScene(C.Structure):
    _fields_ = [("nodes", _vp), ("n_nodes", _u32), ("tris", _vp), ("n_tris", _u32),
                ("points", _vp), ("n_points", _u32), ("tri_lights", _vp), ("n_tri_lights", _u32),
                ("alias", _vp), ("n_alias", _u32)]


SIGS = {
    "oracle_pcg32_stream": (None, [C.c_uint64, C.c_uint64, _u32, _vp]),
    "oracle_pcg32_float_stream": (None, [C.c_uint64, C.c_uint64, _u32, _vp]),
    "oracle_sincos": (None, [_f32, C.POINTER(_f32), C.POINTER(_f32)]),
    "oracle_luminance": (_f32, [_f32, _f32, _f32]),
    "oracle_brdf_luminance": (_f32, [_f32, _f32, _f32, _f32, _f32, _f32, _f32]),
    "oracle_phat": (_f32, [_vp, _vp, _vp, _vp, _vp, _i32, _f32, _f32, _f32, _f32]),
    "oracle_phat_f64": (C.c_double, [_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                     C.c_double]),
    "oracle_add_sample": (None, [_vp, _vp, _vp, _f32, C.c_int32, _f32, _f32, _vp]),
    "oracle_combine": (None, [_vp, _vp, _f32, _vp]),
    "oracle_build_alias_table": (_i32, [_vp, _u32, _vp, _u32, _vp]),
    "oracle_build_alias_table_mixed": (_i32, [_vp, _u32, _vp, _u32, _vp]),
    "oracle_build_aabb_tree": (_i32, [_vp, _u32, _vp, _vp, C.POINTER(_u32)]),
    "oracle_build_light_clusters": (_i32, [_vp, _u32, _u32, C.c_uint64, _u32, _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp]),
    "oracle_ray_aabb": (_i32, [_vp, _vp, _vp, _vp]),
    "oracle_ray_triangle": (_i32, [_vp, _vp, _vp]),
    "oracle_trace_shadow": (None, [C.POINTER(OracleScene), _vp, _vp, _u32, _vp, _vp, _vp]),
    "oracle_trace_brute_f64": (None, [_vp, _u32, _vp, _vp, _u32, _vp, _vp]),
    "oracle_count_reachable": (None, [C.POINTER(OracleScene), _vp, _vp, _u32, _vp]),
    "oracle_pass_a": (None, [C.POINTER(OracleScene), _vp, _vp, _vp, _vp, _vp]),
    "oracle_pass_a_candidates": (None, [C.POINTER(OracleScene), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "oracle_pass_b": (None, [C.POINTER(OracleScene), _vp, _i32, _vp, _vp, _vp]),
    "oracle_pass_c": (None, [C.POINTER(OracleScene), _vp, _i32, _i32, _vp, _vp, _vp]),
    "oracle_cos_threshold": (_f32, [_f32]),
    "oracle_shade": (None, [C.POINTER(OracleScene), _vp, _f32, _vp, _vp, _vp, _u64]),
    "oracle_shade_visibility": (None, [C.POINTER(OracleScene), _vp, _f32, _vp, _vp, _vp, _u64]),
    "oracle_primary_gbuffer": (None, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64,
                                      _u32, _u32]),
    "oracle_projection_view": (None, [_vp, _f32, _vp]),
    "oracle_ground_truth_f64": (None, [C.POINTER(OracleScene), _vp, _vp, _i32, _u32, _vp, _vp,
                                       _u64]),
    "oracle_mlp_features": (None, [_vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "oracle_mlp_forward": (None, [_vp, _vp, _u32, _vp, _u32, _i32, _vp]),
    "oracle_mlp_loss_grad": (C.c_double, [_vp, _vp, _u32, _vp, _vp, _vp, _u32, _i32, _vp, _vp]),
    "oracle_adam": (None, [_vp, _vp, _vp, _vp, C.c_size_t, C.c_double, C.c_double, C.c_double,
                           C.c_double, _u32]),
}

_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


# ----------------------------------------------------------------------------------- helpers
def pcg32(seed: int, seq: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_pcg32_stream(seed, seq, n, out.ctypes.data)
    return out


def pcg32_float(seed: int, seq: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float32)
    lib().oracle_pcg32_float_stream(seed, seq, n, out.ctypes.data)
    return out


def sincos(x: float):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincos(x, C.byref(s), C.byref(c))
    return s.value, c.value


def build_alias_table(points, tri_lights, dtype):
    n = len(points) if points is not None and len(points) else len(tri_lights)
    out = np.zeros(n, dtype=dtype)
    lib().oracle_build_alias_table(_p(points), 0 if points is None else len(points),
                                   _p(tri_lights), 0 if tri_lights is None else len(tri_lights),
                                   out.ctypes.data)
    return out


def build_alias_table_mixed(points, tri_lights, dtype):
    """One table over the point lights followed by the triangle lights (the union light space)."""
    n = len(points) + len(tri_lights)
    out = np.zeros(n, dtype=dtype)
    lib().oracle_build_alias_table_mixed(_p(points), len(points), _p(tri_lights), len(tri_lights),
                                         out.ctypes.data)
    return out


def build_aabb_tree(tri_xyz, node_dtype, tri_dtype):
    tri_xyz = np.ascontiguousarray(tri_xyz, dtype=np.float32).reshape(-1, 9)
    n = len(tri_xyz)
    nodes = np.zeros(n - 1, dtype=node_dtype)
    tris = np.zeros(n, dtype=tri_dtype)
    depth = C.c_uint32()
    rc = lib().oracle_build_aabb_tree(tri_xyz.ctypes.data, n, nodes.ctypes.data, tris.ctypes.data,
                                      C.byref(depth))
    if rc != 0:
        raise RuntimeError(f"oracle_build_aabb_tree failed ({rc})")
    return nodes, tris, depth.value


def build_light_clusters(points, k, seed, iterations, alias_dtype):
    n = len(points)
    out = dict(light_cluster=np.zeros(n, np.int32), light_pwithin=np.zeros(n, np.float32),
               begin=np.zeros(k, np.int32), count=np.zeros(k, np.int32),
               lights=np.zeros(n, np.int32), alias=np.zeros(n, dtype=alias_dtype),
               centroids=np.zeros((k, 3), np.float32))
    pts = np.ascontiguousarray(points)
    rc = lib().oracle_build_light_clusters(
        pts.ctypes.data, n, k, seed, iterations, out["light_cluster"].ctypes.data,
        out["light_pwithin"].ctypes.data, out["begin"].ctypes.data, out["count"].ctypes.data,
        out["lights"].ctypes.data, out["alias"].ctypes.data, out["centroids"].ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_build_light_clusters failed ({rc})")
    return out


class Scene:
    """Keeps the host arrays alive and exposes the oracle_scene struct."""

    def __init__(self, nodes, tris, points=None, tri_lights=None, alias=None):
        self.arrays = [np.ascontiguousarray(a) if a is not None else None
                       for a in (nodes, tris, points, tri_lights, alias)]
        nd, tr, pt, tl, al = self.arrays
        self.s = OracleScene(_p(nd), 0 if nd is None else len(nd), _p(tr), 0 if tr is None else len(tr),
                             _p(pt), 0 if pt is None else len(pt), _p(tl), 0 if tl is None else len(tl),
                             _p(al), 0 if al is None else len(al))

    def ref(self):
        return C.byref(self.s)


def gbuffer(planes: np.ndarray, width: int, height: int):
    """restir_gbuffer over a host float32 array of 13 planes (same order as the product's)."""
    from ctypes import Structure

    class G(Structure):
        _fields_ = [("position", _vp), ("normal", _vp), ("albedo", _vp), ("material", _vp),
                    ("depth", _vp), ("plane_stride", C.c_uint64), ("width", _u32), ("height", _u32),
                    ("normal_format", _u32), ("reserved", _u32)]

    P = width * height
    base = planes.ctypes.data
    g = G(base, base + 12 * P, base + 24 * P, base + 40 * P, base + 48 * P, P, width, height, 0, 0)
    g._keep = planes
    return g


def trace_shadow(scene: Scene, p1, p2, with_counts=True):
    n = len(p1)
    sh = np.zeros(n, dtype=np.uint32)
    v = np.zeros(n, dtype=np.uint32) if with_counts else None
    t = np.zeros(n, dtype=np.uint32) if with_counts else None
    lib().oracle_trace_shadow(scene.ref(), _p(p1), _p(p2), n, sh.ctypes.data, _p(v), _p(t))
    return sh, v, t


def trace_brute_f64(tris, p1, p2):
    n = len(p1)
    sh = np.zeros(n, dtype=np.uint32)
    e = np.zeros(n, dtype=np.float64)
    lib().oracle_trace_brute_f64(tris.ctypes.data, len(tris), _p(p1), _p(p2), n, sh.ctypes.data,
                                 e.ctypes.data)
    return sh, e


def count_reachable(scene: Scene, p1, p2):
    n = len(p1)
    c = np.zeros(n, dtype=np.uint32)
    lib().oracle_count_reachable(scene.ref(), _p(p1), _p(p2), n, c.ctypes.data)
    return c


def pass_a(scene: Scene, u, gcur, gprev, prev_res, res_dtype):
    W, H = u.screenSize[0], u.screenSize[1]
    out = np.zeros(W * H, dtype=res_dtype)
    lib().oracle_pass_a(scene.ref(), C.addressof(u), C.addressof(gcur),
                        C.addressof(gprev) if gprev is not None else None, _p(prev_res),
                        out.ctypes.data)
    return out


def pass_a_candidates(scene: Scene, u, gcur, gprev, prev_res, cand_light, cand_q, res_dtype):
    """Pass A over given candidates: cand_light/cand_q hold M x W*H entries (candidate-major)."""
    W, H = u.screenSize[0], u.screenSize[1]
    cl = np.ascontiguousarray(cand_light, dtype=np.int32).reshape(-1)
    cq = np.ascontiguousarray(cand_q, dtype=np.float32).reshape(-1)
    assert cl.size == cq.size == u.initialLightSampleCount * W * H
    out = np.zeros(W * H, dtype=res_dtype)
    lib().oracle_pass_a_candidates(scene.ref(), C.addressof(u), C.addressof(gcur),
                                   C.addressof(gprev) if gprev is not None else None, _p(prev_res),
                                   cl.ctypes.data, cq.ctypes.data, out.ctypes.data)
    return out


def pass_b(scene: Scene, u, it, g, res_in):
    out = np.zeros_like(res_in)
    lib().oracle_pass_b(scene.ref(), C.addressof(u), it, C.addressof(g), res_in.ctypes.data,
                        out.ctypes.data)
    return out


def pass_c(scene: Scene, u, k, it, g, res_in):
    out = np.zeros_like(res_in)
    lib().oracle_pass_c(scene.ref(), C.addressof(u), k, it, C.addressof(g), res_in.ctypes.data,
                        out.ctypes.data)
    return out


def shade(scene: Scene, cam_pos, gamma, g, res, visibility=False):
    W, H = g.width, g.height
    out = np.zeros(3 * W * H, dtype=np.float32)
    cp = np.asarray(cam_pos[:3], dtype=np.float32)
    fn = lib().oracle_shade_visibility if visibility else lib().oracle_shade
    fn(scene.ref(), cp.ctypes.data, gamma, C.addressof(g), res.ctypes.data, out.ctypes.data, W * H)
    return out.reshape(3, H, W)


def primary_gbuffer(tris, cam_struct, tri_material, materials, width, height):
    P = width * height
    planes = np.zeros(13 * P, dtype=np.float32)
    base = planes.ctypes.data
    lib().oracle_primary_gbuffer(tris.ctypes.data, len(tris), C.addressof(cam_struct),
                                 tri_material.ctypes.data, materials.ctypes.data, base,
                                 base + 12 * P, base + 24 * P, base + 40 * P, base + 48 * P, P,
                                 width, height)
    return planes


def projection_view(cam_struct, aspect):
    out = np.zeros(16, dtype=np.float32)
    lib().oracle_projection_view(C.addressof(cam_struct), aspect, out.ctypes.data)
    return out


def ground_truth_f64(scene: Scene, cam_pos, g, use_bvh=False, n_tri_samples=0, mask=None):
    W, H = g.width, g.height
    out = np.zeros(3 * W * H, dtype=np.float64)
    cp = np.asarray(cam_pos[:3], dtype=np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().oracle_ground_truth_f64(scene.ref(), cp.ctypes.data, C.addressof(g), int(use_bvh),
                                  n_tri_samples, _p(m), out.ctypes.data, W * H)
    return out.reshape(3, H, W)


# ------------------------------------------------------------------------- learned proposal
def mlp_features(g, cam_pos, lo, hi, pixels):
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    out = np.zeros((len(pixels), 128), dtype=np.float32)
    cp = np.asarray(cam_pos[:3], dtype=np.float32)
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    lib().oracle_mlp_features(C.addressof(g), cp.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                              pixels.ctypes.data, len(pixels), out.ctypes.data)
    return out


def mlp_forward(W, B, K, X, emulate_bf16=True):
    W = np.ascontiguousarray(W, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.zeros((len(X), K), dtype=np.float64)
    lib().oracle_mlp_forward(W.ctypes.data, B.ctypes.data, K, X.ctypes.data, len(X),
                             int(emulate_bf16), out.ctypes.data)
    return out


def mlp_loss_grad(W, B, K, X, target, mask, emulate_bf16=True):
    W = np.ascontiguousarray(W, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    target = np.ascontiguousarray(target, dtype=np.int32)
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    gW = np.zeros(W.shape, dtype=np.float64)
    gB = np.zeros(B.shape, dtype=np.float64)
    loss = lib().oracle_mlp_loss_grad(W.ctypes.data, B.ctypes.data, K, X.ctypes.data,
                                      target.ctypes.data, mask.ctypes.data, len(X),
                                      int(emulate_bf16), gW.ctypes.data, gB.ctypes.data)
    return loss, gW, gB


def adam(p, g, m, v, lr, b1, b2, eps, t):
    for a in (p, g, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().oracle_adam(p.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, p.size, lr, b1,
                      b2, eps, t)
