"""ctypes binding of librestir_amd.so (the C ABI in include/restir_abi.h).

The structures below mirror the header byte for byte; their sizes are asserted at import. All
device pointers are plain integers (``tensor.data_ptr()``). Every entry point that returns a
negative code raises :class:`RestirError` carrying the library's message, which is how the
reference's ``vkCheck`` abort (src/misc.cpp:21-26) surfaces on this side.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "librestir_amd.so"

VISIBILITY_REUSE_FLAG = 1 << 0
TEMPORAL_REUSE_FLAG = 1 << 1
STANDARD_WEIGHTS_FLAG = 1 << 2  # RIS with W = sumW / (M p_hat) and triangle lights by area
INPUT_VISIBLE_FLAG = 1 << 3  # Pass C input from a visibility-tested Pass C: known segments are not retraced
SHADE_VISIBILITY_FLAG = 1  # RESTIR_SHADE_VISIBILITY_FLAG: final shading re-tests visibility
TRACE_FIRST_HIT = 0
TRACE_REFERENCE_LOOP = 1
TRACE_ONE_RAY_PER_LANE = 2
ABI_VERSION = 17  # RESTIR_ABI_VERSION of include/restir_abi.h

ERRORS = {
    -1: "RESTIR_E_INVALID_ARGUMENT",
    -2: "RESTIR_E_SIZE_MISMATCH",
    -3: "RESTIR_E_NOT_READY",
    -4: "RESTIR_E_HIP",
    -5: "RESTIR_E_BAD_TREE",
    -6: "RESTIR_E_ALIASING",
    -7: "RESTIR_E_NO_DEVICE",
}

# numpy views of the wire formats (restirStructs.glsl:19-34, aabbTree.glsl:1-13, light.glsl:2-20)
RESERVOIR_DTYPE = np.dtype(
    [("pos_lum", "<f4", 4), ("normal", "<f4", 4), ("lightIndex", "<i4"), ("pHat", "<f4"),
     ("sumWeights", "<f4"), ("w", "<f4"), ("M", "<u4"), ("pad", "<u4", 3)])
NODE_DTYPE = np.dtype(
    [("lmin", "<f4", 4), ("lmax", "<f4", 4), ("rmin", "<f4", 4), ("rmax", "<f4", 4),
     ("left", "<i4"), ("right", "<i4"), ("pad", "<i4", 2)])
TRIANGLE_DTYPE = np.dtype([("p1", "<f4", 4), ("p2", "<f4", 4), ("p3", "<f4", 4)])
POINT_LIGHT_DTYPE = np.dtype([("pos", "<f4", 4), ("color_lum", "<f4", 4)])
TRI_LIGHT_DTYPE = np.dtype(
    [("p1", "<f4", 4), ("p2", "<f4", 4), ("p3", "<f4", 4), ("emission_lum", "<f4", 4),
     ("normal_area", "<f4", 4)])
ALIAS_DTYPE = np.dtype([("prob", "<f4"), ("alias", "<i4"), ("oriProb", "<f4"),
                        ("aliasOriProb", "<f4")])
assert RESERVOIR_DTYPE.itemsize == 64 and NODE_DTYPE.itemsize == 80
assert TRIANGLE_DTYPE.itemsize == 48 and POINT_LIGHT_DTYPE.itemsize == 32
assert TRI_LIGHT_DTYPE.itemsize == 80 and ALIAS_DTYPE.itemsize == 16


class Uniforms(C.Structure):
    """RestirUniforms (restirStructs.glsl:40-56), 128 bytes."""
    _fields_ = [
        ("prevFrameProjectionViewMatrix", C.c_float * 16),
        ("cameraPos", C.c_float * 4),
        ("screenSize", C.c_uint32 * 2),
        ("frame", C.c_uint32),
        ("initialLightSampleCount", C.c_uint32),
        ("temporalSampleCountMultiplier", C.c_uint32),
        ("spatialPosThreshold", C.c_float),
        ("spatialNormalThreshold", C.c_float),
        ("spatialNeighbors", C.c_uint32),
        ("spatialRadius", C.c_float),
        ("flags", C.c_int32),
        ("_pad", C.c_uint32 * 2),
    ]


class GBuffer(C.Structure):
    _fields_ = [
        ("position", C.c_void_p),
        ("normal", C.c_void_p),
        ("albedo", C.c_void_p),
        ("material", C.c_void_p),
        ("depth", C.c_void_p),
        ("plane_stride", C.c_uint64),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("normal_format", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


NORMAL_F32, NORMAL_F16 = 0, 1  # RESTIR_NORMAL_*
# restir_texel_format: the reference's G-buffer attachment formats (gBufferPass.cpp:75-108)
TEXEL = dict(RGBA8_SRGB=1, RGBA8_UNORM=2, RGB16_SNORM=3, RGBA16_SNORM=4, RGB16_SFLOAT=5, RGBA16_SFLOAT=6,
             RGB32_SFLOAT=7, RGBA32_SFLOAT=8, RG16_UNORM=9, D32_SFLOAT=10, D24_UNORM=11)


class Attachment(C.Structure):
    _fields_ = [("data", C.c_void_p), ("format", C.c_uint32), ("row_bytes", C.c_uint32)]


class GBufferAttachments(C.Structure):
    _fields_ = [("albedo", Attachment), ("normal", Attachment), ("material", Attachment),
                ("world_position", Attachment), ("depth", Attachment), ("width", C.c_uint32),
                ("height", C.c_uint32)]
SCHEDULE_QUEUED = 0
SCHEDULE_INLINE = 1
LIGHT_BUFFER_HEADER_BYTES = 16  # RESTIR_LIGHT_BUFFER_HEADER_BYTES


class Launch(C.Structure):
    _fields_ = [("tile_x", C.c_uint32), ("tile_y", C.c_uint32), ("schedule", C.c_uint32),
                ("_reserved", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("lookAt", C.c_float * 3),
        ("worldUp", C.c_float * 3),
        ("fovYRadians", C.c_float),
        ("zNear", C.c_float),
        ("zFar", C.c_float),
    ]


class MlpStruct(C.Structure):
    """restir_mlp: the learned proposal's parameter buffers (device pointers)."""
    _fields_ = [
        ("n_clusters", C.c_uint32),
        ("adam_step", C.c_uint32),
        ("weights_bf16", C.c_void_p),
        ("weights", C.c_void_p),
        ("bias", C.c_void_p),
        ("adam_m", C.c_void_p),
        ("adam_v", C.c_void_p),
        ("adam_step_device", C.c_void_p),
    ]


def mlp_sizes(k: int):
    """(weights, biases) element counts of RESTIR_MLP_N_WEIGHTS / RESTIR_MLP_N_BIAS."""
    return 3 * 128 * 128 + k * 128, 3 * 128 + k


assert C.sizeof(Uniforms) == 128 and C.sizeof(GBuffer) == 64 and C.sizeof(Camera) == 48
assert C.sizeof(MlpStruct) == 56


class View(C.Structure):
    """restir_view: one view of a batched dispatch."""
    _fields_ = [("scene", C.c_void_p), ("uniforms", C.POINTER(Uniforms)), ("cur", C.POINTER(GBuffer)),
                ("prev", C.POINTER(GBuffer)), ("in_", C.c_void_p), ("out", C.c_void_p),
                ("rgb", C.c_void_p), ("rgb_stride", C.c_uint64)]


class KernelTime(C.Structure):
    """restir_kernel_time: one row of restir_profile_read."""
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint32), ("total_ms", C.c_float),
                ("max_ms", C.c_float), ("_reserved", C.c_uint32)]


assert C.sizeof(KernelTime) == 64 and C.sizeof(View) == 64

# Exported symbols with their signatures (restype, argtypes); the CPU test suite checks that the
# library exports every one of them and that include/restir_abi.h declares the same set.
_vp, _u32, _u64, _i32, _f32, _sz = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float, C.c_size_t
SIGNATURES = {
    "restir_version": (_i32, []),
    "restir_create": (_i32, [C.POINTER(_vp), _i32]),
    "restir_destroy": (None, [_vp]),
    "restir_last_error": (C.c_char_p, [_vp]),
    "restir_set_lights": (_i32, [_vp, _vp, _sz, _vp, _sz, _vp, _sz]),
    "restir_set_bvh": (_i32, [_vp, _vp, _u32, _vp, _u32]),
    "restir_build_bvh": (_i32, [_vp, _vp, _u32, _vp, _vp, C.POINTER(_u32), _vp]),
    "restir_initial_temporal": (_i32, [_vp, C.POINTER(Uniforms), C.POINTER(GBuffer),
                                       C.POINTER(GBuffer), _vp, _vp, C.POINTER(Launch), _vp]),
    "restir_initial_temporal_noalloc": (_i32, [_vp, C.POINTER(Uniforms), C.POINTER(GBuffer),
                                       C.POINTER(GBuffer), _vp, _vp, C.POINTER(Launch), _vp]),
    "restir_spatial": (_i32, [_vp, C.POINTER(Uniforms), _i32, C.POINTER(GBuffer), _vp, _vp,
                              C.POINTER(Launch), _vp]),
    "restir_unbiased": (_i32, [_vp, C.POINTER(Uniforms), _i32, _i32, C.POINTER(GBuffer), _vp,
                               _vp, C.POINTER(Launch), _vp]),
    "restir_shade": (_i32, [_vp, C.POINTER(_f32), _f32, C.POINTER(GBuffer), _vp, _vp, _u64, _vp]),
    "restir_shade_ex": (_i32, [_vp, C.POINTER(_f32), _f32, _u32, C.POINTER(GBuffer), _vp, _vp, _u64, _vp]),
    "restir_trace_shadow": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _i32, _vp]),
    "restir_build_aabb_tree": (_i32, [_vp, _u32, _vp, _vp, C.POINTER(_u32)]),
    "restir_build_alias_table": (_i32, [_vp, _u32, _vp, _u32, _vp]),
    "restir_build_alias_table_device": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _vp]),
    "restir_build_alias_table_mixed": (_i32, [_vp, _u32, _vp, _u32, _vp]),
    "restir_build_light_clusters": (_i32, [_vp, _u32, _u32, C.c_uint64, _u32, _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp]),
    "restir_build_light_clusters_device": (_i32, [_vp, _vp, _u32, _u32, C.c_uint64, _u32, _vp, _vp, _vp,
                                                  _vp, _vp, _vp, _vp]),
    "restir_gbuffer_from_attachments": (_i32, [_vp, C.POINTER(GBufferAttachments), _vp, _vp, _u32, _vp, _vp,
                                               _vp, _u64, _vp]),
    "restir_primary_gbuffer": (_i32, [_vp, C.POINTER(Camera), _vp, _vp, _u32, _vp, _vp, _u32, _vp,
                                      _vp, _vp, _u64, _u32, _u32, _vp]),
    "restir_camera_projection_view": (_i32, [C.POINTER(Camera), _f32, C.POINTER(_f32)]),
    "restir_mlp_init": (_i32, [_vp, C.POINTER(MlpStruct), C.c_uint64, _vp, _vp]),
    "restir_set_light_clusters": (_i32, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "restir_learned_candidates": (_i32, [_vp, C.POINTER(Uniforms), C.POINTER(GBuffer),
                                         C.POINTER(MlpStruct), C.POINTER(_f32), C.POINTER(_f32),
                                         _u32, _f32, _vp, _vp, _vp]),
    "restir_initial_temporal_candidates": (_i32, [_vp, C.POINTER(Uniforms), C.POINTER(GBuffer),
                                                  C.POINTER(GBuffer), _vp, _vp, _vp, _vp,
                                                  C.POINTER(Launch), _vp]),
    "restir_mlp_train": (_i32, [_vp, C.POINTER(MlpStruct), C.POINTER(Uniforms), C.POINTER(GBuffer),
                                _vp, C.POINTER(_f32), C.POINTER(_f32), _u32, _f32, _u32, _vp, _vp]),
    "restir_mlp_features": (_i32, [_vp, C.POINTER(GBuffer), C.POINTER(_f32), C.POINTER(_f32),
                                   C.POINTER(_f32), _vp, _u32, _vp, _vp]),
    "restir_initial_temporal_views": (_i32, [_vp, _vp, _u32, C.POINTER(Launch), _vp]),
    "restir_spatial_views": (_i32, [_vp, _vp, _u32, _i32, C.POINTER(Launch), _vp]),
    "restir_unbiased_views": (_i32, [_vp, _vp, _u32, _i32, _i32, C.POINTER(Launch), _vp]),
    "restir_shade_views": (_i32, [_vp, _vp, _u32, _f32, _vp]),
    "restir_ground_truth": (_i32, [_vp, C.POINTER(_f32), C.POINTER(GBuffer), _vp, _u32, _u32, _vp,
                                   _u64, _vp]),
    "restir_profile_enable": (_i32, [_vp, _i32]),
    "restir_profile_read": (_i32, [_vp, _vp, _u32, C.POINTER(_u32)]),
    "restir_set_ray_export": (_i32, [_vp, _i32]),
    "restir_unbiased_rays": (_i32, [_vp, _i32, _vp, _vp, _u32, C.POINTER(_u32), _vp]),
    "restir_mlp_forward": (_i32, [_vp, C.POINTER(MlpStruct), _vp, _u32, _vp, _vp]),
    "restir_mlp_loss_grad": (_i32, [_vp, C.POINTER(MlpStruct), _vp, _vp, _vp, _u32, _vp, _vp, _vp,
                                    _vp]),
    "restir_mlp_adam": (_i32, [_vp, C.POINTER(MlpStruct), _vp, _vp, _f32, _vp]),
}


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v[:3]])


class RestirError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {message}")
        self.code = code


_LIB = None


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load the built library. Fails loudly if it has not been built: there is no fallback."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `make -C {PKG_DIR}` (or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.restir_version() != ABI_VERSION:
        raise RuntimeError(f"{p} implements ABI revision {lib.restir_version()}, this binding expects "
                           f"{ABI_VERSION}: rebuild the library (`make -C {PKG_DIR}`)")
    if path is None:
        _LIB = lib
    return lib


def as_ptr(x) -> int | None:
    """Device pointer of a torch tensor, a raw int pointer, or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


# ------------------------------------------------------------------------- host-side builders
def build_aabb_tree(tri_xyz: np.ndarray):
    """AabbTree::build (aabbTreeBuilder.cpp:52-214) -> (nodes, triangles, depth)."""
    lib = load_library()
    tri_xyz = np.ascontiguousarray(tri_xyz, dtype=np.float32).reshape(-1, 9)
    n = tri_xyz.shape[0]
    nodes = np.zeros(max(n - 1, 0), dtype=NODE_DTYPE)
    tris = np.zeros(n, dtype=TRIANGLE_DTYPE)
    depth = C.c_uint32(0)
    rc = lib.restir_build_aabb_tree(tri_xyz.ctypes.data, n, nodes.ctypes.data, tris.ctypes.data,
                                    C.byref(depth))
    if rc != 0:
        raise RestirError(rc, "restir_build_aabb_tree", "build failed")
    return nodes, tris, depth.value


def build_alias_table(points: np.ndarray | None, tri_lights: np.ndarray | None) -> np.ndarray:
    """createAliasTable (misc.cpp:418-497)."""
    lib = load_library()
    np_ = 0 if points is None else len(points)
    nt = 0 if tri_lights is None else len(tri_lights)
    n = np_ if np_ > 0 else nt
    cols = np.zeros(n, dtype=ALIAS_DTYPE)
    pts = np.ascontiguousarray(points, dtype=POINT_LIGHT_DTYPE) if np_ else None
    tls = np.ascontiguousarray(tri_lights, dtype=TRI_LIGHT_DTYPE) if nt else None
    rc = lib.restir_build_alias_table(pts.ctypes.data if pts is not None else None, np_,
                                      tls.ctypes.data if tls is not None else None, nt,
                                      cols.ctypes.data)
    if rc != 0:
        raise RestirError(rc, "restir_build_alias_table", "build failed")
    return cols


def build_alias_table_mixed(points: np.ndarray, tri_lights: np.ndarray) -> np.ndarray:
    """One table over the point lights followed by the triangle lights (extension)."""
    lib = load_library()
    pts = np.ascontiguousarray(points, dtype=POINT_LIGHT_DTYPE)
    tls = np.ascontiguousarray(tri_lights, dtype=TRI_LIGHT_DTYPE)
    cols = np.zeros(len(pts) + len(tls), dtype=ALIAS_DTYPE)
    rc = lib.restir_build_alias_table_mixed(pts.ctypes.data if len(pts) else None, len(pts),
                                            tls.ctypes.data if len(tls) else None, len(tls),
                                            cols.ctypes.data)
    if rc != 0:
        raise RestirError(rc, "restir_build_alias_table_mixed", "build failed")
    return cols


def uses_union(points, tri_lights, union=None) -> bool:
    """Whether a scene's alias table covers both kinds of lights: by default when both are
    present (the reference's table covers point lights if any, else triangle lights)."""
    both = points is not None and len(points) > 0 and tri_lights is not None and len(tri_lights) > 0
    return both if union is None else bool(union) and both


def scene_alias_table(points, tri_lights, union=None) -> np.ndarray:
    """The scene's alias table: one table over both kinds (uses_union), else createAliasTable's."""
    if uses_union(points, tri_lights, union):
        return build_alias_table_mixed(points, tri_lights)
    return build_alias_table(points if points is not None and len(points) else None,
                             tri_lights if tri_lights is not None and len(tri_lights) else None)


def union_light_records(points: np.ndarray, tri_lights: np.ndarray) -> np.ndarray:
    """Point-light records for the union light space: the point lights, then one record per
    triangle light (centroid position, luminance * area), as the cluster builder takes them."""
    t = np.zeros(len(tri_lights), dtype=POINT_LIGHT_DTYPE)
    if len(tri_lights):
        c = (tri_lights["p1"][:, :3] + tri_lights["p2"][:, :3] + tri_lights["p3"][:, :3]) / np.float32(3.0)
        t["pos"][:, :3] = c
        t["pos"][:, 3] = 1.0
        t["color_lum"][:, :3] = tri_lights["emission_lum"][:, :3]
        t["color_lum"][:, 3] = (tri_lights["emission_lum"][:, 3] * tri_lights["normal_area"][:, 3]).astype(np.float32)
    return np.concatenate([np.asarray(points, dtype=POINT_LIGHT_DTYPE), t])


class LightClusters:
    """Host arrays of restir_build_light_clusters (restir_build_light_clusters_device, whose
    nearest-centroid step runs on the context's GPU, when `ctx` is given: same bytes)."""

    def __init__(self, points: np.ndarray, k: int, seed: int = 7, iterations: int = 16, ctx=None):
        lib = load_library()
        n = len(points)
        pts = np.ascontiguousarray(points, dtype=POINT_LIGHT_DTYPE)
        self.k = k
        self.light_cluster = np.zeros(n, dtype=np.int32)
        self.light_pwithin = np.zeros(n, dtype=np.float32)
        self.begin = np.zeros(k, dtype=np.int32)
        self.count = np.zeros(k, dtype=np.int32)
        self.lights = np.zeros(n, dtype=np.int32)
        self.alias = np.zeros(n, dtype=ALIAS_DTYPE)
        self.centroids = np.zeros((k, 3), dtype=np.float32)
        outs = (self.light_cluster.ctypes.data, self.light_pwithin.ctypes.data, self.begin.ctypes.data,
                self.count.ctypes.data, self.lights.ctypes.data, self.alias.ctypes.data,
                self.centroids.ctypes.data)
        if ctx is None:
            rc = lib.restir_build_light_clusters(pts.ctypes.data, n, k, seed, iterations, *outs)
            if rc != 0:
                raise RestirError(rc, "restir_build_light_clusters", f"n={n} k={k}")
        else:
            ctx._check(lib.restir_build_light_clusters_device(ctx.h, pts.ctypes.data, n, k, seed, iterations, *outs),
                       "restir_build_light_clusters_device")
        lum = pts["color_lum"][:, 3].astype(np.float64)
        power = np.bincount(self.light_cluster, weights=lum, minlength=k)
        self.power_share = (power / power.sum()).astype(np.float32)


def pack_buffer(records: np.ndarray | None, dtype: np.dtype) -> np.ndarray:
    """SSBO image: int32 count + 12 bytes of padding, then the records (sceneBuffers.h:100-124)."""
    n = 0 if records is None else len(records)
    h = LIGHT_BUFFER_HEADER_BYTES
    buf = np.zeros(h + n * dtype.itemsize, dtype=np.uint8)
    buf[:4] = np.frombuffer(np.int32(n).tobytes(), dtype=np.uint8)
    if n:
        buf[h:] = np.frombuffer(np.ascontiguousarray(records, dtype=dtype).tobytes(), dtype=np.uint8)
    return buf


def camera_struct(cam: dict) -> Camera:
    c = Camera()
    c.position[:] = [float(v) for v in cam["position"]]
    c.lookAt[:] = [float(v) for v in cam["lookAt"]]
    c.worldUp[:] = [float(v) for v in cam.get("worldUp", (0.0, 1.0, 0.0))]
    c.fovYRadians = float(cam["fovY"])
    c.zNear = float(cam.get("zNear", 0.01))
    c.zFar = float(cam.get("zFar", 1000.0))
    return c


def projection_view(cam: dict, aspect: float) -> np.ndarray:
    lib = load_library()
    out = (C.c_float * 16)()
    cs = camera_struct(cam)
    rc = lib.restir_camera_projection_view(C.byref(cs), float(aspect), out)
    if rc != 0:
        raise RestirError(rc, "restir_camera_projection_view", "bad camera")
    return np.array(out, dtype=np.float32)


def make_uniforms(width: int, height: int, frame: int, camera_pos, prev_pv=None, *,
                  M: int = 32, temporal_mult: int = 20, pos_threshold: float = 0.1,
                  normal_threshold: float = 25.0, neighbors: int = 4, radius: float = 30.0,
                  flags: int = VISIBILITY_REUSE_FLAG | TEMPORAL_REUSE_FLAG) -> Uniforms:
    """RestirUniforms with the reference's defaults (app.cpp:413-434, app.h:150-174)."""
    u = Uniforms()
    if prev_pv is not None:
        u.prevFrameProjectionViewMatrix[:] = [float(v) for v in np.asarray(prev_pv).ravel()]
    u.cameraPos[:] = [float(camera_pos[0]), float(camera_pos[1]), float(camera_pos[2]), 1.0]
    u.screenSize[:] = [width, height]
    u.frame = frame
    u.initialLightSampleCount = M
    u.temporalSampleCountMultiplier = temporal_mult
    u.spatialPosThreshold = pos_threshold
    u.spatialNormalThreshold = normal_threshold
    u.spatialNeighbors = neighbors
    u.spatialRadius = radius
    u.flags = flags
    return u


def gbuffer_struct(base_ptr: int, width: int, height: int, plane_stride: int | None = None,
                   normal_format: int = NORMAL_F32) -> GBuffer:
    """G-buffer view of one float32 allocation holding 13 planes in the order
    position(3), normal(3), albedo(4: rgb + emissive flag), material(2: roughness, metallic),
    depth(1). With NORMAL_F16 the three normal planes are halves (stride plane_stride halves)
    at the start of the normal region."""
    P = plane_stride or width * height
    g = GBuffer()
    g.position = base_ptr
    g.normal = base_ptr + 4 * 3 * P
    g.albedo = base_ptr + 4 * 6 * P
    g.material = base_ptr + 4 * 10 * P
    g.depth = base_ptr + 4 * 12 * P
    g.plane_stride = P
    g.width = width
    g.height = height
    g.normal_format = normal_format
    return g


GBUFFER_PLANES = 13


class Context:
    """Owns one restir_ctx (one HIP device). Methods map 1:1 onto the C entry points."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.restir_create(C.byref(h), device)
        if rc != 0:
            raise RestirError(rc, "restir_create", f"device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.restir_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, where: str):
        if rc != 0:
            msg = self.lib.restir_last_error(self.h)
            raise RestirError(rc, where, msg.decode() if msg else "")

    def set_lights(self, point_buf, point_bytes, tri_buf, tri_bytes, alias_buf, alias_bytes):
        self._check(self.lib.restir_set_lights(self.h, as_ptr(point_buf), point_bytes,
                                               as_ptr(tri_buf), tri_bytes, as_ptr(alias_buf),
                                               alias_bytes), "restir_set_lights")

    def set_bvh(self, nodes, n_nodes, tris, n_tris):
        self._check(self.lib.restir_set_bvh(self.h, as_ptr(nodes), n_nodes, as_ptr(tris), n_tris),
                    "restir_set_bvh")

    def build_bvh(self, tri_xyz, n_tris, nodes, tris, stream=None) -> int:
        """Device tree build (restir_build_bvh) from device vertices [n_tris, 9]; binds the tree
        to this context and returns its depth."""
        depth = C.c_uint32(0)
        self._check(self.lib.restir_build_bvh(self.h, as_ptr(tri_xyz), n_tris, as_ptr(nodes),
                                              as_ptr(tris), C.byref(depth), stream_ptr(stream)),
                    "restir_build_bvh")
        return depth.value

    def build_alias_table(self, points, n_point, tri_lights, n_tri, columns, stream=None):
        """Device alias-table build (restir_build_alias_table_device) from device light records
        (no SSBO header) into device columns."""
        self._check(self.lib.restir_build_alias_table_device(
            self.h, as_ptr(points), n_point, as_ptr(tri_lights), n_tri, as_ptr(columns),
            stream_ptr(stream)), "restir_build_alias_table_device")

    def initial_temporal(self, u, cur, prev, prev_res, out, launch=None, stream=None):
        self._check(self.lib.restir_initial_temporal(
            self.h, C.byref(u), C.byref(cur), C.byref(prev) if prev is not None else None,
            as_ptr(prev_res), as_ptr(out), C.byref(launch) if launch is not None else None,
            stream_ptr(stream)), "restir_initial_temporal")

    def initial_temporal_noalloc(self, u, cur, prev, prev_res, out, launch=None, stream=None):
        self._check(self.lib.restir_initial_temporal_noalloc(
            self.h, C.byref(u), C.byref(cur), C.byref(prev) if prev is not None else None,
            as_ptr(prev_res), as_ptr(out), C.byref(launch) if launch is not None else None,
            stream_ptr(stream)), "restir_initial_temporal_noalloc")

    def spatial(self, u, it, g, res_in, res_out, launch=None, stream=None):
        self._check(self.lib.restir_spatial(
            self.h, C.byref(u), it, C.byref(g), as_ptr(res_in), as_ptr(res_out),
            C.byref(launch) if launch is not None else None, stream_ptr(stream)), "restir_spatial")

    def unbiased(self, u, k, it, g, res_in, res_out, launch=None, stream=None):
        self._check(self.lib.restir_unbiased(
            self.h, C.byref(u), k, it, C.byref(g), as_ptr(res_in), as_ptr(res_out),
            C.byref(launch) if launch is not None else None, stream_ptr(stream)), "restir_unbiased")

    def shade(self, camera_pos, gamma, g, res, rgb, rgb_stride, stream=None, visibility=False):
        """restir_shade, or with visibility=True restir_shade_ex with the final visibility re-test."""
        cp = (C.c_float * 3)(*[float(v) for v in camera_pos[:3]])
        if visibility:
            self._check(self.lib.restir_shade_ex(self.h, cp, float(gamma), SHADE_VISIBILITY_FLAG, C.byref(g),
                                                 as_ptr(res), as_ptr(rgb), rgb_stride, stream_ptr(stream)),
                        "restir_shade_ex")
            return
        self._check(self.lib.restir_shade(self.h, cp, float(gamma), C.byref(g), as_ptr(res),
                                          as_ptr(rgb), rgb_stride, stream_ptr(stream)),
                    "restir_shade")

    def trace_shadow(self, p1, p2, n, shadowed, visits=None, tests=None, reference_loop=False,
                     one_ray_per_lane=False, stream=None):
        loop = (TRACE_REFERENCE_LOOP if reference_loop else TRACE_FIRST_HIT) | \
               (TRACE_ONE_RAY_PER_LANE if one_ray_per_lane else 0)
        self._check(self.lib.restir_trace_shadow(self.h, as_ptr(p1), as_ptr(p2), n,
                                                 as_ptr(shadowed), as_ptr(visits), as_ptr(tests),
                                                 loop, stream_ptr(stream)), "restir_trace_shadow")

    # ------------------------------------------------------------------- batched views
    @staticmethod
    def _views(views):
        arr = (View * len(views))(*views)
        return arr, len(views)

    def initial_temporal_views(self, views, launch=None, stream=None):
        arr, n = self._views(views)
        self._check(self.lib.restir_initial_temporal_views(
            self.h, arr, n, C.byref(launch) if launch is not None else None, stream_ptr(stream)),
            "restir_initial_temporal_views")

    def spatial_views(self, views, it, launch=None, stream=None):
        arr, n = self._views(views)
        self._check(self.lib.restir_spatial_views(
            self.h, arr, n, it, C.byref(launch) if launch is not None else None, stream_ptr(stream)),
            "restir_spatial_views")

    def unbiased_views(self, views, k, it, launch=None, stream=None):
        arr, n = self._views(views)
        self._check(self.lib.restir_unbiased_views(
            self.h, arr, n, k, it, C.byref(launch) if launch is not None else None, stream_ptr(stream)),
            "restir_unbiased_views")

    def shade_views(self, views, gamma=1.0, stream=None):
        arr, n = self._views(views)
        self._check(self.lib.restir_shade_views(self.h, arr, n, float(gamma), stream_ptr(stream)),
                    "restir_shade_views")

    def ground_truth(self, camera_pos, g, pixels, n, rgb, rgb_stride, n_tri_samples=256, stream=None):
        """FP64 direct lighting of the listed pixels into rgb (device float64, [3][rgb_stride])."""
        cp = (C.c_float * 3)(*[float(v) for v in camera_pos[:3]])
        self._check(self.lib.restir_ground_truth(self.h, cp, C.byref(g), as_ptr(pixels), n,
                                                 n_tri_samples, as_ptr(rgb), rgb_stride,
                                                 stream_ptr(stream)), "restir_ground_truth")

    # ------------------------------------------------------------------- measurement
    def profile_enable(self, on=True):
        self._check(self.lib.restir_profile_enable(self.h, 1 if on else 0), "restir_profile_enable")

    def profile_read(self) -> dict:
        """{kernel name: (launches, total ms, max ms)} of the kernels launched since the last read."""
        cap = 64
        rows = (KernelTime * cap)()
        n = C.c_uint32(0)
        self._check(self.lib.restir_profile_read(self.h, rows, cap, C.byref(n)), "restir_profile_read")
        return {rows[i].name.decode(): (rows[i].launches, rows[i].total_ms, rows[i].max_ms)
                for i in range(min(n.value, cap))}

    def set_ray_export(self, on=True):
        self._check(self.lib.restir_set_ray_export(self.h, 1 if on else 0), "restir_set_ray_export")

    def unbiased_rays(self, which, p1, p2, capacity, stream=None) -> int:
        """Copies the last queued Pass C's centre (0) or neighbour (1) rays; returns the queue length."""
        n = C.c_uint32(0)
        self._check(self.lib.restir_unbiased_rays(self.h, which, as_ptr(p1), as_ptr(p2), capacity,
                                                  C.byref(n), stream_ptr(stream)), "restir_unbiased_rays")
        return n.value

    # ------------------------------------------------------------------- learned proposal
    def mlp_init(self, mlp: "MlpStruct", seed, prior, stream=None):
        self._check(self.lib.restir_mlp_init(self.h, C.byref(mlp), seed, as_ptr(prior),
                                             stream_ptr(stream)), "restir_mlp_init")

    def set_light_clusters(self, k, light_cluster, light_pwithin, begin, count, lights, alias):
        self._check(self.lib.restir_set_light_clusters(
            self.h, k, as_ptr(light_cluster), as_ptr(light_pwithin), as_ptr(begin), as_ptr(count),
            as_ptr(lights), as_ptr(alias)), "restir_set_light_clusters")

    def learned_candidates(self, u, g, mlp, lo, hi, M, mix, cand_light, cand_q, stream=None):
        self._check(self.lib.restir_learned_candidates(
            self.h, C.byref(u), C.byref(g), C.byref(mlp), _f3(lo), _f3(hi), M, float(mix),
            as_ptr(cand_light), as_ptr(cand_q), stream_ptr(stream)), "restir_learned_candidates")

    def initial_temporal_candidates(self, u, cur, prev, prev_res, out, cand_light, cand_q,
                                    launch=None, stream=None):
        self._check(self.lib.restir_initial_temporal_candidates(
            self.h, C.byref(u), C.byref(cur), C.byref(prev) if prev is not None else None,
            as_ptr(prev_res), as_ptr(out), as_ptr(cand_light), as_ptr(cand_q),
            C.byref(launch) if launch is not None else None, stream_ptr(stream)),
            "restir_initial_temporal_candidates")

    def mlp_train(self, mlp, u, g, res, lo, hi, n, lr, seed, loss_out=None, stream=None):
        self._check(self.lib.restir_mlp_train(
            self.h, C.byref(mlp), C.byref(u), C.byref(g), as_ptr(res), _f3(lo), _f3(hi), n,
            float(lr), seed, as_ptr(loss_out), stream_ptr(stream)), "restir_mlp_train")

    def mlp_features(self, g, cam_pos, lo, hi, pixels, n, out, stream=None):
        self._check(self.lib.restir_mlp_features(
            self.h, C.byref(g), _f3(cam_pos), _f3(lo), _f3(hi), as_ptr(pixels), n, as_ptr(out),
            stream_ptr(stream)), "restir_mlp_features")

    def mlp_forward(self, mlp, X, n, logits, stream=None):
        self._check(self.lib.restir_mlp_forward(self.h, C.byref(mlp), as_ptr(X), n, as_ptr(logits),
                                                stream_ptr(stream)), "restir_mlp_forward")

    def mlp_adam(self, mlp, gw, gb, lr, stream=None):
        self._check(self.lib.restir_mlp_adam(self.h, C.byref(mlp), as_ptr(gw), as_ptr(gb), lr,
                                             stream_ptr(stream)), "restir_mlp_adam")

    def mlp_loss_grad(self, mlp, X, target, mask, n, gw, gb, loss, stream=None):
        self._check(self.lib.restir_mlp_loss_grad(
            self.h, C.byref(mlp), as_ptr(X), as_ptr(target), as_ptr(mask), n, as_ptr(gw),
            as_ptr(gb), as_ptr(loss), stream_ptr(stream)), "restir_mlp_loss_grad")

    def gbuffer_from_attachments(self, atts: dict, width, height, gbuf_base, plane_stride=None,
                                 normal_format=NORMAL_F32, stream=None):
        """atts: {"albedo" | "normal" | "material" | "world_position" | "depth": (device buffer,
        TEXEL format name, row bytes)}; fills a G-buffer laid out as gbuffer_struct describes."""
        P = plane_stride or width * height
        ga = GBufferAttachments()
        for k in ("albedo", "normal", "material", "world_position", "depth"):
            buf, fmt, row = atts[k]
            setattr(ga, k, Attachment(as_ptr(buf), TEXEL[fmt], row))
        ga.width, ga.height = width, height
        base = as_ptr(gbuf_base)
        self._check(self.lib.restir_gbuffer_from_attachments(
            self.h, C.byref(ga), base, base + 4 * 3 * P, normal_format, base + 4 * 6 * P, base + 4 * 10 * P,
            base + 4 * 12 * P, P, stream_ptr(stream)), "restir_gbuffer_from_attachments")

    def primary_gbuffer(self, cam: dict, tri_material, materials, n_materials, gbuf_base, width,
                        height, plane_stride=None, stream=None, normal_format=NORMAL_F32):
        """Fills a G-buffer laid out as gbuffer_struct describes (normals as float or half planes)."""
        P = plane_stride or width * height
        cs = camera_struct(cam)
        base = as_ptr(gbuf_base)
        self._check(self.lib.restir_primary_gbuffer(
            self.h, C.byref(cs), as_ptr(tri_material), as_ptr(materials), n_materials,
            base, base + 4 * 3 * P, normal_format, base + 4 * 6 * P, base + 4 * 10 * P, base + 4 * 12 * P, P,
            width, height, stream_ptr(stream)), "restir_primary_gbuffer")
