"""Seeded procedural scenes for the ReSTIR path (synthetic inputs; the reference loads glTF).

Everything is drawn from numpy's PCG64 generator (``np.random.default_rng(seed)``), which is
deterministic across platforms, unlike the reference's ``std::default_random_engine``
(src/misc.cpp:367). Light records follow pointLight / triLight (light.glsl:2-13) with the
luminance in ``w`` computed as ``shader::luminance`` does (common.glsl:7-9).

Scenes:
  * ``cornell_box``  — 5 walls (10 triangles, open front) + 2 yawed boxes (24 triangles),
                       64 point lights in the upper half of the room, log-uniform RGB [0.1, 10].
  * ``terrain``      — a G x G vertex heightfield with value-noise displacement + random oriented
                       boxes; point lights uniform in the scene AABB, log-uniform RGB.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .abi import POINT_LIGHT_DTYPE, TRI_LIGHT_DTYPE


@dataclass
class Scene:
    name: str
    tri_xyz: np.ndarray            # (n, 9) float32, world space, scene order
    tri_material: np.ndarray       # (n,) int32
    materials: np.ndarray          # (m, 8) float32: albedo rgb, emissive flag, rough, metal, 0, 0
    point_lights: np.ndarray       # POINT_LIGHT_DTYPE records
    tri_lights: np.ndarray         # TRI_LIGHT_DTYPE records (may be empty)
    camera: dict                   # position, lookAt, worldUp, fovY, zNear, zFar
    meta: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return int(self.tri_xyz.shape[0])

    def aabb(self):
        v = self.tri_xyz.reshape(-1, 3)
        return v.min(axis=0), v.max(axis=0)


def luminance(r, g, b):
    """common.glsl:7-9 in float32 with the same operation order."""
    r, g, b = (np.asarray(c, dtype=np.float32) for c in (r, g, b))
    return (np.float32(0.2126) * r + np.float32(0.7152) * g) + np.float32(0.0722) * b


def make_point_lights(pos: np.ndarray, rgb: np.ndarray) -> np.ndarray:
    n = len(pos)
    out = np.zeros(n, dtype=POINT_LIGHT_DTYPE)
    out["pos"][:, :3] = pos
    out["pos"][:, 3] = 1.0
    out["color_lum"][:, :3] = rgb
    out["color_lum"][:, 3] = luminance(rgb[:, 0], rgb[:, 1], rgb[:, 2])
    return out


def make_tri_lights(tri_xyz: np.ndarray, emission: np.ndarray) -> np.ndarray:
    """collectTriangleLightsFromScene (misc.cpp:380-414): unit normal and area per triangle."""
    t = np.asarray(tri_xyz, dtype=np.float32).reshape(-1, 3, 3)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]).astype(np.float32)
    norm = np.linalg.norm(n, axis=1).astype(np.float32)
    out = np.zeros(len(t), dtype=TRI_LIGHT_DTYPE)
    for k, name in enumerate(("p1", "p2", "p3")):
        out[name][:, :3] = t[:, k]
        out[name][:, 3] = 1.0
    out["emission_lum"][:, :3] = emission
    out["emission_lum"][:, 3] = luminance(emission[:, 0], emission[:, 1], emission[:, 2])
    out["normal_area"][:, :3] = n / norm[:, None]
    out["normal_area"][:, 3] = norm * np.float32(0.5)
    return out


def _quad(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d)."""
    return [np.concatenate([a, b, c]), np.concatenate([a, c, d])]


def box_triangles(center, size, rot) -> np.ndarray:
    """12 triangles of an oriented box (rot: 3x3 matrix applied to the local corners)."""
    hx, hy, hz = (0.5 * np.asarray(size, dtype=np.float64))
    corners = np.array([[sx * hx, sy * hy, sz * hz] for sx in (-1, 1) for sy in (-1, 1)
                        for sz in (-1, 1)])
    corners = corners @ np.asarray(rot).T + np.asarray(center)
    c = lambda i: corners[i]  # noqa: E731  index = 4*sx + 2*sy + sz
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for f in faces:
        tris += _quad(c(f[0]), c(f[1]), c(f[2]), c(f[3]))
    return np.array(tris, dtype=np.float32)


def _yaw(theta):
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _materials(rng, n, emissive_rows=()):
    m = np.zeros((n, 8), dtype=np.float32)
    m[:, 0:3] = rng.uniform(0.2, 0.9, size=(n, 3))
    m[:, 4] = rng.uniform(0.1, 1.0, size=n)
    for r in emissive_rows:
        m[r, 3] = 1.0
    return m


def cornell_box(seed: int = 1, n_lights: int = 64, light_margin: float = 0.05) -> Scene:
    """Config 1: room x,z in [-1,1], y in [0,2]; camera 60 deg fovY on the open front. Lights lie
    in the upper half of the room at least `light_margin` from the walls and ceiling."""
    rng = np.random.default_rng(seed)
    P = lambda x, y, z: np.array([x, y, z], dtype=np.float64)  # noqa: E731
    walls = [
        _quad(P(-1, 0, -1), P(-1, 0, 1), P(1, 0, 1), P(1, 0, -1)),    # floor
        _quad(P(-1, 2, -1), P(1, 2, -1), P(1, 2, 1), P(-1, 2, 1)),    # ceiling
        _quad(P(-1, 0, -1), P(1, 0, -1), P(1, 2, -1), P(-1, 2, -1)),  # back
        _quad(P(-1, 0, -1), P(-1, 2, -1), P(-1, 2, 1), P(-1, 0, 1)),  # left
        _quad(P(1, 0, -1), P(1, 0, 1), P(1, 2, 1), P(1, 2, -1)),      # right
    ]
    tris, mat = [], []
    for i, w in enumerate(walls):
        tris += w
        mat += [i, i]
    for b, (cx, cz) in enumerate(((-0.4, -0.3), (0.4, 0.35))):
        size = rng.uniform(0.2, 0.5, size=3) * 2.0
        yaw = rng.uniform(0.0, 2.0 * math.pi)
        center = (cx + rng.uniform(-0.1, 0.1), 0.5 * size[1], cz + rng.uniform(-0.1, 0.1))
        bt = box_triangles(center, size, _yaw(yaw))
        tris += list(bt)
        mat += [5 + b] * 12
    tri_xyz = np.array(tris, dtype=np.float32).reshape(-1, 9)
    materials = _materials(rng, 7)
    e = 1.0 - light_margin
    pos = np.stack([rng.uniform(-e, e, n_lights), rng.uniform(1.0, 1.0 + e, n_lights),
                    rng.uniform(-e, e, n_lights)], axis=1)
    rgb = np.exp(rng.uniform(math.log(0.1), math.log(10.0), size=(n_lights, 3)))
    camera = dict(position=(0.0, 1.0, 2.9), lookAt=(0.0, 1.0, 0.0), worldUp=(0.0, 1.0, 0.0),
                  fovY=math.pi / 3.0, zNear=0.01, zFar=1000.0)
    return Scene("cornell_box", tri_xyz, np.array(mat, dtype=np.int32), materials,
                 make_point_lights(pos, rgb), np.zeros(0, dtype=TRI_LIGHT_DTYPE), camera,
                 dict(seed=seed))


def value_noise(rng, xs, zs, cells=8, octaves=3):
    """Sum of smoothstep-interpolated lattice noise octaves over [0,1]^2 coordinates."""
    h = np.zeros_like(xs, dtype=np.float64)
    amp, freq = 1.0, cells
    for _ in range(octaves):
        lat = rng.uniform(0.0, 1.0, size=(freq + 1, freq + 1))
        u, v = xs * freq, zs * freq
        i0 = np.clip(np.floor(u).astype(int), 0, freq - 1)
        j0 = np.clip(np.floor(v).astype(int), 0, freq - 1)
        fu, fv = u - i0, v - j0
        su, sv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
        a = lat[i0, j0] * (1 - su) + lat[i0 + 1, j0] * su
        b = lat[i0, j0 + 1] * (1 - su) + lat[i0 + 1, j0 + 1] * su
        h += amp * (a * (1 - sv) + b * sv)
        amp *= 0.5
        freq *= 2
    return h


def terrain(seed: int = 2, grid: int = 256, n_boxes: int = 500, n_point_lights: int = 10000,
            extent: float = 32.0, height: float = 6.0, light_intensity=(0.01, 10.0),
            n_tri_lights: int = 0, tri_light_emission=(0.01, 10.0)) -> Scene:
    """Config 2 scene: a grid x grid vertex heightfield over [-extent, extent]^2 displaced by
    value noise (2*(grid-1)^2 triangles) plus n_boxes random oriented boxes (12 triangles each);
    point lights uniform in the scene AABB with per-channel log-uniform intensity. With
    n_tri_lights, that many emissive triangles are added as free-floating lamps (from a separate
    random stream, so the rest of the scene is unchanged): random orientation, circumradius
    uniform in [0.3, 1], centres between 1 and 4 units above the highest terrain or box point,
    per-channel log-uniform emission, each with its own emissive material whose albedo is the
    emission. They follow the other triangles and occlude like them."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, grid)
    X, Z = np.meshgrid(t, t, indexing="ij")
    Hh = value_noise(rng, X, Z) * (height / 1.75)
    V = np.stack([(X * 2 - 1) * extent, Hh, (Z * 2 - 1) * extent], axis=-1)
    v00 = V[:-1, :-1].reshape(-1, 3)
    v10 = V[1:, :-1].reshape(-1, 3)
    v01 = V[:-1, 1:].reshape(-1, 3)
    v11 = V[1:, 1:].reshape(-1, 3)
    t1 = np.concatenate([v00, v01, v11], axis=1)
    t2 = np.concatenate([v00, v11, v10], axis=1)
    terrain_tris = np.stack([t1, t2], axis=1).reshape(-1, 9)
    boxes, bmat = [], []
    for b in range(n_boxes):
        cx, cz = rng.uniform(-0.9 * extent, 0.9 * extent, size=2)
        gi = int(round((cx / extent + 1) * 0.5 * (grid - 1)))
        gj = int(round((cz / extent + 1) * 0.5 * (grid - 1)))
        size = rng.uniform(0.5, 3.0, size=3)
        cy = Hh[gi, gj] + rng.uniform(0.0, 2.0)
        boxes.append(box_triangles((cx, cy, cz), size, _random_rotation(rng)))
        bmat += [1 + b] * 12
    tri_xyz = np.concatenate([terrain_tris.astype(np.float32)] +
                             [bx.reshape(-1, 9) for bx in boxes]).astype(np.float32)
    tri_material = np.concatenate([np.zeros(len(terrain_tris), dtype=np.int32),
                                   np.array(bmat, dtype=np.int32)])
    materials = _materials(rng, 1 + n_boxes)
    lo, hi = tri_xyz.reshape(-1, 3).min(axis=0), tri_xyz.reshape(-1, 3).max(axis=0)
    pos = rng.uniform(lo, hi, size=(n_point_lights, 3))
    a, b = light_intensity
    rgb = np.exp(rng.uniform(math.log(a), math.log(b), size=(n_point_lights, 3)))
    camera = dict(position=(0.0, 0.8 * extent, 1.4 * extent), lookAt=(0.0, 0.0, 0.0),
                  worldUp=(0.0, 1.0, 0.0), fovY=math.pi / 3.0, zNear=0.01, zFar=1000.0)
    tri_lights = np.zeros(0, dtype=TRI_LIGHT_DTYPE)
    if n_tri_lights:
        lrng = np.random.default_rng([seed, 7])
        top = float(hi[1])  # above every terrain and box triangle
        centres = np.stack([lrng.uniform(-0.9 * extent, 0.9 * extent, n_tri_lights),
                            lrng.uniform(top + 1.0, top + 4.0, n_tri_lights),
                            lrng.uniform(-0.9 * extent, 0.9 * extent, n_tri_lights)], axis=1)
        radius = lrng.uniform(0.3, 1.0, n_tri_lights)
        corner = np.array([[1.0, 0.0, 0.0], [-0.5, 0.0, 0.75 ** 0.5], [-0.5, 0.0, -(0.75 ** 0.5)]])
        lamps = np.stack([centres[i] + radius[i] * (corner @ _random_rotation(lrng).T)
                          for i in range(n_tri_lights)]).reshape(-1, 9).astype(np.float32)
        ea, eb = tri_light_emission
        emission = np.exp(lrng.uniform(math.log(ea), math.log(eb), size=(n_tri_lights, 3)))
        tri_lights = make_tri_lights(lamps, emission)
        em = np.zeros((n_tri_lights, materials.shape[1]), dtype=np.float32)
        em[:, 0:3] = emission
        em[:, 3] = 1.0  # emissive flag
        em[:, 4] = 1.0  # roughness
        tri_material = np.concatenate([tri_material, len(materials) + np.arange(n_tri_lights, dtype=np.int32)])
        materials = np.concatenate([materials, em])
        tri_xyz = np.concatenate([tri_xyz, lamps])
    return Scene("terrain", tri_xyz, tri_material, materials, make_point_lights(pos, rgb),
                 tri_lights, camera,
                 dict(seed=seed, grid=grid, n_boxes=n_boxes, n_tri_lights=n_tri_lights))


def orbit_camera(camera: dict, frame: int, degrees_per_frame: float = 0.5) -> dict:
    """Rotate the camera position about the world y axis through lookAt by frame*deg."""
    th = math.radians(degrees_per_frame * frame)
    px, py, pz = camera["position"]
    lx, ly, lz = camera["lookAt"]
    dx, dz = px - lx, pz - lz
    c, s = math.cos(th), math.sin(th)
    cam = dict(camera)
    cam["position"] = (lx + c * dx + s * dz, py, lz - s * dx + c * dz)
    return cam


def batched_view_set(seed: int = 3):
    """The batched-views workload: 4 seeded terrains (grids 110..300, 60..1000 boxes: about
    25k-200k triangles; 1k-50k point lights; a third of the box count times 4 emissive lamp
    triangles, in one alias table with the point lights) and 16 views over them (view v shows
    scene v mod 4) at 16:9 resolutions with heights log-uniform in [360, 1440]. Returns
    (scenes, [(scene, width, height)] * 16)."""
    rng = np.random.default_rng(seed)
    specs = [(110, 60, 1000), (160, 200, 5000), (220, 400, 20000), (300, 1000, 50000)]
    scs = [terrain(20 + i, grid=g, n_boxes=b, n_point_lights=nl, extent=32.0, n_tri_lights=4 * b)
           for i, (g, b, nl) in enumerate(specs)]
    views = []
    for v in range(16):
        h = int(round(360 * 4.0 ** rng.random()))
        views.append((scs[v % 4], int(round(h * 16 / 9)), h))
    return scs, views


def stress_scene() -> Scene:
    """The 4K stress workload's scene: a 1024 x 1024-vertex heightfield plus 20,000 boxes
    (2 * 1023^2 + 240,000 = 2,333,058 triangles) and 1M point lights with log-uniform intensity
    in [1e-3, 10]."""
    return terrain(5, grid=1024, n_boxes=20000, n_point_lights=1_000_000, light_intensity=(1e-3, 10.0))


def config_scene(name: str) -> Scene:
    if name in ("c1", "cornell"):
        return cornell_box(1)
    if name in ("c2", "terrain"):
        return terrain(2, n_tri_lights=2000)
    raise ValueError(name)
