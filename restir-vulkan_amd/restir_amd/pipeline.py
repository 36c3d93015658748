"""Frame driver over the C ABI: the pass ordering and buffer ping-pong of the reference's
``App::_recordMainCommandBuffers`` / ``_updateRestirBuffers`` (src/app.h:212-336) and the
per-frame uniform update of ``App::mainLoop`` (src/app.cpp:775-834).

Unbiased mode: Pass A writes a temporary buffer, Pass C writes the frame's reservoir buffer
(app.h:201-215). With several Pass C iterations the buffers alternate so the last one lands in
the frame's buffer. Biased mode: Pass A writes res[i], then pairs of Pass B ping-pong
res[i] -> res[prev] -> res[i] (app.h:248-256, 224-233).

Device memory comes from torch (plumbing only); every computation runs in librestir_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import (GBUFFER_PLANES, RESERVOIR_DTYPE, Context, as_ptr, build_aabb_tree,
                  build_alias_table, gbuffer_struct, make_uniforms, pack_buffer, projection_view)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("restir_amd needs a HIP device: torch.cuda.is_available() is False")
    return torch


def to_device(arr: np.ndarray, torch, device):
    flat = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
    return torch.from_numpy(flat.copy()).to(device)


class SceneOnDevice:
    """BVH, triangles, light tables and materials resident in HBM and bound to a context.

    ``build="host"`` builds the tree on the host (restir_build_aabb_tree) and the alias table
    (restir_build_alias_table) and uploads them; ``build="device"`` uploads the vertices and light
    records and builds the same tree (restir_build_bvh) and, for one kind of lights, the same
    alias table (restir_build_alias_table_device) in HBM; ``nodes`` / ``tris`` / ``alias`` are
    then read back on first access, for checking. A table over both kinds of lights (the union)
    is built on the host in both modes."""

    def __init__(self, ctx: Context, scene, device="cuda", build="host", union_lights=None):
        torch = _torch()
        self.scene = scene
        if build not in ("host", "device"):
            raise ValueError(f"build must be 'host' or 'device', not {build!r}")
        n = scene.n_tris
        if build == "host":
            self._nodes, self._tris, self.depth = build_aabb_tree(scene.tri_xyz)
            self.d_nodes = to_device(self._nodes, torch, device)
            self.d_tris = to_device(self._tris, torch, device)
            ctx.set_bvh(self.d_nodes, len(self._nodes), self.d_tris, len(self._tris))
        else:
            self._nodes = self._tris = None
            self.d_xyz = torch.from_numpy(np.ascontiguousarray(scene.tri_xyz, dtype=np.float32)).to(device)
            self.d_nodes = torch.empty(abi.NODE_DTYPE.itemsize * (n - 1), dtype=torch.uint8, device=device)
            self.d_tris = torch.empty(abi.TRIANGLE_DTYPE.itemsize * n, dtype=torch.uint8, device=device)
            self.depth = ctx.build_bvh(self.d_xyz, n, self.d_nodes, self.d_tris)
        pl = pack_buffer(scene.point_lights, abi.POINT_LIGHT_DTYPE)
        tl = pack_buffer(scene.tri_lights, abi.TRI_LIGHT_DTYPE)
        self.d_points = to_device(pl, torch, device)
        self.d_trilights = to_device(tl, torch, device)
        # with both kinds of lights, one table over both (abi.uses_union) unless
        # union_lights=False asks for the reference's point-light table
        self.union = abi.uses_union(scene.point_lights, scene.tri_lights, union_lights)
        hdr = abi.LIGHT_BUFFER_HEADER_BYTES
        if build == "device" and not self.union:
            # the reference's table (point lights if any, else triangle lights) built in HBM
            n_pt, n_tl = len(scene.point_lights), len(scene.tri_lights)
            n_al = n_pt if n_pt else n_tl
            al = pack_buffer(np.zeros(n_al, dtype=abi.ALIAS_DTYPE), abi.ALIAS_DTYPE)  # header + space
            self.d_alias = to_device(al, torch, device)
            if n_al:
                ctx.build_alias_table(self.d_points.data_ptr() + hdr if n_pt else None, n_pt,
                                      None if n_pt else self.d_trilights.data_ptr() + hdr, 0 if n_pt else n_tl,
                                      self.d_alias.data_ptr() + hdr)
            self._alias = None
        else:
            self._alias = abi.scene_alias_table(scene.point_lights, scene.tri_lights, union_lights)
            al = pack_buffer(self._alias, abi.ALIAS_DTYPE)
            self.d_alias = to_device(al, torch, device)
        self.d_tri_material = to_device(scene.tri_material.astype(np.int32), torch, device)
        self.d_materials = to_device(scene.materials.astype(np.float32), torch, device)
        ctx.set_lights(self.d_points, pl.nbytes, self.d_trilights, tl.nbytes, self.d_alias, al.nbytes)

    @property
    def nodes(self) -> np.ndarray:
        if self._nodes is None:
            self._nodes = self.d_nodes.cpu().numpy().view(abi.NODE_DTYPE)
        return self._nodes

    @property
    def tris(self) -> np.ndarray:
        if self._tris is None:
            self._tris = self.d_tris.cpu().numpy().view(abi.TRIANGLE_DTYPE)
        return self._tris

    @property
    def alias(self) -> np.ndarray:
        if self._alias is None:
            hdr = abi.LIGHT_BUFFER_HEADER_BYTES
            self._alias = self.d_alias[hdr:].cpu().numpy().view(abi.ALIAS_DTYPE)
        return self._alias


class Renderer:
    """Runs frames of the ReSTIR pipeline on one device.

    mode "unbiased": A -> C x spatial_passes (k neighbours) ; mode "biased": A -> (B, B) x
    spatial_iterations. ``launch`` selects the workgroup tile and Pass C schedule
    (restir_launch) of every pass; ``launches`` ({"a"|"b"|"c": Launch}) overrides it per pass, as
    restir_amd.tuning.autotune chooses. ``scene_build`` ("host" or "device") selects where the
    tree is built (SceneOnDevice).
    """

    def __init__(self, scene, width, height, *, device_index=0, mode="unbiased", M=32, k=5,
                 spatial_passes=1, spatial_neighbors=4, radius=30.0, temporal=True,
                 visibility=True, temporal_mult=20, launch=None, ctx=None, learned=None,
                 train_samples=4096, lr=1e-3, train=True, launches=None, scene_build="host",
                 normal_format=abi.NORMAL_F32, standard_weights=False, shade_visibility=False,
                 reuse_verdicts=True):
        torch = _torch()
        self.torch = torch
        self.device = torch.device("cuda", device_index)
        self.ctx = ctx or Context(device_index)
        self.scene = scene
        self.dev = SceneOnDevice(self.ctx, scene, self.device, build=scene_build)
        self.W, self.H = width, height
        self.mode, self.M, self.k = mode, M, k
        self.spatial_passes = spatial_passes
        self.spatial_neighbors = spatial_neighbors
        self.radius = radius
        self.temporal_mult = temporal_mult
        self.flags = (abi.VISIBILITY_REUSE_FLAG if visibility else 0) | \
                     (abi.TEMPORAL_REUSE_FLAG if temporal else 0) | \
                     (abi.STANDARD_WEIGHTS_FLAG if standard_weights else 0)
        self.launch = launch
        self.normal_format = normal_format  # G-buffer normals as float or half planes
        self.shade_visibility = shade_visibility  # final shading re-tests visibility
        # the second and later Pass C do not retrace the segments the previous one found
        # unoccluded (RESTIR_INPUT_VISIBLE_FLAG; the same results)
        self.reuse_verdicts = reuse_verdicts
        self.launches = {"a": launch, "b": launch, "c": launch}
        self.launches.update(launches or {})
        # learned proposal: a restir_amd.learned.LearnedProposal bound to this context, or None
        # (setting the scene's lights above unbound its clusters)
        self.learned = learned
        if learned is not None:
            learned.bind()
        self.train_samples, self.lr, self.train_enabled = train_samples, lr, train
        n = width * height
        self.gbuf = [torch.zeros(GBUFFER_PLANES * n, dtype=torch.float32, device=self.device)
                     for _ in range(2)]
        self.res = [torch.zeros(16 * n, dtype=torch.int32, device=self.device) for _ in range(2)]
        self.tmp = torch.zeros(16 * n, dtype=torch.int32, device=self.device)
        self.rgb = torch.zeros(3 * n, dtype=torch.float32, device=self.device)
        self.frame = 0
        self.cur = 1  # index of the buffers written by the last frame
        self.prev_pv = None
        self.cam = None

    # -------------------------------------------------------------------------------- pieces
    def gstruct(self, i):
        return gbuffer_struct(self.gbuf[i].data_ptr(), self.W, self.H, normal_format=self.normal_format)

    def render_gbuffer(self, cam: dict, i: int, stream=None):
        self.ctx.primary_gbuffer(cam, self.dev.d_tri_material, self.dev.d_materials,
                                 len(self.scene.materials), self.gbuf[i], self.W, self.H,
                                 stream=stream, normal_format=self.normal_format)

    def uniforms(self, cam: dict):
        pv_now = projection_view(cam, self.W / self.H)
        prev_pv = self.prev_pv if self.prev_pv is not None else pv_now
        return make_uniforms(self.W, self.H, self.frame, cam["position"], prev_pv, M=self.M,
                             temporal_mult=self.temporal_mult, neighbors=self.spatial_neighbors,
                             radius=self.radius, flags=self.flags), pv_now

    def begin_frame(self, cam: dict, stream=None, gbuffer=True):
        """Advance the frame counter and render this frame's G-buffer (the raster pass's slot)."""
        self.cur = 1 - self.cur
        self.frame += 1
        self.cam = cam
        if gbuffer:
            self.render_gbuffer(cam, self.cur, stream)

    def restir_passes(self, u, stream=None):
        """The ReSTIR dispatches of one frame (the hot path)."""
        i, p = self.cur, 1 - self.cur
        g, gp = self.gstruct(i), self.gstruct(p)

        def pass_a(out):
            if self.learned is not None:
                cl, cq = self.learned.propose(u, g, self.M, stream)
                self.ctx.initial_temporal_candidates(u, g, gp, self.res[p], out, cl, cq,
                                                     self.launches["a"], stream)
            else:
                self.ctx.initial_temporal(u, g, gp, self.res[p], out, self.launches["a"], stream)

        if self.mode == "unbiased":
            n = self.spatial_passes
            first = self.tmp if n % 2 == 1 else self.res[i]
            pass_a(first)
            src = first
            later = u
            if self.reuse_verdicts and (u.flags & abi.VISIBILITY_REUSE_FLAG):
                later = abi.Uniforms.from_buffer_copy(u)
                later.flags |= abi.INPUT_VISIBLE_FLAG
            for j in range(n):
                dst = self.res[i] if (n - j) % 2 == 1 else self.tmp
                self.ctx.unbiased(later if j > 0 else u, self.k, j, g, src, dst, self.launches["c"], stream)
                src = dst
        else:
            pass_a(self.res[i])
            for j in range(self.spatial_passes):
                self.ctx.spatial(u, 2 * j, g, self.res[i], self.res[p], self.launches["b"], stream)
                self.ctx.spatial(u, 2 * j + 1, g, self.res[p], self.res[i], self.launches["b"],
                                 stream)

    def shade(self, stream=None, gamma=1.0):
        self.ctx.shade(self.cam["position"], gamma, self.gstruct(self.cur), self.res[self.cur],
                       self.rgb, self.W * self.H, stream, visibility=self.shade_visibility)

    def train_step(self, u, stream=None):
        """One online step of the learned proposal on this frame's final reservoirs."""
        if self.learned is not None and self.train_enabled:
            self.learned.train(u, self.gstruct(self.cur), self.res[self.cur], self.train_samples,
                               self.lr, stream)

    def render_frame(self, cam: dict, stream=None):
        self.begin_frame(cam, stream)
        u, pv_now = self.uniforms(cam)
        self.restir_passes(u, stream)
        self.shade(stream)
        self.train_step(u, stream)
        self.prev_pv = pv_now
        return self.rgb

    def reset(self):
        """Back to the state before the first frame: frame counter 0, no temporal history
        (zeroed reservoirs, G-buffers and image, cleared with fill kernels on the current stream)."""
        for t in (*self.gbuf, *self.res, self.tmp, self.rgb):
            t.zero_()
        self.frame, self.cur, self.prev_pv, self.cam = 0, 1, None, None

    # ------------------------------------------------------------------------------ readback
    def image(self) -> np.ndarray:
        return self.rgb.view(3, self.H, self.W).cpu().numpy()

    def reservoirs(self, which=None) -> np.ndarray:
        t = self.res[self.cur] if which is None else which
        return t.cpu().numpy().view(RESERVOIR_DTYPE)

    def gbuffer_host(self, i=None) -> np.ndarray:
        """The 13 G-buffer planes as float32 on the host (half normals decoded to float, exactly)."""
        return float_planes(self.gbuf[self.cur if i is None else i].cpu().numpy(), self.W * self.H,
                            self.normal_format)


def float_planes(planes: np.ndarray, npx: int, normal_format: int) -> np.ndarray:
    """A host G-buffer (13 float32 planes of npx) with its normals as float planes: half normals
    (stored in the first half of the normal region) are widened, which is exact."""
    if normal_format == abi.NORMAL_F32:
        return planes
    out = planes.copy()
    halves = planes[3 * npx: 6 * npx].view(np.float16)[: 3 * npx]
    out[3 * npx: 6 * npx] = halves.astype(np.float32)
    return out


class BatchRenderer:
    """Frames of several views per step. ``views`` is a list of (scene, width, height); views of
    the same Scene object share its device tables (one context per distinct scene). Every view
    has its own G-buffers, reservoirs and image; ``rgb`` holds all images back to back and
    ``images[v]`` is view v's [3 * W * H] slice. Pass B and shading of all views run as one
    persistent launch each (restir_*_views). The passes that trace shadow rays, A and C, run
    either the same way (``rays="persistent"``, the default: one launch per pass over all views,
    rays traced per pixel) or view by view on the queued ray schedule of the view's scene context
    (``rays="queued"``: each view's rays compacted and traced by the persistent tracer). Both give
    each view's output bit-identical to a Renderer of that view alone.

    On the queued schedule the views of different scenes share nothing (each scene context has its
    own scratch and queues), so with ``concurrent_scenes`` (default) each scene's views run in order
    on a stream of their own, forked from the caller's stream after the frame's start and joined
    back before the multi-view shading: small views and the tails of one scene's launches overlap
    another scene's work. A graph capture records the fork and the join as dependencies."""

    def __init__(self, views, *, device_index=0, mode="unbiased", M=32, k=5, spatial_passes=1,
                 spatial_neighbors=4, radius=30.0, temporal=True, visibility=True, temporal_mult=20,
                 launch=None, rays="persistent", normal_format=abi.NORMAL_F32, reuse_verdicts=True,
                 concurrent_scenes=True):
        torch = _torch()
        self.torch = torch
        self.device = torch.device("cuda", device_index)
        self.ctx = Context(device_index)  # launches the batched passes
        self.mode, self.M, self.k = mode, M, k
        self.spatial_passes, self.spatial_neighbors, self.radius = spatial_passes, spatial_neighbors, radius
        self.temporal_mult = temporal_mult
        self.flags = (abi.VISIBILITY_REUSE_FLAG if visibility else 0) | \
                     (abi.TEMPORAL_REUSE_FLAG if temporal else 0)
        self.launch = launch
        if rays not in ("queued", "persistent"):
            raise ValueError(f"rays must be 'queued' or 'persistent', not {rays!r}")
        self.rays = rays
        self.normal_format = normal_format
        self.reuse_verdicts = reuse_verdicts  # later queued Pass C: RESTIR_INPUT_VISIBLE_FLAG
        tx, ty = (launch.tile_x, launch.tile_y) if launch is not None else (0, 0)
        self.launch_q = abi.Launch(tx, ty, abi.SCHEDULE_QUEUED)
        self.scene_dev = {}
        self.views = []
        total = sum(3 * w * h for _, w, h in views)
        self.rgb = torch.zeros(total, dtype=torch.float32, device=self.device)
        off = 0
        for sc, w, h in views:
            if id(sc) not in self.scene_dev:
                c = Context(device_index)
                self.scene_dev[id(sc)] = (c, SceneOnDevice(c, sc, self.device))
            c, dev = self.scene_dev[id(sc)]
            n = w * h
            self.views.append({
                "scene": sc, "ctx": c, "dev": dev, "W": w, "H": h,
                "gbuf": [torch.zeros(GBUFFER_PLANES * n, dtype=torch.float32, device=self.device) for _ in range(2)],
                "res": [torch.zeros(16 * n, dtype=torch.int32, device=self.device) for _ in range(2)],
                "tmp": torch.zeros(16 * n, dtype=torch.int32, device=self.device),
                "rgb": self.rgb[off: off + 3 * n], "prev_pv": None, "cam": None,
            })
            off += 3 * n
        self.images = [v["rgb"] for v in self.views]
        # view indices per scene context, in order of first appearance; a stream per scene
        groups = {}
        for vi, v in enumerate(self.views):
            groups.setdefault(id(v["ctx"]), []).append(vi)
        self.groups = list(groups.values())
        self.concurrent = concurrent_scenes and rays == "queued" and len(self.groups) > 1
        self.side = [torch.cuda.Stream(device=self.device) for _ in self.groups] if self.concurrent else []
        self.frame = 0
        self.cur = 1

    def reset(self):
        for v in self.views:
            for t in (*v["gbuf"], *v["res"], v["tmp"]):
                t.zero_()
            v["prev_pv"], v["cam"] = None, None
        self.rgb.zero_()
        self.frame, self.cur = 0, 1

    def _g(self, v, i):
        return gbuffer_struct(v["gbuf"][i].data_ptr(), v["W"], v["H"], normal_format=self.normal_format)

    def render_frame(self, cams, stream=None):
        """One frame of every view; cams[v] is view v's camera."""
        self.cur = 1 - self.cur
        self.frame += 1
        i, p = self.cur, 1 - self.cur
        keep = []  # ctypes structures referenced by pointer until the calls return

        def view(v, u, cur, prev, res_in, out, rgb=None):
            gs = [cur, prev]
            keep.extend([u, *gs])
            return abi.View(v["ctx"].h.value, C.pointer(u), C.pointer(cur), C.pointer(prev),
                            as_ptr(res_in), as_ptr(out), as_ptr(rgb) if rgb is not None else None,
                            v["W"] * v["H"])

        us = []
        for v, cam in zip(self.views, cams):
            v["cam"] = cam
            pv_now = projection_view(cam, v["W"] / v["H"])
            prev_pv = v["prev_pv"] if v["prev_pv"] is not None else pv_now
            u = make_uniforms(v["W"], v["H"], self.frame, cam["position"], prev_pv, M=self.M,
                              temporal_mult=self.temporal_mult, neighbors=self.spatial_neighbors,
                              radius=self.radius, flags=self.flags)
            us.append(u)
            v["prev_pv"] = pv_now
        n_sp = self.spatial_passes
        if self.mode == "unbiased":
            first = [v["tmp"] if n_sp % 2 == 1 else v["res"][i] for v in self.views]
        else:
            first = [v["res"][i] for v in self.views]

        def gbuffer(v, s):
            v["ctx"].primary_gbuffer(v["cam"], v["dev"].d_tri_material, v["dev"].d_materials,
                                     len(v["scene"].materials), v["gbuf"][i], v["W"], v["H"], stream=s,
                                     normal_format=self.normal_format)

        def queued_passes(v, u, f, s):
            """Pass A and (unbiased) every Pass C of one view on the queued schedule, on stream s."""
            v["ctx"].initial_temporal(u, self._g(v, i), self._g(v, p), v["res"][p], f, self.launch_q, s)
            if self.mode != "unbiased":
                return
            src = f
            for j in range(n_sp):
                dst = v["res"][i] if (n_sp - j) % 2 == 1 else v["tmp"]
                uj = u
                if j > 0 and self.reuse_verdicts and (u.flags & abi.VISIBILITY_REUSE_FLAG):
                    uj = abi.Uniforms.from_buffer_copy(u)
                    uj.flags |= abi.INPUT_VISIBLE_FLAG
                v["ctx"].unbiased(uj, self.k, j, self._g(v, i), src, dst, self.launch_q, s)
                src = dst

        if self.concurrent:
            # fork: every scene's stream starts after the work already on the caller's stream
            main = stream if stream is not None else self.torch.cuda.current_stream(self.device)
            for g, side in zip(self.groups, self.side):
                side.wait_stream(main)
                for vi in g:
                    gbuffer(self.views[vi], side)
                    queued_passes(self.views[vi], us[vi], first[vi], side)
            for side in self.side:  # join
                main.wait_stream(side)
        else:
            for v in self.views:
                gbuffer(v, stream)
            if self.rays == "persistent":
                va = [view(v, u, self._g(v, i), self._g(v, p), v["res"][p], f)
                      for v, u, f in zip(self.views, us, first)]
                self.ctx.initial_temporal_views(va, self.launch, stream)
                if self.mode == "unbiased":
                    src = first
                    for j in range(n_sp):
                        dst = [v["res"][i] if (n_sp - j) % 2 == 1 else v["tmp"] for v in self.views]
                        vc = [view(v, u, self._g(v, i), self._g(v, i), a, b)
                              for v, u, a, b in zip(self.views, us, src, dst)]
                        self.ctx.unbiased_views(vc, self.k, j, self.launch, stream)
                        src = dst
            else:
                for v, u, f in zip(self.views, us, first):
                    queued_passes(v, u, f, stream)
        if self.mode != "unbiased":
            for j in range(self.spatial_passes):
                vb = [view(v, u, self._g(v, i), self._g(v, i), v["res"][i], v["res"][p])
                      for v, u in zip(self.views, us)]
                self.ctx.spatial_views(vb, 2 * j, self.launch, stream)
                vb = [view(v, u, self._g(v, i), self._g(v, i), v["res"][p], v["res"][i])
                      for v, u in zip(self.views, us)]
                self.ctx.spatial_views(vb, 2 * j + 1, self.launch, stream)
        vs = [view(v, u, self._g(v, i), self._g(v, i), v["res"][i], None, v["rgb"])
              for v, u in zip(self.views, us)]
        self.ctx.shade_views(vs, 1.0, stream)
        return self.rgb


class FrameSequenceGraph:
    """A sequence of consecutive frames of a Renderer or BatchRenderer (G-buffers, ReSTIR passes,
    shading) captured into one HIP graph, so the whole sequence is submitted with a single launch
    (for a BatchRenderer, cams[f] lists every view's camera of frame f).

    The renderer must have rendered one frame eagerly at this size first, so the library's
    scratch already exists (allocation is not allowed under capture, and the library reports it
    as an error rather than capturing it). Capturing records work without running it: the
    renderer's host state advances to the end of the sequence, and each replay runs the frames
    with the uniforms, seeds and buffers recorded at capture (frame f's image is copied into
    ``images[f]``). Nothing in the graph clears memory with hipMemsetAsync: the library clears its
    counters with kernels. A learned proposal is captured with its training steps when its Adam
    step count lives on the device (restir_mlp.adam_step_device, as LearnedProposal keeps it):
    each replay then continues training from the parameters it finds, with the random seeds
    recorded at capture.
    """

    def __init__(self, r: Renderer, cams):
        torch = r.torch
        learned = getattr(r, "learned", None)
        if learned is not None and not learned.mlp.adam_step_device:
            raise ValueError("FrameSequenceGraph: the learned proposal keeps its Adam step count on the host")
        self.r, self.cams = r, list(cams)
        self.images = torch.zeros(len(self.cams), r.rgb.numel(), dtype=torch.float32, device=r.device)
        self.stream = torch.cuda.Stream(device=r.device)
        self.graph = torch.cuda.CUDAGraph()
        self.stream.wait_stream(torch.cuda.current_stream(r.device))
        with torch.cuda.graph(self.graph, stream=self.stream):
            for f, cam in enumerate(self.cams):
                r.render_frame(cam, stream=self.stream)
                self.images[f].copy_(r.rgb)

    def replay(self):
        self.graph.replay()
