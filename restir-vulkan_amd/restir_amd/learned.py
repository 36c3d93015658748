"""Device state of the learned light proposal: k-means light clusters with per-cluster alias
tables (built by the library's host builder, or with its nearest-centroid step on the GPU), the MLP parameters (bf16 working copy, fp32 master
copy, Adam moments) and per-resolution candidate buffers. All compute runs in librestir_amd.so.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .abi import LightClusters, MlpStruct, mlp_sizes


class LearnedProposal:
    def __init__(self, ctx: abi.Context, scene, k: int = 256, seed: int = 7, iterations: int = 16,
                 mix: float = 0.1, device="cuda", device_clustering: bool = False):
        import torch
        self.torch = torch
        self.ctx, self.k, self.mix = ctx, k, mix
        # the clusters cover the alias table's index space: with one table over both kinds of
        # lights, the triangle lights follow the point lights as (centroid, luminance * area)
        union = abi.uses_union(scene.point_lights, scene.tri_lights)
        records = abi.union_light_records(scene.point_lights, scene.tri_lights) if union else scene.point_lights
        cl = LightClusters(records, k, seed=seed, iterations=iterations, ctx=ctx if device_clustering else None)
        self.clusters = cl
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        self.d_light_cluster = t(cl.light_cluster)
        self.d_light_pwithin = t(cl.light_pwithin)
        self.d_begin = t(cl.begin)
        self.d_count = t(cl.count)
        self.d_lights = t(cl.lights)
        self.d_alias = t(cl.alias.view(np.uint8))
        self.bind()
        nW, nB = mlp_sizes(k)
        self.w_bf16 = torch.zeros(nW, dtype=torch.int16, device=device)
        self.w = torch.zeros(nW, dtype=torch.float32, device=device)
        self.b = torch.zeros(nB, dtype=torch.float32, device=device)
        self.m = torch.zeros(nW + nB, dtype=torch.float32, device=device)
        self.v = torch.zeros(nW + nB, dtype=torch.float32, device=device)
        # the Adam step count lives on the device, so training steps can be graph-captured
        self.step = torch.zeros(1, dtype=torch.int32, device=device)
        self.mlp = MlpStruct(k, 0, self.w_bf16.data_ptr(), self.w.data_ptr(), self.b.data_ptr(),
                             self.m.data_ptr(), self.v.data_ptr(), self.step.data_ptr())
        # output bias starts at the log power share of each cluster; empty clusters are masked
        share = cl.power_share.astype(np.float64)
        prior = np.where(share > 0, np.log(np.maximum(share, 1e-30)), -30000.0).astype(np.float32)
        self.d_prior = t(prior)
        ctx.mlp_init(self.mlp, seed, self.d_prior)
        lo, hi = scene.aabb()
        pad = 1e-3 * (hi - lo) + 1e-3
        self.lo, self.hi = (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)
        self.cand = {}
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.train_seed = seed * 7919

    def bind(self):
        """(Re)binds the clusters to the context: restir_set_lights unbinds them, so a renderer
        that sets the scene's lights on a shared context calls this afterwards."""
        self.ctx.set_light_clusters(self.k, self.d_light_cluster, self.d_light_pwithin, self.d_begin,
                                    self.d_count, self.d_lights, self.d_alias)

    def buffers(self, npx: int, M: int):
        key = (npx, M)
        if key not in self.cand:
            t = self.torch
            self.cand[key] = (t.zeros(M * npx, dtype=t.int32, device=self.w.device),
                              t.zeros(M * npx, dtype=t.float32, device=self.w.device))
        return self.cand[key]

    def propose(self, u, g, M: int, stream=None):
        cl, cq = self.buffers(g.width * g.height, M)
        self.ctx.learned_candidates(u, g, self.mlp, self.lo, self.hi, M, self.mix, cl, cq, stream)
        return cl, cq

    def train(self, u, g, res, n=4096, lr=1e-3, stream=None):
        self.train_seed += 1
        self.ctx.mlp_train(self.mlp, u, g, res, self.lo, self.hi, n, lr, self.train_seed & 0xffffffff,
                           self.loss, stream)

    def state(self):
        """Parameters and Adam moments as host arrays (the proposal's checkpoint)."""
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy(), "m": self.m.cpu().numpy(),
                "v": self.v.cpu().numpy(), "w_bf16": self.w_bf16.cpu().numpy(),
                "adam_step": int(self.step.item()), "train_seed": self.train_seed}

    def load_state(self, st):
        t = self.torch
        self.w.copy_(t.from_numpy(st["w"]))
        self.b.copy_(t.from_numpy(st["b"]))
        self.m.copy_(t.from_numpy(st["m"]))
        self.v.copy_(t.from_numpy(st["v"]))
        if "w_bf16" in st:
            self.w_bf16.copy_(t.from_numpy(st["w_bf16"]))
        else:
            bf = t.from_numpy(st["w"]).to(self.w.device).to(t.bfloat16)
            self.w_bf16.copy_(bf.view(t.int16))
        self.mlp.adam_step = int(st["adam_step"])
        self.step.fill_(int(st["adam_step"]))
        if "train_seed" in st:
            self.train_seed = int(st["train_seed"])


def log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=size))
