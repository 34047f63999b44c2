"""NeuSModel (models/neus.py:46-318), registered as 'neus', with
VarianceNetwork (:15-43) and an occupancy grid in the role of nerfacc 0.3.3's
OccupancyGrid (:63-68,108-111).

forward(rays) returns the reference's output dict (:242-293).  The training
path is: ray-AABB + occupancy marching (HIP) -> sample positions (HIP) ->
fused VolumeSDF (FD normal + curvature, HIP) -> fused VolumeRadiance (HIP)
-> fused alpha / transmittance / accumulation (HIP).
"""
import math

import torch
import torch.nn as nn
import torch.nn.functional as F

from . import _native as nat
from . import descs, make, register
from . import functions as fn
from .utils import chunk_batch, update_module_step


class VarianceNetwork(nn.Module):
    """models/neus.py:15-43."""

    def __init__(self, config):
        super().__init__()
        self.config = config
        self.init_val = config.init_val
        self.register_parameter("variance", nn.Parameter(torch.tensor(float(config.init_val))))
        self.modulate = config.get("modulate", False)
        if self.modulate:
            self.mod_start_steps = config.mod_start_steps
            self.reach_max_steps = config.reach_max_steps
            self.max_inv_s = config.max_inv_s
        self.do_mod = False

    @property
    def inv_s(self):
        val = torch.exp(self.variance * 10.0)
        if self.modulate and self.do_mod:
            val = val.clamp_max(self.mod_val)
        return val

    def forward(self, x):
        return torch.ones([len(x), 1], device=self.variance.device) * self.inv_s

    def update_step(self, epoch, global_step):
        if self.modulate:
            self.do_mod = global_step > self.mod_start_steps
            if not self.do_mod:
                self.prev_inv_s = self.inv_s.item()
            else:
                self.mod_val = min((global_step / self.reach_max_steps) * (self.max_inv_s - self.prev_inv_s)
                                   + self.prev_inv_s, self.max_inv_s)


class OccupancyGrid(nn.Module):
    """Occupancy grid of nerfacc 0.3.3 (AABB contraction): occs [R^3] float32
    and a uint8 binary [R,R,R] (x-major), updated every n steps by
    am_occ_update; the random cells / jitter are drawn with torch."""

    def __init__(self, roi_aabb, resolution=128):
        super().__init__()
        self.resolution = int(resolution)
        self.register_buffer("roi_aabb", torch.as_tensor(roi_aabb, dtype=torch.float32))
        n = self.resolution ** 3
        self.num_cells = n
        self.register_buffer("occs", torch.zeros(n))
        self.register_buffer("binary", torch.ones(self.resolution, self.resolution, self.resolution,
                                                  dtype=torch.uint8))

    @torch.no_grad()
    def every_n_step(self, step, update_fn, occ_thre=1e-2, ema_decay=0.95, warmup_steps=256, n=16):
        if step % n == 0:
            self.update(step, update_fn, occ_thre, ema_decay, warmup_steps)

    @torch.no_grad()
    def update(self, step, update_fn, occ_thre=1e-2, ema_decay=0.95, warmup_steps=256, gen=None):
        dev = self.occs.device
        if step < warmup_steps:
            cells = None
            m = self.num_cells
        else:
            N = self.num_cells // 4
            uniform = torch.randint(self.num_cells, (N,), device=dev, generator=gen)
            occupied = torch.nonzero(self.binary.flatten())[:, 0]
            if N < len(occupied):
                sel = torch.randint(len(occupied), (N,), device=dev, generator=gen)
                occupied = occupied[sel]
            cells = torch.cat([uniform, occupied]).to(torch.int32)
            m = cells.numel()
        jitter = torch.rand(m, 3, device=dev, generator=gen)
        update_fn(cells, jitter, m, occ_thre, ema_decay)


@register("neus")
class NeuSModel(nn.Module):
    def __init__(self, config):
        super().__init__()
        self.config = config
        self.setup()
        if config.get("weights", None):
            self.load_state_dict(torch.load(config.weights, weights_only=True))

    def setup(self):
        cfg = self.config
        self.geometry = make(cfg.geometry.name, cfg.geometry)
        self.texture = make(cfg.texture.name, cfg.texture)
        if cfg.get("learned_background", False):
            raise NotImplementedError("learned_background (forward_bg_) is not part of this build")
        self.variance = VarianceNetwork(cfg.variance)
        r = float(cfg.radius)
        self.register_buffer("scene_aabb", torch.as_tensor([-r, -r, -r, r, r, r], dtype=torch.float32))
        if cfg.get("grid_prune", False):
            self.occupancy_grid = OccupancyGrid(self.scene_aabb, resolution=cfg.get("grid_resolution", 128))
        self.randomized = cfg.randomized
        self.background_color = None
        step = cfg.get("render_step_size", None)
        self.render_step_size = float(step) if step is not None else 1.732 * 2 * r / cfg.num_samples_per_ray
        self.max_samples_per_ray = int(cfg.get("max_samples_per_ray", 0))
        self.cos_anneal_ratio = 1.0
        self.kahan = int(cfg.get("kahan", 0))
        self.precision = "fp32"
        self.set_precision(cfg.get("precision", "fp32"))

    def set_precision(self, precision):
        """"fp32": the plain fp32 form.  "mixed": fp16 table rows, bf16 MFMA
        matmuls with fp32 accumulation (SDF and radiance MLPs), fp16
        table-gradient pairs summed in fp32 (the trainer applies dynamic loss
        scaling), and Kahan-compensated transmittance and compositing sums."""
        self.geometry.set_precision(precision)
        self.texture.set_precision(precision)
        self.precision = precision
        self.kahan = 1 if precision == "mixed" else int(self.config.get("kahan", 0))

    # -- per-step protocol (models/neus.py:79-111) --------------------------
    def update_step(self, epoch, global_step):
        update_module_step(self.geometry, epoch, global_step)
        update_module_step(self.texture, epoch, global_step)
        update_module_step(self.variance, epoch, global_step)
        cos_anneal_end = self.config.get("cos_anneal_end", 0)
        self.cos_anneal_ratio = 1.0 if cos_anneal_end == 0 else min(1.0, global_step / cos_anneal_end)
        if self.training and self.config.get("grid_prune", False):
            self.occupancy_grid.every_n_step(step=global_step, update_fn=self._occ_update,
                                             occ_thre=self.config.get("grid_prune_occ_thre", 0.01))

    def inv_s(self):
        return self.variance(torch.zeros([1, 3]))[:, :1].clip(1e-6, 1e6)

    @torch.no_grad()
    def _occ_update(self, cells, jitter, m, occ_thre, ema_decay):
        """occ_eval_fn (models/neus.py:90-101) + the EMA / threshold update."""
        geo = self.geometry
        g = geo.grid_module
        og = self.occupancy_grid
        sd = geo._sdf_desc(False, False)
        lib = nat.load()
        inv_s = self.inv_s().reshape(1).float().contiguous()
        ws = nat.workspace(lib.am_occ_update_workspace_size(og.resolution, m), og.occs.device)
        params = geo.network.flat_params().detach().contiguous()
        nat.check(lib.am_occ_update(nat.sptr(g.grid), nat.sptr(geo.network.desc), nat.sptr(sd),
                                    nat.ptr(g.kernel_table()), nat.ptr(params), nat.ptr(inv_s),
                                    float(self.render_step_size), nat.ptr(cells),
                                    nat.ptr(jitter.contiguous()), m, og.resolution,
                                    float(ema_decay), float(occ_thre), nat.ptr(og.occs),
                                    nat.ptr(og.binary), nat.ptr(ws), ws.numel(), nat.stream()),
                  "am_occ_update")

    def get_alpha(self, sdf, normal, dirs, dists):
        """models/neus.py:117-139 (torch form, for callers outside the fused
        compositing; the training path computes alpha inside am_composite)."""
        inv_s = self.inv_s().expand(sdf.shape[0], 1)
        true_cos = (dirs * normal).sum(-1, keepdim=True)
        r = self.cos_anneal_ratio
        iter_cos = -(F.relu(-true_cos * 0.5 + 0.5) * (1.0 - r) + F.relu(-true_cos) * r)
        next_sdf = sdf[..., None] + iter_cos * dists.reshape(-1, 1) * 0.5
        prev_sdf = sdf[..., None] - iter_cos * dists.reshape(-1, 1) * 0.5
        prev_cdf = torch.sigmoid(prev_sdf * inv_s)
        next_cdf = torch.sigmoid(next_sdf * inv_s)
        return ((prev_cdf - next_cdf + 1e-5) / (prev_cdf + 1e-5)).view(-1).clip(0.0, 1.0)

    # -- sampling ---------------------------------------------------------
    def march(self, rays):
        """nerfacc ray_marching with the scene AABB and the occupancy grid."""
        aabb = self.scene_aabb.tolist()
        t_near, t_far = fn.ray_aabb(rays, aabb)
        jitter = torch.rand_like(t_near) if self.randomized else None
        grid = self.occupancy_grid.binary if self.config.get("grid_prune", False) else None
        roi = aabb
        return fn.ray_marching(rays, t_near, t_far, roi, grid, self.render_step_size, jitter,
                               self.max_samples_per_ray)

    # -- forward ------------------------------------------------------------
    def forward_(self, rays, samples=None, randn=None):
        """models/neus.py:205-283.  samples: optional (ray_indices int32,
        t_starts, t_ends, packed) replacing marching."""
        rays = rays.float().contiguous()
        n_rays = rays.shape[0]
        if samples is None:
            with torch.no_grad():
                ray_indices, t_starts, t_ends, packed = self.march(rays)
        else:
            ray_indices, t_starts, t_ends, packed = samples
        pts, mid, dists = fn.sample_points(rays, ray_indices, t_starts, t_ends)
        geo = self.geometry
        sdf, sdf_grad, normal, feature, lap = geo.fused(pts, with_grad=True, with_laplace=True,
                                                        randn=randn)
        rgb = self.texture.fused(feature, normal, rays, ray_indices)
        cd = nat.CompositeDesc()
        cd.cos_anneal_ratio = float(self.cos_anneal_ratio)
        cd.kahan = self.kahan
        weights, opacity, depth, comp_rgb, comp_normal = fn.composite(
            sdf, normal, rgb, self.inv_s().reshape(1), rays, ray_indices, t_starts, t_ends, packed, cd)
        opacity = opacity[:, None]
        depth = depth[:, None]
        n = ray_indices.shape[0]
        out = {
            "comp_rgb": comp_rgb,
            "comp_normal": comp_normal,
            "opacity": opacity,
            "depth": depth,
            "rays_valid": opacity > 0,
            "num_samples": torch.as_tensor([n], dtype=torch.int32, device=rays.device),
        }
        if self.training:
            out.update({
                "sdf_samples": sdf,
                "sdf_grad_samples": sdf_grad,
                "weights": weights.view(-1),
                "points": mid.view(-1),
                "intervals": dists.view(-1),
                "ray_indices": ray_indices.long().view(-1),
                "sdf_laplace_samples": lap.view(-1, 1),
            })
        bg = self.background_color
        if bg is None:
            bg = torch.ones(3, device=rays.device)
        out_bg = {
            "comp_rgb": bg[None, :].expand(*comp_rgb.shape),
            "num_samples": torch.zeros_like(out["num_samples"]),
            "rays_valid": torch.zeros_like(out["rays_valid"]),
        }
        out_full = {
            "comp_rgb": out["comp_rgb"] + out_bg["comp_rgb"] * (1.0 - out["opacity"]),
            "num_samples": out["num_samples"] + out_bg["num_samples"],
            "rays_valid": out["rays_valid"] | out_bg["rays_valid"],
        }
        return {**out, **{k + "_bg": v for k, v in out_bg.items()},
                **{k + "_full": v for k, v in out_full.items()}}

    def forward(self, rays, samples=None, randn=None):
        """models/neus.py:285-293."""
        if self.training:
            out = self.forward_(rays, samples, randn)
        else:
            out = chunk_batch(lambda r: self.forward_(r), self.config.ray_chunk, False, rays) \
                if samples is None else self.forward_(rays, samples, randn)
        return {**out, "inv_s": self.variance.inv_s}

    def train(self, mode=True):
        self.randomized = mode and self.config.randomized
        return super().train(mode=mode)

    def eval(self):
        self.randomized = False
        return super().eval()

    def regularizations(self, out):
        losses = {}
        losses.update(self.geometry.regularizations(out))
        losses.update(self.texture.regularizations(out))
        return losses
