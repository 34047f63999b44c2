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


def contract_inv_unbounded(x, roi):
    """Inverse of the unbounded-sphere contraction (nerfacc 0.3.3
    contract_inv, ContractionType.UN_BOUNDED_SPHERE): y = 4x - 2; for |y| > 1,
    y <- y / (|y| (2 - |y|)); then y/2 + 1/2 scaled onto roi.  Defined for
    |x - 1/2| < 1/2 (the contracted ball)."""
    roi = torch.as_tensor(roi, dtype=x.dtype, device=x.device)
    y = x * 4.0 - 2.0
    n = y.norm(dim=-1, keepdim=True)
    outside = n > 1.0
    m = torch.where(outside, n, torch.ones_like(n))
    y = torch.where(outside, y / (m * (2.0 - m)), y)
    u = y * 0.5 + 0.5
    return u * (roi[3:] - roi[:3]) + roi[:3]


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

    @staticmethod
    def due(step, n=16):
        """Whether every_n_step updates the grid at this step."""
        return step % n == 0

    @torch.no_grad()
    def every_n_step(self, step, update_fn, occ_thre=1e-2, ema_decay=0.95, warmup_steps=256, n=16):
        if self.due(step, n):
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
        self.learned_background = bool(cfg.get("learned_background", False))
        if self.learned_background:
            self.geometry_bg = make(cfg.geometry_bg.name, cfg.geometry_bg)
            self.texture_bg = make(cfg.texture_bg.name, cfg.texture_bg)
            self.geometry_bg.contraction_type = "un_bounded_sphere"
            self.near_plane_bg, self.far_plane_bg = 0.1, 1e3
            self.cone_angle_bg = 10 ** (math.log10(self.far_plane_bg) / cfg.num_samples_per_ray_bg) - 1.0
            self.render_step_size_bg = 0.01
            self.max_samples_per_ray_bg = int(cfg.get("max_samples_per_ray_bg", 0))
        self.variance = VarianceNetwork(cfg.variance)
        r = float(cfg.radius)
        self.register_buffer("scene_aabb", torch.as_tensor([-r, -r, -r, r, r, r], dtype=torch.float32))
        self._aabb_host = [-r, -r, -r, r, r, r]   # launch arguments: no device read per step
        if cfg.get("grid_prune", False):
            self.occupancy_grid = OccupancyGrid(self.scene_aabb, resolution=cfg.get("grid_resolution", 128))
            if self.learned_background:
                self.occupancy_grid_bg = OccupancyGrid(self.scene_aabb, resolution=256)
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
        if self.learned_background:
            update_module_step(self.geometry_bg, epoch, global_step)
            update_module_step(self.texture_bg, epoch, global_step)
        update_module_step(self.variance, epoch, global_step)
        cos_anneal_end = self.config.get("cos_anneal_end", 0)
        self.cos_anneal_ratio = 1.0 if cos_anneal_end == 0 else min(1.0, global_step / cos_anneal_end)
        if self.training and self.config.get("grid_prune", False):
            self.occupancy_grid.every_n_step(step=global_step, update_fn=self._occ_update,
                                             occ_thre=self.config.get("grid_prune_occ_thre", 0.01))
            if self.learned_background:
                self.occupancy_grid_bg.every_n_step(
                    step=global_step, update_fn=self._occ_update_bg,
                    occ_thre=self.config.get("grid_prune_occ_thre_bg", 0.01))

    def grid_update_due(self, global_step):
        """Whether update_step(global_step) rewrites the occupancy grids
        (which marching reads)."""
        return (self.training and bool(self.config.get("grid_prune", False))
                and self.occupancy_grid.due(global_step))

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
        ws = nat.held_workspace("occ_update", lib.am_occ_update_workspace_size(og.resolution, m),
                                 og.occs.device)
        params = geo.network.flat_params().detach().contiguous()
        nat.check(lib.am_occ_update(nat.sptr(g.grid), nat.sptr(geo.network.desc), nat.sptr(sd),
                                    nat.ptr(g.kernel_table()), nat.ptr(params), nat.ptr(inv_s),
                                    float(self.render_step_size), nat.ptr(cells),
                                    nat.ptr(jitter.contiguous()), m, og.resolution,
                                    float(ema_decay), float(occ_thre), nat.ptr(og.occs),
                                    nat.ptr(og.binary), nat.ptr(ws), ws.numel(), nat.stream()),
                  "am_occ_update")

    @torch.no_grad()
    def _occ_update_bg(self, cells, jitter, m, occ_thre, ema_decay, chunk=1 << 22):
        """occ_eval_fn_bg (models/neus.py:103-106: density * step) on the
        unbounded-sphere grid, with nerfacc 0.3.3's update rule: jittered
        cell points inside the contracted ball, mapped back to the world by
        `contract_inv_unbounded`; occs = max(occs * decay, occ) at the
        sampled cells; binary = occs > min(mean(occs), occ_thre).  A cell
        drawn more than once takes the max of its estimates (nerfacc's
        indexed assignment keeps whichever write lands last; the scatter-max
        is one of those outcomes, the same on every run, and the rule of the
        foreground grid's update)."""
        og = self.occupancy_grid_bg
        R = og.resolution
        dev = og.occs.device
        idx = torch.arange(og.num_cells, device=dev) if cells is None else cells.long()
        coords = torch.stack([idx // (R * R), (idx // R) % R, idx % R], -1).float()
        x = (coords + jitter) / R
        inside = (x - 0.5).norm(dim=1) < 0.5
        x, idx = x[inside], idx[inside]
        pts = contract_inv_unbounded(x, og.roi_aabb)
        occ = torch.empty(idx.shape[0], device=dev)
        for a in range(0, idx.shape[0], chunk):
            density, _ = self.geometry_bg(pts[a:a + chunk])
            occ[a:a + chunk] = density * self.render_step_size_bg
        cand = torch.maximum(og.occs[idx] * ema_decay, occ)
        og.occs.scatter_reduce_(0, idx, cand, reduce="amax", include_self=False)
        thre = torch.clamp(og.occs.mean(), max=occ_thre)
        og.binary.copy_((og.occs > thre).view(R, R, R).to(torch.uint8))

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
        aabb = self._aabb_host
        t_near, t_far = fn.ray_aabb(rays, aabb)
        jitter = torch.rand_like(t_near) if self.randomized else None
        grid = self.occupancy_grid.binary if self.config.get("grid_prune", False) else None
        roi = aabb
        return fn.ray_marching(rays, t_near, t_far, roi, grid, self.render_step_size, jitter,
                               self.max_samples_per_ray)

    def march_bg_candidates(self, rays):
        """The background marching of forward_bg_ up to the early-stop filter
        (models/neus.py:141-168): near = the far AABB hit (0.1 on a miss),
        far 1e3, stratified jitter of one step, cone-angle marching over the
        unbounded grid.  Reads the occupancy grid, not the weights."""
        rays = rays.float().contiguous()
        _, t_max = fn.ray_aabb(rays, self._aabb_host)
        near = torch.where(t_max > 1e9, torch.full_like(t_max, self.near_plane_bg), t_max)
        near = torch.clamp(near, min=0.0)
        far = torch.full_like(near, self.far_plane_bg)
        if self.randomized:
            near = near + torch.rand_like(near) * self.render_step_size_bg
        grid = self.occupancy_grid_bg.binary if self.config.get("grid_prune", False) else None
        return fn.ray_marching_bg_candidates(rays, near, far, self._aabb_host, grid,
                                             self.render_step_size_bg, self.cone_angle_bg,
                                             self.max_samples_per_ray_bg)

    def march_bg(self, rays, candidates=None):
        """Background sampling of forward_bg_ (models/neus.py:141-171): the
        candidates (march_bg_candidates, or the ones given), then the samples
        behind transmittance 1e-4 dropped (sigma_fn = the background
        density, models/neus.py:144-150)."""
        rays = rays.float().contiguous()
        if candidates is None:
            candidates = self.march_bg_candidates(rays)

        def sigma_fn(t0, t1, ridx):
            pts, _, _ = fn.sample_points(rays, ridx, t0, t1)
            density, _ = self.geometry_bg(pts)
            return density

        return fn.ray_marching_bg_filter(candidates, sigma_fn, 1e-4)

    def forward_bg_(self, rays, samples=None, candidates=None):
        """models/neus.py:141-203.  samples: optional (ray_indices int32,
        t_starts, t_ends, packed) replacing the sampling; candidates: the
        marching before the early-stop filter, made ahead
        (march_bg_candidates)."""
        rays = rays.float().contiguous()
        n_rays = rays.shape[0]
        if samples is None:
            with torch.no_grad():
                ray_indices, t_starts, t_ends, packed = self.march_bg(rays, candidates)
        else:
            ray_indices, t_starts, t_ends, packed = samples
        pts, mid, dists = fn.sample_points(rays, ray_indices, t_starts, t_ends)
        density, feature = self.geometry_bg(pts)
        t_dirs = rays[ray_indices.long(), 3:6]
        rgb = self.texture_bg(feature, t_dirs)
        weights, opacity, depth, comp_rgb = fn.composite_density(density, rgb, t_starts, t_ends,
                                                                 packed)
        opacity = opacity[:, None]
        depth = depth[:, None]
        bg = self.background_color
        if bg is None:
            bg = torch.ones(3, device=rays.device)
        comp_rgb = comp_rgb + bg * (1.0 - opacity)
        out = {
            "comp_rgb": comp_rgb,
            "opacity": opacity,
            "depth": depth,
            "rays_valid": opacity > 0,
            "num_samples": torch.full((1,), ray_indices.shape[0], dtype=torch.int32,
                                      device=rays.device),
        }
        if self.training:
            out.update({
                "weights": weights.view(-1),
                "points": mid.view(-1),
                "intervals": dists.view(-1),
                "ray_indices": ray_indices.long().view(-1),
            })
        return out

    # -- forward ------------------------------------------------------------
    def forward_(self, rays, samples=None, randn=None, samples_bg=None, bg_candidates=None):
        """models/neus.py:205-283.  samples / samples_bg: optional
        (ray_indices int32, t_starts, t_ends, packed) replacing the
        foreground / background sampling."""
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
            "num_samples": torch.full((1,), n, dtype=torch.int32, device=rays.device),
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
        if self.learned_background:
            out_bg = self.forward_bg_(rays, samples_bg, bg_candidates)
        else:
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

    def forward(self, rays, samples=None, randn=None, samples_bg=None, bg_candidates=None):
        """models/neus.py:285-293.  bg_candidates: the learned background's
        marching before its early-stop filter, made ahead for these rays
        (march_bg_candidates; a pipelined training loop)."""
        if self.training:
            out = self.forward_(rays, samples, randn, samples_bg, bg_candidates)
        else:
            out = chunk_batch(lambda r: self.forward_(r), self.config.ray_chunk, False, rays) \
                if samples is None else self.forward_(rays, samples, randn, samples_bg,
                                                      bg_candidates)
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
