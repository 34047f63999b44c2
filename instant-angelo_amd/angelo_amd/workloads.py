"""Benchmark workloads: model / loss configurations with the reference's
config keys (configs/neuralangelo-colmap_sparse.yaml:16-134 layout) and the
synthetic data that feeds them.

  sphere-fp32   plain NeuS step: hash grid 8 levels x 2 features, 2^15 rows
                per level, base 16, scale 1.5; SDF MLP 1x64 -> 1+15; radiance
                2x64 sigmoid; 1024 rays x 64 uniform samples inside the unit
                sphere; Lambertian sphere of radius 0.5 as target.
  neuralangelo-fp32
                16 levels x 2 features, 2^21 rows, base 32, finest 2049;
                SDF MLP 1x64, radiance 2x64; 8192 rays per step marched
                through a 128^3 occupancy grid with step 2/1024 (scene
                diameter 2), at most 512 samples per ray; finite-difference
                normals, eikonal and curvature losses; sphere + torus scene
                (100 views at 512x512, rendered once; each step samples its
                rays on the device).
"""
import torch

from . import make, scenes
from .sampling import TrainRaySampler
from .system import DynamicLossScaler, NeuSTrainer
from .utils import to_config

NEURALANGELO_SCALE = 1.3195079107728942


def model_config(levels, log2T, base, scale, rays, samples_per_ray, step=None, cap=0,
                 grid_prune=False, geo_hidden=1, feature_dim=16, radius=1.0, start_level=None,
                 kahan=0, seed=0, precision="fp32", grad_type="finite_difference",
                 learned_background=False):
    cfg = to_config({
        "name": "neus", "radius": radius, "num_samples_per_ray": samples_per_ray,
        "train_num_rays": rays, "max_train_num_rays": rays, "dynamic_ray_sampling": False,
        "ray_chunk": 2048, "randomized": True, "grid_prune": grid_prune,
        "grid_prune_occ_thre": 0.001, "cos_anneal_end": 20000, "learned_background": False,
        "background_color": "white", "render_step_size": step, "max_samples_per_ray": cap,
        "kahan": kahan, "precision": precision,
        "variance": {"init_val": 0.3, "modulate": False},
        "geometry": {
            "name": "volume-sdf", "radius": radius, "feature_dim": feature_dim,
            "grad_type": grad_type, "finite_difference_eps": "progressive",
            "seed": seed, "isosurface": None,
            "xyz_encoding_config": {
                "otype": "ProgressiveBandHashGrid", "n_levels": levels, "n_features_per_level": 2,
                "log2_hashmap_size": log2T, "base_resolution": base, "per_level_scale": scale,
                "include_xyz": True, "start_level": start_level if start_level else levels,
                "start_step": 0, "update_steps": 1000},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": geo_hidden, "sphere_init": True,
                                   "sphere_init_radius": 0.5, "weight_norm": True}},
        "texture": {
            "name": "volume-radiance", "input_feature_dim": feature_dim + 6,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": 2},
            "color_activation": "sigmoid"},
    })
    if learned_background:
        # the background model of configs/neuralangelo-colmap_sparse.yaml:79-120
        cfg.learned_background = True
        cfg.background_color = "random"
        cfg.num_samples_per_ray_bg = 256
        cfg.grid_prune_occ_thre_bg = 0.01
        cfg.geometry_bg = to_config({
            "name": "volume-density", "radius": radius, "feature_dim": 8,
            "density_activation": "trunc_exp", "density_bias": -1, "isosurface": None,
            "seed": seed + 1,
            "xyz_encoding_config": {
                "otype": "ProgressiveBandHashGrid", "n_levels": 16, "n_features_per_level": 2,
                "log2_hashmap_size": 19, "base_resolution": 32,
                "per_level_scale": NEURALANGELO_SCALE, "include_xyz": True, "start_level": 4,
                "start_step": 5000, "update_steps": 1000},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": 1}})
        cfg.texture_bg = to_config({
            "name": "volume-dual-color", "input_feature_dim": 8,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": 2},
            "color_activation": "sigmoid"})
    return cfg


def neuralangelo_configs(radius=1.5, train_num_rays=256, log2_hashmap_size=19):
    """The model and system sections of configs/neuralangelo-colmap_sparse.yaml
    (:16-150) with its interpolations resolved: radius 1.5, 512 + 256
    samples per ray, dynamic ray sampling up to 8192 rays, analytic-gradient
    VolumeSDF with 65 outputs and two hidden layers on a 16-level 2^19-row
    progressive hash grid, VolumeDualColor texture (71 + 16 inputs), the
    learned background (VolumeDensity + VolumeRadiance), the sparse-point,
    colour, eikonal and curvature losses, AdamW.  Returns (model, system)."""
    band = {"otype": "ProgressiveBandHashGrid", "n_levels": 16, "n_features_per_level": 2,
            "log2_hashmap_size": log2_hashmap_size, "base_resolution": 32,
            "per_level_scale": NEURALANGELO_SCALE, "include_xyz": True, "start_level": 4,
            "start_step": 5000, "update_steps": 1000}
    mlp = lambda nh, **kw: dict({"otype": "VanillaMLP", "activation": "ReLU",
                                 "output_activation": "none", "n_neurons": 64,
                                 "n_hidden_layers": nh}, **kw)
    model = to_config({
        "name": "neus", "radius": radius, "num_samples_per_ray": 512,
        "train_num_rays": train_num_rays, "max_train_num_rays": 8192, "grid_prune": True,
        "grid_prune_occ_thre": 0.001, "dynamic_ray_sampling": True,
        "batch_image_sampling": True, "randomized": True, "ray_chunk": 2048,
        "cos_anneal_end": 20000, "learned_background": True, "background_color": "random",
        "variance": {"init_val": 0.3, "modulate": False},
        "geometry": {
            "name": "volume-sdf", "radius": radius, "feature_dim": 65, "grad_type": "analytic",
            "finite_difference_eps": "progressive",
            "isosurface": {"method": "mc", "resolution": 512, "chunk": 2097152,
                           "threshold": 0.001},
            "xyz_encoding_config": dict(band),
            "mlp_network_config": mlp(2, sphere_init=True, sphere_init_radius=0.5,
                                      weight_norm=True)},
        "texture": {
            "name": "volume-dual-color", "input_feature_dim": 65 + 6,
            "diffuse_warmup_steps": 5000,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": mlp(2), "color_activation": "sigmoid"},
        "num_samples_per_ray_bg": 256,
        "geometry_bg": {
            "name": "volume-density", "radius": radius, "feature_dim": 8,
            "density_activation": "trunc_exp", "density_bias": -1, "isosurface": None,
            "xyz_encoding_config": dict(band), "mlp_network_config": mlp(1)},
        "texture_bg": {
            "name": "volume-radiance", "input_feature_dim": 8,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": mlp(2), "color_activation": "sigmoid"},
    })
    system = to_config({
        "loss": {"lambda_sdf_l1": [0, 1, 0, 20000], "lambda_normal": 0.0,
                 "lambda_rgb_mse": 10.0, "lambda_rgb_l1": 0.0, "lambda_mask": 0.0,
                 "lambda_eikonal": 0.1, "lambda_curvature": [0, 0, 0.5, 5000],
                 "lambda_sparsity": 0.0, "lambda_distortion": 0.0, "lambda_distortion_bg": 0.0,
                 "lambda_opaque": 0.0, "sparsity_scale": 1.0},
        "optimizer": {"lr": 0.01, "betas": (0.9, 0.99), "eps": 1e-15, "variance_lr": 0.001},
    })
    return model, system


def system_config():
    """Loss weights of configs/neuralangelo-colmap_sparse.yaml:121-133 (the
    zero-weighted terms omitted) and the AdamW settings of :134-150."""
    return to_config({
        "loss": {"lambda_rgb_mse": 10.0, "lambda_eikonal": 0.1,
                 "lambda_curvature": [0, 0, 0.5, 5000]},
        "optimizer": {"lr": 0.01, "betas": (0.9, 0.99), "eps": 1e-15, "variance_lr": 0.001},
    })


def make_optimizer(model, scfg):
    o = scfg.optimizer
    groups = [{"params": list(model.geometry.parameters()), "lr": o.lr},
              {"params": list(model.texture.parameters()), "lr": o.lr},
              {"params": list(model.variance.parameters()), "lr": o.variance_lr}]
    if getattr(model, "learned_background", False):
        groups += [{"params": list(model.geometry_bg.parameters()), "lr": o.lr},
                   {"params": list(model.texture_bg.parameters()), "lr": o.lr}]
    try:
        return torch.optim.AdamW(groups, lr=o.lr, betas=tuple(o.betas), eps=o.eps, fused=True)
    except (RuntimeError, TypeError):
        return torch.optim.AdamW(groups, lr=o.lr, betas=tuple(o.betas), eps=o.eps, foreach=True)


WORKLOADS = {
    "neuralangelo-fp32": dict(levels=16, log2T=21, base=32, scale=NEURALANGELO_SCALE, rays=8192,
                              samples_per_ray=512, step=2.0 / 1024, cap=512, grid_prune=True),
    "sphere-fp32": dict(levels=8, log2T=15, base=16, scale=1.5, rays=1024, samples_per_ray=64,
                        step=None, cap=0, grid_prune=False),
}
# The mixed-precision form of the Neuralangelo step (same shapes): fp16
# table rows, bf16 MFMA matmuls with fp32 accumulation, fp16 table-gradient
# pairs with dynamic loss scaling, Kahan-compensated compositing, fp64 loss
# reductions.
WORKLOADS["neuralangelo-mixed"] = dict(WORKLOADS["neuralangelo-fp32"], precision="mixed")
# grad_type 'analytic' (the default of configs/neuralangelo-colmap_sparse.yaml:38):
# one base evaluation with d sdf / d x instead of six central differences,
# curvature still from six evaluations around the shifted point.
WORKLOADS["neuralangelo-analytic"] = dict(WORKLOADS["neuralangelo-fp32"], grad_type="analytic")
# With the learned background of the COLMAP configurations (learned_background
# true, models/neus.py:141-203): unbounded-sphere cone-angle marching with 256
# samples per ray, the early-stop filter, VolumeDensity + VolumeDualColor and
# a 256^3 background occupancy grid.
WORKLOADS["neuralangelo-bg"] = dict(WORKLOADS["neuralangelo-fp32"], learned_background=True)
# The start of training: the progressive band at its first 4 active levels
# (start_level 4; the timed steps stay below the first band update) and the
# finite-difference step of that band.
WORKLOADS["neuralangelo-fp32-l4"] = dict(WORKLOADS["neuralangelo-fp32"], start_level=4,
                                         global_step=0)
# The reference's own configuration (configs/neuralangelo-colmap_sparse.yaml,
# neuralangelo_configs): radius 1.5, render step 1.732 * 3 / 512 with no
# per-ray cap, analytic gradients, 65 SDF outputs and two hidden layers,
# 2^19 rows per level, VolumeDualColor, the learned background; 8192 rays a
# step (its max_train_num_rays, held fixed instead of the dynamic count).
WORKLOADS["neuralangelo-config"] = dict(levels=16, log2T=19, base=32, scale=NEURALANGELO_SCALE,
                                        rays=8192, samples_per_ray=512, step=None, cap=0,
                                        grid_prune=True, grad_type="analytic",
                                        learned_background=True, feature_dim=65,
                                        geo_hidden=2, radius=1.5, reference_config=True)


class Workload:
    """Model, trainer, optimizer and the step's data: a pool of
    device-resident batches with uniform samples (sphere workload), or rays
    drawn on the device each step from a resident image set (Neuralangelo
    workloads).  pipeline=True draws and marches the next step's rays on a
    second stream while the current step's backward runs (marched workloads
    only; bit-identical to the sequential loop)."""

    def __init__(self, name, device, pool=8, seed=0, rays=None, global_step=20000, band_cells=1.5,
                 pipeline=False):
        spec = dict(WORKLOADS[name])
        if rays is not None:
            spec["rays"] = rays
        self.name, self.spec, self.device = name, spec, device
        self.uniform = not spec["grid_prune"]
        if spec.get("reference_config"):
            mcfg, scfg = neuralangelo_configs(train_num_rays=spec["rays"],
                                              log2_hashmap_size=spec["log2T"])
            mcfg.dynamic_ray_sampling = False
            mcfg.max_train_num_rays = spec["rays"]
            mcfg.precision = spec.get("precision", "fp32")
        else:
            mcfg = model_config(spec["levels"], spec["log2T"], spec["base"], spec["scale"],
                                spec["rays"], spec["samples_per_ray"], spec["step"], spec["cap"],
                                spec["grid_prune"], seed=seed,
                                precision=spec.get("precision", "fp32"),
                                grad_type=spec.get("grad_type", "finite_difference"),
                                learned_background=spec.get("learned_background", False),
                                start_level=spec.get("start_level"))
            scfg = system_config()
        torch.manual_seed(seed)
        self.mcfg, self.scfg = mcfg, scfg
        self.model = make("neus", mcfg).to(device)
        self.model.train()
        self.trainer = NeuSTrainer(self.model, self.scfg, mcfg)
        self.optimizer = make_optimizer(self.model, self.scfg)
        self.mixed = spec.get("precision", "fp32") == "mixed"
        self.scaler = DynamicLossScaler() if self.mixed else None
        self.skipped_steps = 0
        self.bg_samples = 0   # background samples of the steps so far (in step()'s totals)
        global_step = spec.get("global_step", global_step)
        self.global_step = global_step
        self.batches = []
        gen = torch.Generator().manual_seed(seed + 17)
        if self.uniform:
            for b in range(pool):
                rays, rgb = scenes.config0_batch(spec["rays"], n_views=16, seed=seed + b)
                rays, rgb = rays.to(device), rgb.to(device)
                samples = scenes.uniform_samples(rays, spec["samples_per_ray"])
                self.batches.append(dict(rays=rays, rgb=rgb, samples=samples))
        else:
            # 100 seeded 512x512 views rendered once and kept on the device;
            # each step draws its rays there (systems/neus.py:35-118)
            cams, whf = scenes.config1_pool(100, 512, 512, 40.0, seed)
            images, dirs, c2w = scenes.config1_images(cams, whf, device)
            self.sampler = TrainRaySampler(images, c2w, dirs, None, device=device,
                                           background_color=mcfg.background_color)
            self.sample_gen = torch.Generator(device=device).manual_seed(seed + 17)
            og = self.model.occupancy_grid
            binary = scenes.prewarm_occupancy(scenes.sdf_union, og.resolution,
                                              float(mcfg.radius), band_cells, device)
            og.binary.copy_(binary)
            og.occs.copy_(binary.flatten().float())
        self.model.update_step(0, global_step - 1)
        # pipeline: the next step's rays are drawn and marched on a side
        # stream while this step's backward runs (marching reads the
        # occupancy grid only), so its sample-count read does not drain the
        # main stream.  Not before a step that updates the grid.  The global
        # generator's draws (march jitter, grid updates) keep their order, so
        # the step sequence is the sequential loop's.
        self.pipeline = bool(pipeline) and not self.uniform
        self._side = torch.cuda.Stream(device) if self.pipeline else None
        self._ahead = None      # (step, batch, event) marched on the side stream
        self._grid_ev = None    # the main stream after this step's grid update

    def _draw(self, set_model=True):
        b = dict(self.sampler.sample(self.spec["rays"], self.sample_gen,
                                     self.model if set_model else None), samples=None)
        b["bg"] = self.sampler.last_draw[3]
        return b

    def _prefetch(self, i):
        main = torch.cuda.current_stream(self.device)
        side = self._side
        with torch.cuda.stream(side):
            if self._grid_ev is not None:
                side.wait_event(self._grid_ev)
            b = self._draw(set_model=False)
            b["samples"] = self.model.march(b["rays"])
            # the step's other global-generator draws, in the forward's
            # order: the curvature directions, then the background jitter
            # (with the background's marching up to its early-stop filter)
            b["randn"] = torch.randn(b["samples"][0].shape[0], 3, device=self.device)
            if self.model.learned_background:
                b["bg_candidates"] = self.model.march_bg_candidates(b["rays"])
            ev = torch.cuda.Event()
            ev.record(side)
        for v in b.values():
            vals = v.values() if isinstance(v, dict) else (v if isinstance(v, (tuple, list)) else (v,))
            for t in vals:
                if isinstance(t, torch.Tensor) and t.is_cuda:
                    t.record_stream(main)   # freed memory waits for the main stream's use
        self._ahead = (i, b, ev)

    def drain(self):
        """Make the current stream wait for a march already started for the
        next step (so a timed region that ends here includes it)."""
        if self._ahead is not None:
            torch.cuda.current_stream(self.device).wait_event(self._ahead[2])

    def step(self, i):
        """One training step; returns the number of samples processed."""
        if self.uniform:
            b = self.batches[i % len(self.batches)]
        elif self._ahead is not None and self._ahead[0] == i:
            _, b, ev = self._ahead
            self._ahead = None
            torch.cuda.current_stream(self.device).wait_event(ev)
            self.model.background_color = b["bg"]
        else:
            self._ahead = None
            b = self._draw()
        gs = self.global_step + i
        self.trainer.global_step = gs
        self.model.update_step(0, gs)
        if self.pipeline:
            self._grid_ev = torch.cuda.Event()
            self._grid_ev.record()
        out = self.model(b["rays"], samples=b["samples"], randn=b.get("randn"),
                         bg_candidates=b.get("bg_candidates"))
        loss, _ = self.trainer.losses(out, b)
        self.optimizer.zero_grad(set_to_none=True)
        if self.scaler is not None:
            self.scaler.scale_loss(loss).backward()
        else:
            loss.backward()
        if self.pipeline and not self.model.grid_update_due(gs + 1):
            self._prefetch(i + 1)
        if self.scaler is not None:
            if not self.scaler.step(self.optimizer, list(self.model.parameters())):
                self.skipped_steps += 1
        else:
            self.optimizer.step()
        n = out["ray_indices"].shape[0]
        if "ray_indices_bg" in out:
            nb = out["ray_indices_bg"].shape[0]
            self.bg_samples += nb
            n += nb
        return n
