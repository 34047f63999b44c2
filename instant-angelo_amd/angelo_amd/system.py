"""Training-step glue of NeuSSystem (systems/neus.py:24-206) without
PyTorch-Lightning: the loss terms of training_step (:120-206) with the
scheduled weights of BaseSystem.C (systems/base.py:28-45), the dynamic ray
count (:126-128), and a dynamic loss scaler for the reduced-precision form.
"""
import torch

from .utils import scheduled_value


def binary_cross_entropy(input, target):
    """systems/criterions.py:155-159 (stable BCE)."""
    return -(target * torch.log(input) + (1 - target) * torch.log(1 - input)).mean()


class NeuSTrainer:
    """Holds the per-step state of NeuSSystem that the hot path needs."""

    def __init__(self, model, system_config, model_config, has_mask=False):
        self.model = model
        # dataset.has_mask of the reference: the mask loss counts only when
        # the dataset has foreground masks (systems/neus.py:143-145); a batch
        # may say so itself ("has_mask", TrainRaySampler.sample does)
        self.has_mask = bool(has_mask)
        self.cfg = system_config
        self.mcfg = model_config
        self.train_num_rays = int(model_config.train_num_rays)
        self.train_num_samples = self.train_num_rays * (
            model_config.num_samples_per_ray + model_config.get("num_samples_per_ray_bg", 0))
        self.global_step = 0
        self.current_epoch = 0

    def C(self, value):
        return scheduled_value(value, self.global_step, self.current_epoch)

    def update_num_rays(self, out):
        """systems/neus.py:126-128 (one host read, as in the reference)."""
        if self.mcfg.get("dynamic_ray_sampling", False):
            n = int(out["num_samples_full"].sum().item())
            train_num_rays = int(self.train_num_rays * (self.train_num_samples / max(n, 1)))
            self.train_num_rays = min(int(self.train_num_rays * 0.9 + train_num_rays * 0.1),
                                      self.mcfg.max_train_num_rays)

    def _mean(self, t):
        """Loss reductions: fp32, or fp64 in the mixed-precision form (the
        compensated reductions of that form run in fp64 on the GPU)."""
        if getattr(self.model, "precision", "fp32") == "mixed":
            return t.double().mean().float()
        return t.mean()

    def _masked_mean(self, t, valid):
        """mean of t[valid] (the rows of the valid rays) without the host
        read a boolean index makes: sum(t * valid) / (count * columns); an
        empty selection gives NaN as the indexed mean does."""
        m = valid.to(t.dtype).unsqueeze(-1)
        if getattr(self.model, "precision", "fp32") == "mixed":
            return ((t * m).double().sum() / (m.double().sum() * t.shape[-1])).float()
        return (t * m).sum() / (m.sum() * t.shape[-1])

    def losses(self, out, batch):
        """systems/neus.py:130-194; returns (loss, dict of terms)."""
        L = self.cfg.loss
        terms = {}
        loss = 0.0
        valid = out["rays_valid_full"][..., 0]
        diff = out["comp_rgb_full"] - batch["rgb"]
        terms["rgb_mse"] = self._masked_mean(diff ** 2, valid)
        loss = loss + terms["rgb_mse"] * self.C(L.lambda_rgb_mse)
        if self.C(L.get("lambda_rgb_l1", 0.0)) > 0:
            terms["rgb_l1"] = self._masked_mean(diff.abs(), valid)
            loss = loss + terms["rgb_l1"] * self.C(L.lambda_rgb_l1)
        terms["eikonal"] = self._mean(
            (torch.linalg.norm(out["sdf_grad_samples"], ord=2, dim=-1) - 1.0) ** 2)
        loss = loss + terms["eikonal"] * self.C(L.lambda_eikonal)
        opacity = torch.clamp(out["opacity"].squeeze(-1), 1e-3, 1.0 - 1e-3)
        has_mask = bool(batch.get("has_mask", self.has_mask))
        if self.C(L.get("lambda_mask", 0.0)) > 0 and has_mask:
            terms["mask"] = binary_cross_entropy(opacity, batch["fg_mask"].float())
            loss = loss + terms["mask"] * self.C(L.lambda_mask)
        if self.C(L.get("lambda_opaque", 0.0)) > 0:
            terms["opaque"] = binary_cross_entropy(opacity, opacity)
            loss = loss + terms["opaque"] * self.C(L.lambda_opaque)
        if self.C(L.get("lambda_sparsity", 0.0)) > 0:
            terms["sparsity"] = torch.exp(-L.sparsity_scale * out["sdf_samples"].abs()).mean()
            loss = loss + terms["sparsity"] * self.C(L.lambda_sparsity)
        if self.C(L.get("lambda_curvature", 0.0)) > 0:
            terms["curvature"] = self._mean(out["sdf_laplace_samples"].abs())
            loss = loss + terms["curvature"] * self.C(L.lambda_curvature)
        distortion = ["lambda_distortion"]
        if self.mcfg.get("learned_background", False):
            distortion.append("lambda_distortion_bg")   # counted only then (:172)
        for k in distortion:
            if self.C(L.get(k, 0.0)) > 0:
                raise NotImplementedError(f"{k} > 0 is not part of this build")
        if self.C(L.get("lambda_sdf_l1", 0.0)) > 0:
            w_normal = self.C(L.get("lambda_normal", L.lambda_sdf_l1))
            sdf_l1, normal_cos = self.sparse_point_losses(batch, with_normal=w_normal > 0)
            terms["sdf_l1"] = sdf_l1
            loss = loss + sdf_l1 * self.C(L.lambda_sdf_l1)
            if normal_cos is not None:
                terms["normal_cos"] = normal_cos
                loss = loss + normal_cos * w_normal
        return loss, terms

    def sparse_point_losses(self, batch, with_normal=True):
        """systems/neus.py:180-194: the SDF at the dataset's sparse points
        (l1 towards 0, scaled by their confidence) and the cosine distance of
        its gradient to the points' normals (None when not asked for: the
        Neuralangelo configurations weight it 0, lambda_normal).  As there,
        the l1 term is the mean |sdf| (F.l1_loss) times the weights, averaged
        over dim 0."""
        pts, weights = batch.get("pts"), batch.get("pts_weights")
        normal = batch.get("pts_normal")
        if pts is None or pts.numel() == 0:
            raise ValueError("lambda_sdf_l1 > 0 needs the batch's sparse points ('pts')")
        if with_normal and (normal is None or normal.numel() == 0):
            raise ValueError("lambda_normal > 0 needs the sparse points' normals ('pts_normal')")
        sdf, sdf_grad = self.model.geometry(pts, with_grad=True, with_feature=False)
        loss_sdf = (sdf.abs().mean() * weights).mean(dim=0)
        if not with_normal:
            return loss_sdf, None
        n_gt = torch.nn.functional.normalize(normal, p=2, dim=-1)
        n_pred = torch.nn.functional.normalize(sdf_grad, p=2, dim=-1)
        loss_normal = (1.0 - torch.sum(n_pred * n_gt, dim=-1)).mean()
        return loss_sdf, loss_normal


class DynamicLossScaler:
    """Dynamic loss scaling (the GradScaler semantics behind precision: 16):
    scale the loss, skip the step and halve the scale on inf/nan gradients,
    double it after `growth_interval` clean steps."""

    def __init__(self, init_scale=2.0 ** 16, growth_factor=2.0, backoff_factor=0.5,
                 growth_interval=2000):
        self.scale = float(init_scale)
        self.growth_factor, self.backoff_factor = growth_factor, backoff_factor
        self.growth_interval = growth_interval
        self._good = 0

    def scale_loss(self, loss):
        return loss * self.scale

    def unscale_and_check(self, params):
        """Unscales grads in place; returns True if all are finite."""
        grads = [p.grad for p in params if p.grad is not None]
        if not grads:
            return True
        inv = 1.0 / self.scale
        for g in grads:
            g.mul_(inv)
        finite = torch.stack([torch.isfinite(g).all() for g in grads]).all()
        return bool(finite.item())

    def update(self, found_finite):
        if found_finite:
            self._good += 1
            if self._good >= self.growth_interval:
                self.scale *= self.growth_factor
                self._good = 0
        else:
            self.scale *= self.backoff_factor
            self._good = 0

    def state_dict(self):
        """What a checkpoint keeps (GradScaler's state_dict keys)."""
        return {"scale": self.scale, "growth_factor": self.growth_factor,
                "backoff_factor": self.backoff_factor, "growth_interval": self.growth_interval,
                "_growth_tracker": self._good}

    def load_state_dict(self, state):
        self.scale = float(state["scale"])
        self.growth_factor = float(state["growth_factor"])
        self.backoff_factor = float(state["backoff_factor"])
        self.growth_interval = int(state["growth_interval"])
        self._good = int(state["_growth_tracker"])

    def step(self, optimizer, params):
        ok = self.unscale_and_check(params)
        if ok:
            optimizer.step()
        self.update(ok)
        return ok
