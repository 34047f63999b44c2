"""Training-ray sampling on the device: systems/neus.py:35-118 (the train
stage of NeuSSystem.preprocess_data).

The view / pixel draws are torch.randint on the device; the gather of the
pixel's direction, camera and colour, get_rays (models/ray_utils.py:23-31),
the direction normalisation and the mask blend run in one HIP launch
(am_sample_rays).  Images, masks, cameras and directions stay resident, so a
step makes no host round trip for its batch.
"""
import torch

from . import _native as nat


class TrainRaySampler:
    """images [N,H,W,C] (C >= 3), c2w [N,3,4] (or [N,4,4]), directions
    [H,W,3] or [N,H,W,3], fg_masks [N,H,W] or None; all moved to `device`.

    batch_image_sampling: every ray draws its own view (else one view per
    batch, systems/neus.py:56-64).  sample_foreground_ratio < 1 with
    fg_indices / bg_indices ([M,3] rows of (view, y, x)) draws 80 % of the
    rays from the foreground pixels (:40-48).  background_color 'white' or
    'random' (:99-105); apply_mask blends the target with it (:108-109)."""

    def __init__(self, images, c2w, directions, fg_masks=None, device="cuda", apply_mask=False,
                 background_color="white", batch_image_sampling=True,
                 sample_foreground_ratio=1.0, fg_indices=None, bg_indices=None,
                 points=None, points_confidence=None, points_normal=None):
        if background_color not in ("white", "random"):
            raise NotImplementedError(f"background_color {background_color!r}")
        self.device = torch.device(device)
        self.images = images.to(self.device, torch.float32).contiguous()
        n, H, W, C = self.images.shape
        if C < 3:
            raise ValueError("images need at least 3 channels")
        self.c2w = c2w[:, :3, :4].to(self.device, torch.float32).contiguous()
        self.directions = directions.to(self.device, torch.float32).contiguous()
        if self.directions.shape[-3:-1] != (H, W) or self.directions.ndim not in (3, 4):
            raise ValueError("directions must be [H,W,3] or [N,H,W,3] for the images' H, W")
        self.fg_masks = None if fg_masks is None else \
            fg_masks.to(self.device, torch.float32).reshape(n, H, W).contiguous()
        self.n_views, self.H, self.W, self.C = n, H, W, C
        self.apply_mask = bool(apply_mask)
        self.background_color = background_color
        self.batch_image_sampling = batch_image_sampling
        self.ratio = float(sample_foreground_ratio)
        self.fg_indices = None if fg_indices is None else fg_indices.to(self.device).long()
        self.bg_indices = None if bg_indices is None else bg_indices.to(self.device).long()
        # sparse scene points of the dataset (all_points, all_points_confidence,
        # pts3d_normal), drawn one per ray for the sparse-point SDF loss
        self.points = None if points is None else points.to(self.device, torch.float32)
        self.points_confidence = None
        self.points_normal = None
        if self.points is not None:
            m = self.points.shape[0]
            self.points_confidence = (torch.ones(m, device=self.device) if points_confidence is None
                                      else points_confidence.to(self.device, torch.float32))
            if points_normal is not None:
                self.points_normal = points_normal.to(self.device, torch.float32)
        self.has_mask = fg_masks is not None
        self._bad = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.last_draw = None

    def draw(self, n, generator=None):
        """(view, y, x) int32 [n] each (systems/neus.py:36-64)."""
        dev, g = self.device, generator
        if self.batch_image_sampling and self.ratio < 1 and self.fg_indices is not None:
            n_fg = int(n * 0.8)
            fg = self.fg_indices[torch.randint(0, self.fg_indices.shape[0], (n_fg,), device=dev,
                                               generator=g)]
            bg = self.bg_indices[torch.randint(0, self.bg_indices.shape[0], (n - n_fg,), device=dev,
                                               generator=g)]
            idx = torch.cat([fg, bg], 0)
            index, y, x = idx[:, 0], idx[:, 1], idx[:, 2]
        elif self.batch_image_sampling:
            index = torch.randint(0, self.n_views, (n,), device=dev, generator=g)
            x = torch.randint(0, self.W, (n,), device=dev, generator=g)
            y = torch.randint(0, self.H, (n,), device=dev, generator=g)
        else:
            index = torch.randint(0, self.n_views, (1,), device=dev, generator=g).expand(n)
            x = torch.randint(0, self.W, (n,), device=dev, generator=g)
            y = torch.randint(0, self.H, (n,), device=dev, generator=g)
        i32 = lambda t: t.to(torch.int32).contiguous()
        return i32(index), i32(y), i32(x)

    def gather(self, index, y, x, bg_color):
        """The HIP launch: rays [n,6], rgb [n,3], fg_mask [n]."""
        n = index.shape[0]
        dev = self.device
        rays = torch.empty(n, 6, device=dev, dtype=torch.float32)
        rgb = torch.empty(n, 3, device=dev, dtype=torch.float32)
        fg = torch.empty(n, device=dev, dtype=torch.float32)
        bg = bg_color.to(dev, torch.float32).contiguous()
        nat.check(nat.load().am_sample_rays(
            nat.ptr(self.images), nat.ptr(self.fg_masks), nat.ptr(self.c2w),
            nat.ptr(self.directions), self.n_views, self.H, self.W, self.C,
            int(self.directions.ndim == 4), nat.ptr(bg), int(self.apply_mask), nat.ptr(index),
            nat.ptr(y), nat.ptr(x), n, nat.ptr(rays), nat.ptr(rgb), nat.ptr(fg),
            nat.ptr(self._bad), nat.stream()), "am_sample_rays")
        return rays, rgb, fg

    def sample(self, n, generator=None, model=None):
        """One training batch: dict(rays, rgb, fg_mask, has_mask, and with
        sparse points pts, pts_weights, pts_normal); the background colour of
        this step is set on `model` (systems/neus.py:99-107)."""
        if self.background_color == "white":
            bg = torch.ones(3, device=self.device)
        else:
            bg = torch.rand(3, device=self.device, generator=generator)
        index, y, x = self.draw(n, generator)
        self.last_draw = (index, y, x, bg)
        rays, rgb, fg = self.gather(index, y, x, bg)
        if model is not None:
            model.background_color = bg
        batch = dict(rays=rays, rgb=rgb, fg_mask=fg, has_mask=self.has_mask)
        if self.points is not None:
            # systems/neus.py:67-74: as many points as rays
            k = torch.randint(0, self.points.shape[0], (n,), device=self.device, generator=generator)
            batch["pts"] = self.points[k]
            batch["pts_weights"] = self.points_confidence[k]
            batch["pts_normal"] = (self.points_normal[k] if self.points_normal is not None
                                   else self.points.new_empty(0))
        return batch

    def out_of_range(self):
        """Draws clamped so far (0 unless an index list is bad).  Waits for
        the whole device, so draws made on another stream (a pipelined
        training loop) are counted too."""
        torch.cuda.synchronize(self.device)
        return int(self._bad.item())
