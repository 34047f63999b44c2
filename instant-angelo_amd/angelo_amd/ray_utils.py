"""Pinhole rays (models/ray_utils.py:9-43): OpenGL convention, -z forward."""
import numpy as np
import torch


def get_ray_directions(W, H, fx, fy, cx, cy, use_pixel_centers=True):
    pixel_center = 0.5 if use_pixel_centers else 0
    i, j = np.meshgrid(np.arange(W, dtype=np.float32) + pixel_center,
                       np.arange(H, dtype=np.float32) + pixel_center, indexing="xy")
    i, j = torch.from_numpy(i), torch.from_numpy(j)
    return torch.stack([(i - cx) / fx, -(j - cy) / fy, -torch.ones_like(i)], -1)


def get_rays(directions, c2w, keepdim=False):
    assert directions.shape[-1] == 3
    if directions.ndim == 2:
        assert c2w.ndim == 3
        rays_d = (directions[:, None, :] * c2w[:, :3, :3]).sum(-1)
        rays_o = c2w[:, :, 3].expand(rays_d.shape)
    elif directions.ndim == 3:
        if c2w.ndim == 2:
            rays_d = (directions[:, :, None, :] * c2w[None, None, :3, :3]).sum(-1)
            rays_o = c2w[None, None, :, 3].expand(rays_d.shape)
        else:
            rays_d = (directions[None, :, :, None, :] * c2w[:, None, None, :3, :3]).sum(-1)
            rays_o = c2w[:, None, None, :, 3].expand(rays_d.shape)
    else:
        raise ValueError("directions must be [N,3] or [H,W,3]")
    if not keepdim:
        rays_o, rays_d = rays_o.reshape(-1, 3), rays_d.reshape(-1, 3)
    return rays_o, rays_d
