"""Pins of the training-ray sampling oracle (systems/neus.py:66-109)."""
import math

import numpy as np
import torch
import torch.nn.functional as F

from oracle import rays as orays


def _scene(seed=0, n=4, H=5, W=7, C=3):
    rng = np.random.default_rng(seed)
    images = rng.uniform(size=(n, H, W, C)).astype(np.float32)
    masks = (rng.uniform(size=(n, H, W)) < 0.5).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, a, b, c = q.T
    R = np.stack([np.stack([1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)], -1),
                  np.stack([2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)], -1),
                  np.stack([2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)], -1)],
                 1)
    t = rng.normal(size=(n, 3, 1)) * 2
    c2w = np.concatenate([R, t], -1).astype(np.float32)
    f = 0.5 * W / math.tan(math.radians(20))
    i, j = np.meshgrid(np.arange(W, dtype=np.float32) + 0.5, np.arange(H, dtype=np.float32) + 0.5,
                       indexing="xy")
    dirs = np.stack([(i - W / 2) / f, -(j - H / 2) / f, -np.ones_like(i)], -1).astype(np.float32)
    m = 64
    index = rng.integers(0, n, m)
    ys = rng.integers(0, H, m)
    xs = rng.integers(0, W, m)
    return images, masks, c2w, dirs, index, ys, xs


def test_matches_reference_torch_expressions():
    images, masks, c2w, dirs, index, ys, xs = _scene()
    rays, rgb, fg = orays.sample_rays(images, c2w, dirs, index, ys, xs, masks)
    d = torch.from_numpy(dirs).double()[ys, xs]
    C = torch.from_numpy(c2w).double()[index]
    rays_d = (d[:, None, :] * C[:, :3, :3]).sum(-1)
    rays_o = C[:, :, 3].expand(rays_d.shape)
    ref = torch.cat([rays_o, F.normalize(rays_d, p=2, dim=-1)], -1).numpy()
    assert np.allclose(rays, ref, rtol=0, atol=2.5e-7)
    assert np.array_equal(rgb, images[index, ys, xs])
    assert np.array_equal(fg, masks[index, ys, xs])


def test_identity_and_quarter_turn():
    images, masks, c2w, dirs, index, ys, xs = _scene(1)
    c2w[:, :3, :3] = np.eye(3, dtype=np.float32)
    rays, _, _ = orays.sample_rays(images, c2w, dirs, index, ys, xs)
    assert np.array_equal(rays[:, :3], c2w[index, :, 3])
    d = dirs[ys, xs]
    assert np.allclose(rays[:, 3:], d / np.linalg.norm(d.astype(np.float64), axis=1)[:, None],
                       rtol=0, atol=1.5e-7)
    c2w[:, :3, :3] = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)   # +90 deg about z
    rays2, _, _ = orays.sample_rays(images, c2w, dirs, index, ys, xs)
    assert np.array_equal(rays2[:, 3], -rays[:, 4])
    assert np.array_equal(rays2[:, 4], rays[:, 3])
    assert np.array_equal(rays2[:, 5], rays[:, 5])


def test_mask_blend_hand_cases():
    images, masks, c2w, dirs, index, ys, xs = _scene(2)
    bg = np.array([0.25, 0.5, 1.0], np.float32)
    _, rgb, fg = orays.sample_rays(images, c2w, dirs, index, ys, xs, masks, bg, apply_mask=True)
    px = images[index, ys, xs]
    assert np.array_equal(rgb[fg == 1], px[fg == 1])
    assert np.array_equal(rgb[fg == 0], np.broadcast_to(bg, px[fg == 0].shape))
    assert (fg == 0).any() and (fg == 1).any()
