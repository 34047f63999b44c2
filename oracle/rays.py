"""Training-ray sampling oracle -- test infrastructure only.

Restates the train stage of systems/neus.py:66-109: the gather of the drawn
pixel's direction (models/ray_utils.py:9-20 layout), get_rays
(models/ray_utils.py:23-31: rays_d_i = sum_j d_j c2w[i][j], origin =
c2w[:, 3]), F.normalize(rays_d) (x / max(|x|, 1e-12)), the pixel's colour and
foreground mask, and the mask blend with the background colour -- in
float32 numpy, one rounded operation at a time, with the sums taken left to
right.

Pins (tests/test_oracle_rays.py): the reference's torch expressions
((d[:, None, :] * c2w[:, :3, :3]).sum(-1), F.normalize) evaluated in
float64 agree to 2 ulp; an identity rotation returns the translation and
d/|d| exactly; a quarter turn about z permutes the axes with the sign; mask
0 and 1 give the background and the pixel colour.
"""
import numpy as np


def sample_rays(images, c2w, directions, index, ys, xs, fg_masks=None, bg_color=None,
                apply_mask=False):
    f32 = np.float32
    images = np.asarray(images, f32)
    c2w = np.asarray(c2w, f32)[:, :3, :4]
    directions = np.asarray(directions, f32)
    v, y, x = (np.asarray(t, np.int64) for t in (index, ys, xs))
    d = directions[y, x] if directions.ndim == 3 else directions[v, y, x]
    m = c2w[v]
    r = np.empty((v.shape[0], 3), f32)
    for i in range(3):
        r[:, i] = (d[:, 0] * m[:, i, 0] + d[:, 1] * m[:, i, 1]) + d[:, 2] * m[:, i, 2]
    nrm = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
    den = np.maximum(nrm, f32(1e-12))
    rays = np.concatenate([m[:, :, 3], r / den[:, None]], -1).astype(f32)
    rgb = images[v, y, x, :3].copy()
    fg = np.ones(v.shape[0], f32) if fg_masks is None else np.asarray(fg_masks, f32)[v, y, x]
    if apply_mask:
        bg = np.asarray(bg_color, f32)
        rgb = (rgb * fg[:, None] + bg[None, :] * (f32(1.0) - fg[:, None])).astype(f32)
    return rays, rgb, fg
