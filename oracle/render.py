"""NeuS rendering oracle -- test infrastructure only.

  alpha          NeuSModel.get_alpha (models/neus.py:117-139), torch ops in
                 the reference's order.
  weights        nerfacc 0.3.3 render_weight_from_alpha: w_i = T_i*a_i,
                 T_i = prod_{j<i} (1 - a_j) per ray (sequential product).
  accumulate     accumulate_along_rays: per-ray sums of w, w*mid, w*rgb,
                 w*normal (index_add; the reference scatter_add_s),
                 comp_normal = normalize, zeroed where opacity <= 0.1
                 (models/neus.py:235-241).
  neus_forward   the whole NeuSModel.forward_ (:205-283) from given samples,
                 with the foreground-only output dict (:242-283).
  losses         systems/neus.py:130-159: 10*MSE(rgb_full[valid]),
                 lambda_eik*mean((|grad|-1)^2), lambda_curv*mean(|laplace|).

  weights_from_density / forward_bg   nerfacc 0.3.3 render_weight_from_density
                 (T_i = exp(-sum_{j<i} sigma_j dt_j), w_i = T_i (1 - exp(-s_i)))
                 and the background pass NeuSModel.forward_bg_ (:172-203).

Pins (tests/test_oracle_render.py): closed forms of the transmittance (one
sample per ray, alpha = 1 stops the ray, alpha = 0 contributes nothing, the
sum of weights equals 1 - prod(1 - a)); scalar hand cases of get_alpha; the
per-ray sums against explicit Python loops; gradcheck in float64.  For the
density weights: constant density gives T_i = exp(-sigma i dt) in closed
form, sum w = 1 - exp(-sum s), a per-ray Python loop of the sequential
product form, and gradcheck of forward_bg in float64.
"""
import numpy as np
import torch
import torch.nn.functional as F

from . import field as ofield


def get_alpha(sdf, normal, dirs, dists, inv_s, cos_anneal_ratio):
    inv_s = inv_s.reshape(1, 1).expand(sdf.shape[0], 1)
    true_cos = (dirs * normal).sum(-1, keepdim=True)
    r = cos_anneal_ratio
    iter_cos = -(F.relu(-true_cos * 0.5 + 0.5) * (1.0 - r) + F.relu(-true_cos) * r)
    estimated_next_sdf = sdf[..., None] + iter_cos * dists.reshape(-1, 1) * 0.5
    estimated_prev_sdf = sdf[..., None] - iter_cos * dists.reshape(-1, 1) * 0.5
    prev_cdf = torch.sigmoid(estimated_prev_sdf * inv_s)
    next_cdf = torch.sigmoid(estimated_next_sdf * inv_s)
    p = prev_cdf - next_cdf
    c = prev_cdf
    return ((p + 1e-5) / (c + 1e-5)).view(-1).clip(0.0, 1.0)


def packed_from_indices(ray_indices, n_rays):
    counts = torch.bincount(ray_indices.long(), minlength=n_rays)
    starts = torch.cumsum(counts, 0) - counts
    return torch.stack([starts, counts], -1).to(torch.int32)


def transmittance(alpha, ray_indices, n_rays):
    """Exclusive per-ray product of (1 - alpha), sequential within the ray."""
    packed = packed_from_indices(ray_indices, n_rays)
    counts = packed[:, 1].long()
    L = int(counts.max()) if counts.numel() and int(counts.max()) > 0 else 0
    if L == 0:
        return torch.ones_like(alpha)
    starts = packed[:, 0].long()
    pos = torch.arange(alpha.shape[0], device=alpha.device) - starts[ray_indices.long()]
    om = torch.ones(n_rays, L + 1, dtype=alpha.dtype, device=alpha.device)
    om = om.index_put((ray_indices.long(), pos + 1), 1.0 - alpha)
    T = torch.cumprod(om, dim=1)
    return T[ray_indices.long(), pos]


def accumulate(weights, ray_indices, n_rays, values=None):
    src = weights[:, None] if values is None else weights[:, None] * values
    out = torch.zeros(n_rays, src.shape[-1], dtype=src.dtype, device=src.device)
    return out.index_add(0, ray_indices.long(), src)


def composite(sdf, normal, rgb, dirs, t_starts, t_ends, ray_indices, n_rays, inv_s, ratio):
    mid = (t_starts + t_ends) / 2.0
    dists = t_ends - t_starts
    alpha = get_alpha(sdf, normal, dirs, dists, inv_s, ratio)
    T = transmittance(alpha, ray_indices, n_rays)
    w = T * alpha
    opacity = accumulate(w, ray_indices, n_rays)
    depth = accumulate(w, ray_indices, n_rays, mid[:, None])
    comp_rgb = accumulate(w, ray_indices, n_rays, rgb)
    rays_fg = opacity > 0.1
    cn = accumulate(w, ray_indices, n_rays, normal)
    cn = F.normalize(cn, p=2, dim=-1)
    cn = torch.where(rays_fg, cn, torch.zeros_like(cn))
    return dict(alpha=alpha, weights=w, trans=T, opacity=opacity, depth=depth, comp_rgb=comp_rgb,
                comp_normal=cn)


def neus_forward(rays, ray_indices, t_starts, t_ends, n_rays, field, radiance, inv_s, radius, eps,
                 ratio, randn, background=None, recontract=True, mode="f32",
                 grad_type="finite_difference"):
    """NeuSModel.forward_ from given samples.  field(x01) -> [M, n_out];
    radiance(feature, dirs, normal) -> rgb."""
    ridx = ray_indices.long()
    rays_o, rays_d = rays[:, 0:3], rays[:, 3:6]
    t_origins = rays_o[ridx]
    t_dirs = rays_d[ridx]
    midpoints = (t_starts + t_ends) / 2.0
    positions = t_origins + t_dirs * midpoints[:, None]
    geo = ofield.volume_sdf(positions, field, radius, eps, True, True, randn, recontract, mode,
                            grad_type=grad_type)
    normal = geo["normal"]
    rgb = radiance(geo["feature"], t_dirs, normal)
    c = composite(geo["sdf"], normal, rgb, t_dirs, t_starts, t_ends, ridx, n_rays, inv_s, ratio)
    bg = background if background is not None else torch.ones(3, dtype=rays.dtype,
                                                                 device=rays.device)
    full = c["comp_rgb"] + bg[None, :] * (1.0 - c["opacity"])
    return dict(c, sdf=geo["sdf"], grad=geo["grad"], normal=normal, feature=geo["feature"],
                laplace=geo["laplace"], rgb=rgb, positions=positions, comp_rgb_full=full,
                rays_valid=c["opacity"] > 0)


def weights_from_density(sigma, t_starts, t_ends, ray_indices, n_rays):
    """nerfacc 0.3.3 render_weight_from_density (models/neus.py:174):
    s = sigma dt, T_i = exp(-sum_{j<i} s_j) per ray, w_i = T_i (1 - exp(-s_i))."""
    s = sigma * (t_ends - t_starts)
    packed = packed_from_indices(ray_indices, n_rays)
    starts = packed[:, 0].long()[ray_indices.long()]
    csum = torch.cumsum(s, 0)
    # exclusive per-ray sum: global inclusive cumsum minus the ray's offset
    offset = torch.where(starts > 0, csum[(starts - 1).clamp(min=0)], torch.zeros_like(csum))
    excl = csum - s - offset
    T = torch.exp(-excl)
    return T * (1.0 - torch.exp(-s)), T


def forward_bg(rays, ray_indices, t_starts, t_ends, n_rays, field, radiance, radius,
               density_bias=-1.0, background=None, mode="f32"):
    """NeuSModel.forward_bg_ (models/neus.py:172-203) from given samples.
    field(x01) -> [M, n_out] (the background encoding + MLP);
    radiance(feature, dirs) -> rgb."""
    ridx = ray_indices.long()
    rays_o, rays_d = rays[:, 0:3], rays[:, 3:6]
    t_dirs = rays_d[ridx]
    midpoints = (t_starts + t_ends) / 2.0
    positions = rays_o[ridx] + t_dirs * midpoints[:, None]
    density, feature = ofield.volume_density(positions, field, radius, density_bias, mode)
    rgb = radiance(feature, t_dirs)
    w, T = weights_from_density(density, t_starts, t_ends, ridx, n_rays)
    opacity = accumulate(w, ridx, n_rays)
    depth = accumulate(w, ridx, n_rays, midpoints[:, None])
    comp_rgb = accumulate(w, ridx, n_rays, rgb)
    bg = background if background is not None else torch.ones(3, dtype=rays.dtype,
                                                                 device=rays.device)
    comp_rgb = comp_rgb + bg[None, :] * (1.0 - opacity)
    return dict(density=density, feature=feature, rgb=rgb, weights=w, trans=T, opacity=opacity,
                depth=depth, comp_rgb=comp_rgb, positions=positions)


def losses(out, gt_rgb, lambda_rgb=10.0, lambda_eik=0.1, lambda_curv=0.5):
    valid = out["rays_valid"][..., 0]
    l_rgb = F.mse_loss(out["comp_rgb_full"][valid], gt_rgb[valid])
    l_eik = ((torch.linalg.norm(out["grad"], ord=2, dim=-1) - 1.0) ** 2).mean()
    l_curv = out["laplace"].abs().mean()
    total = l_rgb * lambda_rgb + l_eik * lambda_eik + l_curv * lambda_curv
    return total, dict(rgb_mse=l_rgb, eikonal=l_eik, curvature=l_curv)
