"""VolumeSDF oracle -- test infrastructure only.

Restates models/geometry.py:195-287 (VolumeSDF.forward, grad_type
'finite_difference') line by line with torch CPU ops:

  contract       scale_anything(x, (-r, r), (0, 1)) (models/utils.py:109-114)
  base           out = MLP(cat(x01*2-1, enc(x01))); sdf = out[...,0];
                 feature = cat(out, x01*2-1)                         (:204-208)
  FD normal      clamp(P + offsets, -r, r) -> contract -> sdf;
                 grad = 0.5*(f+ - f-)/eps                            (:219-234)
  analytic grad  grad_type 'analytic' (:214-218): grad = autograd.grad(sdf,
                 points, create_graph=True), so losses on it differentiate to
                 second order
  curvature      n = normalize(grad); t = cross(n, normalize(randn));
                 ps = x01 + t*eps (the contracted point, :246); the six
                 offsets are added to ps, clamped and contracted again
                 (:264-265); laplace = acos(clamp(n.ns, -1+1e-6, 1-1e-6))/pi
                                                                     (:236-275)

Two arithmetic modes.  'f32' reproduces what the reference's PyTorch runs on
the GPU: a Python-scalar divisor is applied as a multiplication by its fp32
reciprocal (torch's CUDA div-by-scalar), so contracted positions -- and the
hash indices computed from them -- are bit-identical with the HIP kernel.
'f64' is exact real arithmetic in float64, used for error budgets.

Pins (tests/test_oracle_field.py): a field that is linear in the world
position gives the exact gradient by central differences; the curvature of a
sphere SDF is 0 with recontract off and matches a hand-computed value for the
re-contraction quirk; `torch.autograd.gradcheck` in float64 on tiny inputs.
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

OFFSETS = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]


def _scalar_div(x, s, mode):
    if mode == "f32":
        return x * torch.tensor(np.float32(1.0) / np.float32(s), dtype=x.dtype, device=x.device)
    return x / s


def contract(x, r, mode="f32"):
    return _scalar_div(x - (-r), r - (-r), mode) * (1 - 0) + 0


def volume_sdf(points, field, r, eps, with_grad=True, with_laplace=False, randn=None,
               recontract=True, mode="f32", grad_type="finite_difference"):
    """field(x01 [M,3]) -> out [M, n_out].  Returns dict of sdf [N], grad [N,3],
    normal [N,3], feature [N, n_out+3], laplace [N,1]."""
    dt = points.dtype
    analytic = with_grad and grad_type == "analytic"
    if analytic and not points.requires_grad:
        points = points.detach().requires_grad_(True)
    x = contract(points, r, mode)
    out = field(x.view(-1, 3))
    sdf = out[..., 0]
    feature = torch.cat([out, x * 2 - 1], dim=-1)
    res = {"sdf": sdf, "feature": feature}
    if not with_grad:
        return res
    e = np.float32(eps) if mode == "f32" else float(eps)
    offsets = (torch.tensor(OFFSETS, dtype=dt, device=points.device)
               * torch.tensor(e, dtype=dt, device=points.device))
    if analytic:
        grad = torch.autograd.grad(sdf, points, grad_outputs=torch.ones_like(sdf),
                                   create_graph=True, retain_graph=True, only_inputs=True)[0]
    else:
        pd = (points[..., None, :] + offsets).clamp(-r, r)
        xd = contract(pd, r, mode)
        fd = field(xd.view(-1, 3))[..., 0].view(-1, 6)
        grad = _scalar_div(0.5 * (fd[..., 0::2] - fd[..., 1::2]), float(e), mode)
    normal = F.normalize(grad, p=2, dim=-1)
    res.update(grad=grad, normal=normal)
    if not with_laplace:
        return res
    rd = F.normalize(randn, dim=-1)
    t = torch.linalg.cross(normal, rd, dim=-1)
    base = x if recontract else points
    ps = base.clone() + t * torch.tensor(e, dtype=dt, device=points.device)
    psd = (ps[..., None, :] + offsets).clamp(-r, r)
    xsd = contract(psd, r, mode)
    fsd = field(xsd.view(-1, 3))[..., 0].view(-1, 6)
    gs = _scalar_div(0.5 * (fsd[..., 0::2] - fsd[..., 1::2]), float(e), mode)
    ns = F.normalize(gs, dim=-1)
    dot = (normal * ns).sum(dim=-1, keepdim=True)
    angle = torch.acos(torch.clamp(dot, -1.0 + 1e-6, 1.0 - 1e-6))
    res["laplace"] = _scalar_div(angle, math.pi, mode)
    return res


def field_from(encode, mlp):
    """Composes encode(x01) -> [M, L*F] and mlp(cat) as VolumeSDF does
    (CompositeEncoding with include_xyz, xyz_scale 2, offset -1)."""
    def f(x):
        return mlp(torch.cat([x * 2.0 - 1.0, encode(x)], dim=-1))
    return f


def contract_unbounded(x, r, mode="f32"):
    """models/geometry.py:22-28 (ContractionType.UN_BOUNDED_SPHERE):
    [-r, r]^3 -> [0,1]^3 -> [-1,1]^3; norm m > 1 -> (2 - 1/m) x/m; x/4 + 1/2."""
    x = contract(x, r, mode)
    x = x * 2 - 1
    mag = x.norm(dim=-1, keepdim=True)
    mask = mag.squeeze(-1) > 1
    x = x.clone()
    x[mask] = (2 - 1 / mag[mask]) * (x[mask] / mag[mask])
    return x / 4 + 0.5


def volume_density(points, field, r, density_bias=-1.0, mode="f32"):
    """VolumeDensity.forward (models/geometry.py:159-167) with
    density_activation trunc_exp (exp in the forward, models/utils.py:54-69):
    returns density [N], feature [N, n_out]."""
    x = contract_unbounded(points, r, mode)
    out = field(x.view(-1, 3))
    return torch.exp(out[..., 0] + float(density_bias)), out
