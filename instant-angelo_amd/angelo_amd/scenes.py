"""Seeded synthetic scenes, cameras and ray batches for the benchmark
workloads (the reference trains on COLMAP captures; none ship with it and
there is no network, so the inputs are analytic scenes of the same shape).

  sphere  Lambertian sphere of radius 0.5 under a fixed light (64x64 views,
          40 deg fov, cameras on a radius-2.5 sphere looking at the origin).
  union   sphere + torus with a procedural albedo (512x512 views).

Rays follow models/ray_utils.py (OpenGL camera, -z forward) with unit
directions (systems/neus.py:95).
"""
import math

import torch

from .ray_utils import get_ray_directions, get_rays

LIGHT_DIR = (0.4, 0.8, 0.45)


def look_at_c2w(eye):
    eye = torch.as_tensor(eye, dtype=torch.float32)
    fwd = -eye / eye.norm()            # camera looks along -z: z axis = -fwd
    up = torch.tensor([0.0, 0.0, 1.0])
    if abs(float(fwd @ up)) > 0.99:
        up = torch.tensor([0.0, 1.0, 0.0])
    right = torch.linalg.cross(fwd, up)
    right = right / right.norm()
    true_up = torch.linalg.cross(right, fwd)
    c2w = torch.zeros(3, 4)
    c2w[:, 0] = right
    c2w[:, 1] = true_up
    c2w[:, 2] = -fwd
    c2w[:, 3] = eye
    return c2w


def sphere_cameras(n_views, radius=2.5, seed=0):
    g = torch.Generator().manual_seed(seed)
    v = torch.randn(n_views, 3, generator=g)
    v = v / v.norm(dim=-1, keepdim=True) * radius
    return torch.stack([look_at_c2w(e) for e in v])


def camera_rays(c2w, W, H, fov_deg):
    f = 0.5 * W / math.tan(0.5 * math.radians(fov_deg))
    dirs = get_ray_directions(W, H, f, f, W / 2, H / 2)
    o, d = get_rays(dirs, c2w)
    d = torch.nn.functional.normalize(d, p=2, dim=-1)
    return torch.cat([o, d], -1)  # [H*W, 6]


# ---- analytic SDFs (torch, any device) -----------------------------------
def sdf_sphere(p, r=0.5):
    return p.norm(dim=-1) - r


def sdf_union(p):
    s = (p - p.new_tensor([0.25, 0.0, 0.0])).norm(dim=-1) - 0.35
    q = p - p.new_tensor([-0.2, 0.0, 0.1])
    ring = torch.stack([q[..., :2].norm(dim=-1) - 0.45, q[..., 2]], -1).norm(dim=-1) - 0.12
    return torch.minimum(s, ring)


def albedo_union(p):
    return 0.5 + 0.4 * torch.stack([torch.sin(6 * p[..., 0]), torch.sin(6 * p[..., 1] + 1.0),
                                    torch.cos(6 * p[..., 2])], -1)


def sphere_trace(rays, sdf, n_iter=96, t_max=6.0):
    o, d = rays[:, :3], rays[:, 3:]
    t = torch.zeros(rays.shape[0], device=rays.device)
    for _ in range(n_iter):
        t = t + sdf(o + d * t[:, None]).clamp_min(0)
        t = t.clamp_max(t_max)
    p = o + d * t[:, None]
    hit = sdf(p).abs() < 1e-3
    return p, hit


def shade(rays, sdf, albedo_fn=None, bg=1.0):
    p, hit = sphere_trace(rays, sdf)
    eps = 1e-4
    n = torch.stack([(sdf(p + p.new_tensor(e)) - sdf(p - p.new_tensor(e))) for e in
                     ([eps, 0, 0], [0, eps, 0], [0, 0, eps])], -1)
    n = torch.nn.functional.normalize(n, dim=-1)
    light = torch.nn.functional.normalize(p.new_tensor(LIGHT_DIR), dim=0)
    lam = (n @ light).clamp_min(0.0)[:, None] * 0.8 + 0.2
    alb = albedo_fn(p) if albedo_fn is not None else torch.full_like(p, 0.7)
    rgb = alb * lam
    return torch.where(hit[:, None], rgb, torch.full_like(rgb, bg)), hit


# ---- configs -------------------------------------------------------------
def sphere_hits(rays, radius=1.0):
    o, d = rays[:, :3], rays[:, 3:]
    b = (o * d).sum(-1)
    c = (o * o).sum(-1) - radius * radius
    disc = b * b - c
    s = disc.clamp_min(0).sqrt()
    return disc > 0, (-b - s), (-b + s)


def uniform_samples(rays, n_samples, radius=1.0):
    """n_samples uniform intervals over each ray's chord of the radius sphere
    (rays are chosen to hit it).  Returns ray_indices int32, t_starts, t_ends,
    packed [R,2]."""
    _, tn, tf = sphere_hits(rays, radius)
    tn = tn.clamp_min(0)
    R = rays.shape[0]
    k = torch.arange(n_samples + 1, device=rays.device, dtype=torch.float32) / n_samples
    t = tn[:, None] + (tf - tn)[:, None] * k[None, :]
    t0 = t[:, :-1].reshape(-1).contiguous()
    t1 = t[:, 1:].reshape(-1).contiguous()
    ridx = torch.arange(R, device=rays.device, dtype=torch.int32).repeat_interleave(n_samples)
    packed = torch.stack([torch.arange(R, device=rays.device, dtype=torch.int32) * n_samples,
                          torch.full((R,), n_samples, device=rays.device, dtype=torch.int32)], -1)
    return ridx, t0, t1, packed.contiguous()


def config0_batch(n_rays=1024, n_views=16, W=64, H=64, fov=40.0, seed=0):
    """Plain NeuS workload batch: rays from seeded views that hit the unit
    sphere, targets from the Lambertian radius-0.5 sphere."""
    g = torch.Generator().manual_seed(seed + 1)
    cams = sphere_cameras(n_views, 2.5, seed)
    rays = torch.cat([camera_rays(c, W, H, fov) for c in cams])
    hit, _, _ = sphere_hits(rays, 1.0)
    rays = rays[hit]
    sel = torch.randperm(rays.shape[0], generator=g)[:n_rays]
    rays = rays[sel].contiguous()
    rgb, _ = shade(rays, sdf_sphere)
    return rays, rgb


def config1_pool(n_views=100, W=512, H=512, fov=40.0, seed=0):
    cams = sphere_cameras(n_views, 2.5, seed)
    return cams, (W, H, fov)


def config1_batch(cams, whf, n_rays=8192, gen=None, device="cpu"):
    """Neuralangelo workload batch: random pixels of random views (host draws);
    targets by sphere tracing the sphere + torus union."""
    W, H, fov = whf
    v = torch.randint(0, cams.shape[0], (n_rays,), generator=gen)
    x = torch.randint(0, W, (n_rays,), generator=gen)
    y = torch.randint(0, H, (n_rays,), generator=gen)
    f = 0.5 * W / math.tan(0.5 * math.radians(fov))
    dirs = torch.stack([(x + 0.5 - W / 2) / f, -(y + 0.5 - H / 2) / f, -torch.ones(n_rays)], -1)
    c2w = cams[v]
    o, d = get_rays(dirs, c2w)
    d = torch.nn.functional.normalize(d, p=2, dim=-1)
    rays = torch.cat([o, d], -1).to(device).contiguous()
    rgb, _ = shade(rays, sdf_union, albedo_union)
    return rays, rgb


def config1_images(cams, whf, device, chunk_views=8):
    """The Neuralangelo workload's image set rendered once: images
    [N,H,W,3] (sphere + torus union, white where the rays miss), per-pixel
    camera directions [H,W,3] (models/ray_utils.py:9-20) and c2w [N,3,4], all
    on `device`, for on-device training-ray sampling."""
    W, H, fov = whf
    f = 0.5 * W / math.tan(0.5 * math.radians(fov))
    dirs = get_ray_directions(W, H, f, f, W / 2, H / 2)
    n = cams.shape[0]
    images = torch.empty(n, H, W, 3, device=device)
    for a in range(0, n, chunk_views):
        c = cams[a:a + chunk_views].to(device)
        o, d = get_rays(dirs.to(device), c)
        d = torch.nn.functional.normalize(d, p=2, dim=-1)
        rgb, _ = shade(torch.cat([o, d], -1).contiguous(), sdf_union, albedo_union)
        images[a:a + c.shape[0]] = rgb.view(c.shape[0], H, W, 3)
    return images, dirs.to(device), cams[:, :3, :4].to(device).contiguous()


def prewarm_occupancy(sdf, res=128, radius=1.0, band_cells=1.5, device="cpu"):
    """Binary grid of cells whose centre is within band_cells cells of the
    surface (x-major [res,res,res], uint8)."""
    c = (torch.arange(res, device=device, dtype=torch.float32) + 0.5) / res * 2 * radius - radius
    X, Y, Z = torch.meshgrid(c, c, c, indexing="ij")
    p = torch.stack([X, Y, Z], -1).reshape(-1, 3)
    d = sdf(p).abs()
    cell = 2 * radius / res
    return (d < band_cells * cell).reshape(res, res, res).to(torch.uint8)
