"""Ray-marching oracle -- test infrastructure only.

Wraps the C restatement (oracle.c) of nerfacc 0.3.3 ray_aabb_intersect and
ray_marching (AABB contraction, cone angle 0) as called at
models/neus.py:209-220.  nerfacc's source is not in the reference; its
published algorithm is restated (constant step, occupancy test at the
interval midpoint, DDA skip to the next voxel boundary in whole steps).

Pins (tests/test_oracle_march.py): with every cell occupied the samples are
exactly the sequential float32 stepping t0, t0+step, ... while the midpoint
is before t_far; samples are ray-sorted, contiguous per ray, have
t_end - t_start == step and midpoints inside occupied cells; the DDA never
skips an occupied cell that a fine brute-force stepping finds; slab-test
hand cases (axis-aligned, grazing, miss -> 1e10, inside -> t_min = 0).
Exact agreement with NVIDIA's nerfacc binary: parity unpinned (its FMA
contraction choices are not in any file here).

Background marching (march_bg; models/neus.py:141-171, unbounded-sphere
contraction, cone-angle steps, no voxel skip) and the early-stop filter
(visible_prefix) are restated the same way.  Pins
(tests/test_oracle_march.py): without a grid the steps follow the
geometric closed form t_{k+1} = t_k (1 + cone) once t cone > step, so the
sample count from the near to the far plane matches the analytic count;
with a grid the samples are exactly the grid-free samples whose midpoint,
contracted by a torch float64 restatement of models/geometry.py:22-28,
falls in an occupied cell (boundary ties excluded); the filter's kept
counts match the float64 closed form exp(-sigma sum dt) for constant
density (ties excluded).
"""
import numpy as np

from . import clib


def ray_aabb(rays, aabb):
    rays = clib.f32(rays)
    o = np.ascontiguousarray(rays[:, 0:3])
    d = np.ascontiguousarray(rays[:, 3:6])
    n = rays.shape[0]
    tn = np.zeros(n, np.float32)
    tf = np.zeros(n, np.float32)
    box = clib.f32(aabb)
    clib.load().or_ray_aabb(clib.p(o), clib.p(d), n, clib.p(box), clib.p(tn), clib.p(tf))
    return tn, tf


def ray_marching(rays, t_near, t_far, roi, grid, step, cap=0, jitter=None):
    """Returns ray_indices int32 [N], t_starts [N], t_ends [N], counts [R].
    grid: uint8 [res,res,res] (x-major) or None (all occupied, roi +-1e10)."""
    rays = clib.f32(rays)
    o = np.ascontiguousarray(rays[:, 0:3])
    d = np.ascontiguousarray(rays[:, 3:6])
    n = rays.shape[0]
    near = clib.f32(t_near).copy()
    if jitter is not None:
        near = (near + (clib.f32(jitter) * np.float32(step)).astype(np.float32)).astype(np.float32)
    far = clib.f32(t_far)
    if grid is None:
        roi = [-1e10, -1e10, -1e10, 1e10, 1e10, 1e10]
        grid = np.ones((1, 1, 1), np.uint8)
    roi = clib.f32(roi)
    g = np.ascontiguousarray(grid, dtype=np.uint8)
    res = g.shape[-1]
    counts = np.zeros(n, np.int32)
    lib = clib.load()
    total = lib.or_ray_marching_count(clib.p(o), clib.p(d), clib.p(near), clib.p(far), n,
                                      clib.p(roi), res, clib.p(g), float(np.float32(step)), int(cap),
                                      clib.p(counts))
    ridx = np.zeros(total, np.int32)
    t0 = np.zeros(total, np.float32)
    t1 = np.zeros(total, np.float32)
    lib.or_ray_marching_fill(clib.p(o), clib.p(d), clib.p(near), clib.p(far), n, clib.p(roi), res,
                             clib.p(g), float(np.float32(step)), int(cap), clib.p(ridx), clib.p(t0),
                             clib.p(t1))
    return ridx, t0, t1, counts


def march_bg(rays, t_near, t_far, roi, grid, step, cone, cap=0):
    """Background marching (nerfacc 0.3.3 ray_marching, UN_BOUNDED_SPHERE,
    cone angle; models/neus.py:141-171) from given per-ray [t_near, t_far]
    (near plane, far plane and jitter already applied).  grid: uint8
    [res,res,res] over the contracted roi, or None (every cell occupied).
    Returns ray_indices int32, t_starts, t_ends, counts."""
    rays = clib.f32(rays)
    o = np.ascontiguousarray(rays[:, 0:3])
    d = np.ascontiguousarray(rays[:, 3:6])
    n = rays.shape[0]
    near, far = clib.f32(t_near), clib.f32(t_far)
    roi = clib.f32(roi)
    if grid is None:
        g, res, gp = None, 0, None
    else:
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        res, gp = g.shape[-1], clib.p(g)
    counts = np.zeros(n, np.int32)
    lib = clib.load()
    st, cn = float(np.float32(step)), float(np.float32(cone))
    total = lib.or_march_bg_count(clib.p(o), clib.p(d), clib.p(near), clib.p(far), n, clib.p(roi),
                                  res, gp, st, cn, int(cap), clib.p(counts))
    ridx = np.zeros(total, np.int32)
    t0 = np.zeros(total, np.float32)
    t1 = np.zeros(total, np.float32)
    lib.or_march_bg_fill(clib.p(o), clib.p(d), clib.p(near), clib.p(far), n, clib.p(roi), res, gp,
                         st, cn, int(cap), clib.p(ridx), clib.p(t0), clib.p(t1))
    return ridx, t0, t1, counts


def visible_prefix(sigma, t_starts, t_ends, counts, eps=1e-4):
    """Per-ray kept prefix length of the early-stop filter (nerfacc 0.3.3
    ray_marching with sigma_fn, alpha_thre 0): T_j = prod_{i<j}(1 - alpha_i),
    alpha = 1 - exp(-sigma dt), kept while T_j >= eps."""
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    starts = (np.cumsum(counts) - counts).astype(np.int32)
    packed = np.ascontiguousarray(np.stack([starts, counts], -1), dtype=np.int32)
    n = counts.shape[0]
    keep = np.zeros(n, np.int32)
    s, a, b = clib.f32(sigma), clib.f32(t_starts), clib.f32(t_ends)
    clib.load().or_visible_prefix(clib.p(s), clib.p(a), clib.p(b), clib.p(packed), n,
                                  float(np.float32(eps)), clib.p(keep))
    return keep


def keep_prefix(ridx, t0, t1, counts, keep):
    """The samples of each ray's kept prefix, ray-sorted."""
    starts = np.cumsum(counts) - counts
    pos = np.arange(ridx.shape[0]) - starts[ridx]
    m = pos < keep[ridx]
    return ridx[m], t0[m], t1[m]
