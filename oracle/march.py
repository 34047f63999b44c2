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
