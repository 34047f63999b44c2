"""Pins of the ray-marching oracle (nerfacc 0.3.3 restatement)."""
import numpy as np

from oracle import march as omarch

AABB = [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]


def _rays(n, seed=0):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=-1, keepdims=True) * 2.5
    tgt = rng.uniform(-0.6, 0.6, size=(n, 3))
    d = tgt - o
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return np.concatenate([o, d], -1).astype(np.float32)


def test_slab_hand_cases():
    rays = np.array([[0, 0, -3, 0, 0, 1], [0, 3, 0, 1, 0, 0], [0.1, 0.2, 0.3, 0.6, 0, 0.8],
                     [-3, -3, -3, 0.57735026, 0.57735026, 0.57735026]], np.float32)
    tn, tf = omarch.ray_aabb(rays, AABB)
    assert tn[0] == 2.0 and tf[0] == 4.0
    assert tn[1] == 1e10 and tf[1] == 1e10          # miss
    assert tn[2] == 0.0 and tf[2] > 0               # origin inside: t_min clipped to 0
    assert abs(tn[3] - 2 * np.sqrt(3)) < 1e-5 and abs(tf[3] - 4 * np.sqrt(3)) < 1e-5


def test_all_occupied_is_sequential_stepping():
    rays = _rays(50)
    tn, tf = omarch.ray_aabb(rays, AABB)
    step = np.float32(2.0 / 128)
    grid = np.ones((8, 8, 8), np.uint8)
    ridx, t0, t1, counts = omarch.ray_marching(rays, tn, tf, AABB, grid, step)
    base = 0
    for r in range(50):
        a = np.float32(tn[r])
        b = np.float32(a + step)
        ref = []
        while np.float32((a + b) * np.float32(0.5)) < tf[r]:
            ref.append((a, b))
            a = b
            b = np.float32(a + step)
        assert counts[r] == len(ref)
        got = list(zip(t0[base:base + counts[r]], t1[base:base + counts[r]]))
        assert got == ref
        base += counts[r]


def _cell_of(p, res):
    u = (p - (-1.0)) / 2.0 * res
    return np.clip(u.astype(np.int64), 0, res - 1)


def test_invariants_and_no_skipped_occupancy():
    rays = _rays(200, 1)
    tn, tf = omarch.ray_aabb(rays, AABB)
    res = 16
    rng = np.random.default_rng(3)
    grid = (rng.random((res, res, res)) < 0.15).astype(np.uint8)
    step = np.float32(2.0 / 256)
    ridx, t0, t1, counts = omarch.ray_marching(rays, tn, tf, AABB, grid, step)
    assert (np.diff(ridx) >= 0).all()                                 # ray-sorted
    assert np.array_equal(np.bincount(ridx, minlength=200), counts)   # contiguous per ray
    np.testing.assert_allclose(t1 - t0, step, rtol=1e-3)
    for i in range(len(ridx)):
        o, d = rays[ridx[i], :3], rays[ridx[i], 3:]
        mid = (t0[i] + t1[i]) * 0.5
        p = o + d * mid
        assert (p >= -1 - 1e-5).all() and (p <= 1 + 1e-5).all()
        c = _cell_of(p, res)
        assert grid[c[0], c[1], c[2]]
    # naive stepping (no DDA) finds the same lattice of occupied midpoints
    for r in range(200):
        o, d = rays[r, :3].astype(np.float64), rays[r, 3:].astype(np.float64)
        t = float(tn[r]) + float(step) / 2
        naive = []
        while t < tf[r]:
            p = o + d * t
            c = _cell_of(p, res)
            if (p >= -1).all() and (p <= 1).all() and grid[c[0], c[1], c[2]]:
                naive.append(t)
            t += float(step)
        got = (t0[ridx == r] + t1[ridx == r]) / 2
        assert abs(len(got) - len(naive)) <= 2
        if len(got) and len(naive):
            # every marched midpoint sits on the naive lattice
            dist = np.abs(got[:, None] - np.array(naive)[None, :]).min(1)
            assert dist.max() < 1e-4


def test_cap_and_jitter():
    rays = _rays(60, 2)
    tn, tf = omarch.ray_aabb(rays, AABB)
    step = np.float32(2.0 / 512)
    jit = np.random.default_rng(4).random(60).astype(np.float32)
    _, t0, _, c = omarch.ray_marching(rays, tn, tf, AABB, None, step, cap=37, jitter=jit)
    assert c.max() == 37
    first = np.cumsum(c) - c
    np.testing.assert_array_equal(t0[first[c > 0]], (tn + jit * step).astype(np.float32)[c > 0])


# ---- background marching (unbounded-sphere contraction, cone angle) ----
def _bg_rays(n, seed=0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.5, 0.5, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return np.concatenate([o, d], -1).astype(np.float32)


def test_bg_steps_follow_cone_closed_form():
    n_per_ray = 64
    cone = np.float32(10 ** (np.log10(1e3) / n_per_ray) - 1.0)
    step = np.float32(0.01)
    rays = _bg_rays(8)
    near = np.full(8, 0.1, np.float32)
    far = np.full(8, 1e3, np.float32)
    ridx, t0, t1, counts = omarch.march_bg(rays, near, far, [-1, -1, -1, 1, 1, 1], None, step, cone)
    # linear region up to t* = step / cone, then geometric growth by (1 + cone)
    t_star = float(step) / float(cone)
    expect = (t_star - 0.1) / float(step) + np.log(1e3 / t_star) / np.log1p(float(cone))
    assert np.all(np.abs(counts - expect) <= 2), (counts, expect)
    base = 0
    for r in range(8):
        a, b = t0[base:base + counts[r]], t1[base:base + counts[r]]
        assert a[0] == near[r]
        assert np.all(a[1:] == b[:-1])                      # contiguous intervals
        dt = np.maximum(a.astype(np.float64) * float(cone), float(step))
        assert np.allclose(b - a, dt, rtol=2e-5, atol=0)    # dt = clamp(t0 cone, step, 1e10)
        assert ((a[-1] + b[-1]) / 2) < 1e3
        assert np.all(ridx[base:base + counts[r]] == r)
        base += counts[r]


def test_bg_grid_keeps_exactly_occupied_midpoints():
    import torch
    from oracle import field as ofield
    R = 16
    rng = np.random.default_rng(3)
    grid = (rng.uniform(size=(R, R, R)) < 0.5).astype(np.uint8)
    cone = np.float32(10 ** (3.0 / 32) - 1.0)
    step = np.float32(0.01)
    rays = _bg_rays(64, seed=4)
    near = np.full(64, 0.1, np.float32)
    far = np.full(64, 1e3, np.float32)
    roi = [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]
    a_r, a0, a1, ac = omarch.march_bg(rays, near, far, roi, None, step, cone)
    g_r, g0, g1, gc = omarch.march_bg(rays, near, far, roi, grid, step, cone)
    mid = (a0.astype(np.float64) + a1) / 2
    pts = rays[a_r, :3].astype(np.float64) + rays[a_r, 3:].astype(np.float64) * mid[:, None]
    u = ofield.contract_unbounded(torch.from_numpy(pts), 1.0, mode="f64").numpy() * R
    frac = u - np.floor(u)
    tie = np.any((frac < 1e-4) | (frac > 1 - 1e-4), axis=-1)
    ix = np.clip(np.floor(u).astype(np.int64), 0, R - 1)
    occ = grid[ix[:, 0], ix[:, 1], ix[:, 2]] != 0
    kept = set(zip(g_r.tolist(), g0.tolist(), g1.tolist()))
    for k in range(a_r.shape[0]):
        if tie[k]:
            continue
        assert ((a_r[k], a0[k], a1[k]) in kept) == occ[k], k
    assert gc.sum() == g_r.shape[0] and gc.sum() < ac.sum()
    # ray-sorted output
    assert np.all(np.diff(g_r) >= 0)


def test_visible_prefix_constant_density_closed_form():
    step = np.float32(0.01)
    rays = _bg_rays(40, seed=5)
    near = np.zeros(40, np.float32)
    far = np.full(40, 3.0, np.float32)
    ridx, t0, t1, counts = omarch.march_bg(rays, near, far, [-1, -1, -1, 1, 1, 1], None, step, 0.0)
    sig = np.geomspace(1.0, 1000.0, 40).astype(np.float32)
    keep = omarch.visible_prefix(sig[ridx], t0, t1, counts, 1e-4)
    for r in range(40):
        s = float(sig[r]) * float(step)
        j = np.log(1e4) / s          # T_j = exp(-j s) >= 1e-4  <=>  j <= ln(1e4)/s
        if abs(j - round(j)) < 1e-3:
            continue
        assert keep[r] == min(counts[r], int(np.floor(j)) + 1), (r, keep[r], j, counts[r])
    kr, k0, k1 = omarch.keep_prefix(ridx, t0, t1, counts, keep)
    assert kr.shape[0] == keep.sum()
