"""Differentiation checks of the hand-written backward kernels.

* FP64 central differences (h = 1e-4) of the oracle, evaluated in float64 on
  the GPU, against the HIP backward of the hash-grid encode (table and
  position), the VanillaMLP (parameters and input) and the NeuS compositing
  (sdf, normal, rgb, inv_s).  4,096 seeded points in [-1,1]^3 and 512 rays x
  256 samples; hash grid 16 levels x 2 features, 2^16 rows per level, base
  32, scale 1.3195...
* The table-gradient reduction is run-to-run bit-identical at these shapes.
* The stored-encoding and the recompute backward of the fused SDF field give
  bit-identical gradients at 2^20 samples.

Every central difference is taken where the function is smooth in the
perturbed coordinate (points whose perturbation crosses a cell boundary at
some level, and samples whose alpha sits at a clip bound, are left out), so
the differences carry only rounding and O(h^2) error.
"""
import numpy as np
import pytest
import torch

from oracle import hashgrid as ohg
from oracle import nets as onets
from oracle import render as orender

import helpers

pytestmark = pytest.mark.gpu

GRID = dict(L=16, F=2, log2T=16, base=32, scale=1.3195079107728942)
H = 1e-4


def _central(f, theta, idx, h=H):
    """d f / d theta[idx] by central differences in float64 (f: theta -> scalar)."""
    out = []
    flat = theta.view(-1)
    for i in idx:
        old = float(flat[i])
        flat[i] = old + h
        fp = float(f(theta))
        flat[i] = old - h
        fm = float(f(theta))
        flat[i] = old
        out.append((fp - fm) / (2.0 * h))
    return np.array(out)


def _points(n, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(n, 3, generator=g) * 2 - 1   # [-1,1]^3


def test_hashgrid_backward_vs_fp64_central_differences(dev):
    from angelo_amd import descs
    from angelo_amd.functions import hashgrid_encode
    L, F, log2T, base, s = GRID["L"], GRID["F"], GRID["log2T"], GRID["base"], GRID["scale"]
    grid = descs.grid_desc(L, F, log2T, base, s)
    lt = ohg.level_table(L, base, s, log2T, F)
    x01 = (_points(4096, 31) + 1.0) * 0.5
    g = torch.Generator().manual_seed(32)
    table = torch.rand(lt.n_entries, F, generator=g) * 2 - 1
    w = torch.randn(x01.shape[0], L * F, generator=g)
    # HIP backward
    xd = x01.to(dev).requires_grad_(True)
    td = table.to(dev).requires_grad_(True)
    (hashgrid_encode(xd, td, grid, L) * w.to(dev)).sum().backward()
    dtab = td.grad.double().cpu()
    dx = xd.grad.double().cpu()
    # FP64 oracle on the GPU
    x64 = x01.double().to(dev)
    w64 = w.double().to(dev)
    t64 = table.double().to(dev)

    def f_table(t):
        return (ohg.encode_torch(x64, t, lt, L, mode="f32") * w64).sum()

    touched = torch.nonzero(dtab.abs().sum(-1) > 0)[:, 0]
    untouched = torch.nonzero(dtab.abs().sum(-1) == 0)[:, 0]
    rng = np.random.default_rng(33)
    rows = np.concatenate([rng.choice(touched.numpy(), 48, replace=False),
                           rng.choice(untouched.numpy(), 16, replace=False)])
    idx = np.concatenate([rows * F, rows * F + 1])
    fd = _central(f_table, t64, idx)
    an = dtab.view(-1)[idx].numpy()
    scale = np.abs(an).max()
    assert np.abs(fd - an).max() <= 1e-5 * scale, np.abs(fd - an).max() / scale
    assert np.all(fd[np.isin(idx // F, rows[48:])] == 0.0)

    # position derivatives: trilinear blending is linear along one axis inside
    # a cell, so the FP64 central difference is exact away from cell
    # boundaries; the kernel's weights along the other two axes come from
    # fp32 positions (fmaf(scale, x, 0.5) carries up to ~2.4e-4 of a cell at
    # the finest level, scale 2048), which bounds its agreement at ~1e-3
    scales = torch.from_numpy(lt.scale.astype(np.float64))
    pts = rng.choice(x01.shape[0], 64, replace=False)
    checked = 0
    for p in pts:
        for d in range(3):
            xp = x64[p:p + 1].clone()
            c0 = torch.floor(scales * float(xp[0, d]) + 0.5)
            cp = torch.floor(scales * (float(xp[0, d]) + H) + 0.5)
            cm = torch.floor(scales * (float(xp[0, d]) - H) + 0.5)
            if not (torch.equal(c0, cp) and torch.equal(c0, cm)):
                continue
            wp = w64[p:p + 1]

            def f_x(xx):
                return (ohg.encode_torch(xx, t64, lt, L, mode="f64") * wp).sum()

            fdx = _central(f_x, xp, [d])[0]
            assert abs(fdx - float(dx[p, d])) <= 2e-3 * max(1.0, abs(fdx)), (p, d, fdx, dx[p, d])
            checked += 1
    assert checked >= 20


def test_mlp_backward_vs_fp64_central_differences(dev):
    from angelo_amd import descs
    from angelo_amd.functions import mlp_apply
    n_in, n_out, nh = 35, 16, 1
    desc = descs.mlp_desc(n_in, n_out, nh, 64, "softplus", "none", 100.0)
    n_params = descs.mlp_n_params(desc)
    g = torch.Generator().manual_seed(41)
    x = torch.randn(4096, n_in, generator=g) * 0.5
    params = torch.randn(n_params, generator=g) * 0.2
    w = torch.randn(4096, n_out, generator=g)
    xd = x.to(dev).requires_grad_(True)
    pd = params.to(dev).requires_grad_(True)
    (mlp_apply(xd, pd, desc) * w.to(dev)).sum().backward()
    dims = onets.mlp_dims(n_in, n_out, nh)
    x64, w64 = x.double().to(dev), w.double().to(dev)

    def f_p(p):
        return (onets.vanilla_mlp(x64, onets.unflatten(p, dims), "softplus", "none", 100.0)
                * w64).sum()

    rng = np.random.default_rng(42)
    idx = rng.choice(n_params, 96, replace=False)
    fd = _central(f_p, params.double().to(dev), idx)
    an = pd.grad.double().cpu().numpy()[idx]
    scale = np.abs(an).max()
    assert np.abs(fd - an).max() <= 2e-4 * scale, np.abs(fd - an).max() / scale
    p64 = params.double().to(dev)

    def f_x(xx):
        return (onets.vanilla_mlp(xx, onets.unflatten(p64, dims), "softplus", "none", 100.0)
                * w64).sum()

    idx_x = rng.choice(x.numel(), 64, replace=False)
    fdx = _central(f_x, x.double().to(dev), idx_x)
    anx = xd.grad.double().cpu().view(-1).numpy()[idx_x]
    assert np.abs(fdx - anx).max() <= 2e-4 * np.abs(anx).max()


def test_composite_backward_vs_fp64_central_differences(dev):
    from angelo_amd import _native as nat, functions as fn
    from angelo_amd import scenes
    n_rays, S = 512, 256
    rays, _ = scenes.config0_batch(n_rays=n_rays, n_views=8, seed=51)
    ridx, t0, t1, packed = scenes.uniform_samples(rays, S)
    n = ridx.shape[0]
    g = torch.Generator().manual_seed(52)
    pos = rays[ridx.long(), 0:3] + rays[ridx.long(), 3:6] * ((t0 + t1) * 0.5)[:, None]
    sdf = pos.norm(dim=-1) - 0.5 + 0.01 * torch.randn(n, generator=g)
    normal = torch.nn.functional.normalize(pos + 0.1 * torch.randn(n, 3, generator=g), dim=-1)
    rgb = torch.rand(n, 3, generator=g)
    inv_s = torch.tensor([20.0])
    ratio = 0.4
    wo, wd = torch.randn(n_rays, generator=g), torch.randn(n_rays, generator=g)
    wr, wn = torch.randn(n_rays, 3, generator=g), torch.randn(n_rays, 3, generator=g)
    cd = nat.CompositeDesc()
    cd.cos_anneal_ratio = ratio
    cd.kahan = 0
    leaves = [t.to(dev).requires_grad_(True) for t in (sdf, normal, rgb, inv_s)]
    _, op, dep, crgb, cn = fn.composite(leaves[0], leaves[1], leaves[2], leaves[3], rays.to(dev),
                                        ridx.to(dev), t0.to(dev), t1.to(dev), packed.to(dev), cd)
    ((op * wo.to(dev)).sum() + (dep * wd.to(dev)).sum() + (crgb * wr.to(dev)).sum()
     + (cn * wn.to(dev)).sum()).backward()
    dirs = rays[ridx.long(), 3:6].double().to(dev)
    t0d, t1d, rid = t0.double().to(dev), t1.double().to(dev), ridx.to(dev)
    W = [x.double().to(dev) for x in (wo, wd, wr, wn)]
    base = [x.double().to(dev) for x in (sdf, normal, rgb, inv_s)]

    def loss_of(vals):
        o = orender.composite(vals[0], vals[1], vals[2], dirs, t0d, t1d, rid, n_rays, vals[3], ratio)
        return ((o["opacity"][:, 0] * W[0]).sum() + (o["depth"][:, 0] * W[1]).sum()
                + (o["comp_rgb"] * W[2]).sum() + (o["comp_normal"] * W[3]).sum())

    with torch.no_grad():
        o0 = orender.composite(base[0], base[1], base[2], dirs, t0d, t1d, rid, n_rays, base[3],
                               ratio)
    alpha = o0["alpha"].cpu().numpy().reshape(-1)
    smooth = np.nonzero((alpha > 1e-3) & (alpha < 1 - 1e-3))[0]
    rng = np.random.default_rng(53)
    for k, name in enumerate(("sdf", "normal", "rgb", "inv_s")):
        theta = base[k]
        if name == "inv_s":
            idx = [0]
        else:
            per = theta.view(n, -1).shape[1]
            samp = rng.choice(smooth, 48, replace=False)
            idx = list(samp * per + rng.integers(0, per, samp.shape[0]))

        def f(t, k=k):
            vals = list(base)
            vals[k] = t
            return loss_of(vals)

        fd = _central(f, theta, idx, h=H if name != "inv_s" else 1e-3)
        an = leaves[k].grad.double().cpu().view(-1).numpy()[idx]
        err = np.abs(fd - an).max() / max(np.abs(an).max(), 1e-12)
        assert err < 1e-3, (name, err)


def test_hashgrid_backward_bit_identical_at_check_shapes(dev):
    from angelo_amd import descs
    from angelo_amd.functions import hashgrid_encode
    L, F, log2T, base, s = GRID["L"], GRID["F"], GRID["log2T"], GRID["base"], GRID["scale"]
    grid = descs.grid_desc(L, F, log2T, base, s)
    x01 = ((_points(4096, 61) + 1.0) * 0.5).to(dev)
    g = torch.Generator().manual_seed(62)
    table = (torch.rand(int(grid.n_entries), F, generator=g) * 2 - 1).to(dev)
    w = torch.randn(4096, L * F, generator=g).to(dev)
    grads = []
    for _ in range(3):
        td = table.clone().requires_grad_(True)
        xd = x01.clone().requires_grad_(True)
        (hashgrid_encode(xd, td, grid, L) * w).sum().backward()
        grads.append((td.grad.clone(), xd.grad.clone()))
    for a, b in grads[1:]:
        assert torch.equal(a, grads[0][0]) and torch.equal(b, grads[0][1])


def test_store_and_recompute_backward_identical_at_2pow20_samples(dev):
    from angelo_amd import functions
    import angelo_amd
    cfg = helpers.model_config(L=16, log2T=16, base=32, scale=GRID["scale"])
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=71)
    model.train()
    model = model.to(dev)
    model.update_step(0, 30000)
    geo = model.geometry
    n = 1 << 20
    g = torch.Generator().manual_seed(72)
    pts = ((torch.rand(n, 3, generator=g) * 2 - 1) * 0.95).to(dev)
    randn = torch.randn(n, 3, generator=g).to(dev)
    grads = {}
    try:
        for mode in ("store", "recompute"):
            functions.set_sdf_backward(mode)
            model.zero_grad(set_to_none=True)
            sdf, grad, normal, feature, lap = geo.fused(pts, True, True, randn)
            (sdf.sum() + grad.square().sum() + lap.sum() + feature.sum()).backward()
            grads[mode] = [p.grad.clone() for p in geo.parameters()]
            del sdf, grad, normal, feature, lap
    finally:
        functions.set_sdf_backward("auto")
    helpers.assert_nontrivial(geo.grid_module.params.grad, "table gradient")
    for a, b in zip(grads["store"], grads["recompute"]):
        assert torch.equal(a, b)
