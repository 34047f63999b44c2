"""Learned background (NeuSModel.forward_bg_, models/neus.py:141-203) on the
HIP path against the oracle: unbounded-sphere cone-angle marching
(bit-exact), the early-stop filter (kept prefix per ray), density
compositing forward and backward, the VolumeDensity + VolumeDualColor
background pass with shared parameters, and NeuS training steps with the
background and both occupancy grids on."""
import math

import numpy as np
import pytest
import torch

from oracle import field as ofield
from oracle import hashgrid as ohg
from oracle import march as omarch
from oracle import nets as onets
from oracle import render as orender

import helpers

pytestmark = pytest.mark.gpu

ROI = [-1.0, -1.0, -1.0, 1.0, 1.0, 1.0]


def _cone(n_per_ray):
    return 10 ** (math.log10(1e3) / n_per_ray) - 1.0


def _rays(n, seed):
    g = torch.Generator().manual_seed(seed)
    o = torch.rand(n, 3, generator=g) - 0.5
    d = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)
    return torch.cat([o, d], -1).float()


def _bg_model_config(L=8, log2T=15, grid_prune=False):
    cfg = helpers.model_config(grid_prune=grid_prune)
    from angelo_amd.utils import to_config
    extra = to_config({
        "learned_background": True, "num_samples_per_ray_bg": 64,
        "grid_prune_occ_thre_bg": 0.01,
        "geometry_bg": {
            "name": "volume-density", "radius": 1.0, "feature_dim": 8,
            "density_activation": "trunc_exp", "density_bias": -1, "isosurface": None, "seed": 5,
            "xyz_encoding_config": {
                "otype": "ProgressiveBandHashGrid", "n_levels": L, "n_features_per_level": 2,
                "log2_hashmap_size": log2T, "base_resolution": 16, "per_level_scale": 1.5,
                "include_xyz": True, "start_level": L, "start_step": 0, "update_steps": 1000},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": 1}},
        "texture_bg": {
            "name": "volume-dual-color", "input_feature_dim": 8,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": 2},
            "color_activation": "sigmoid"}})
    for k in extra:
        cfg[k] = extra[k]
    return cfg


@pytest.mark.parametrize("grid_res,cap", [(0, 0), (64, 0), (32, 9)])
def test_bg_marching_bit_exact(dev, grid_res, cap):
    from angelo_amd import functions as fn
    R = 300
    rays = _rays(R, seed=1)
    g = torch.Generator().manual_seed(2)
    near = (0.1 + torch.rand(R, generator=g) * 0.01).float()
    far = torch.full((R,), 1e3)
    grid = None
    if grid_res:
        grid = (torch.rand(grid_res, grid_res, grid_res, generator=g) < 0.3).to(torch.uint8)
    cone = _cone(64)
    ridx, t0, t1, packed = fn.ray_marching_bg(rays.to(dev), near.to(dev), far.to(dev), ROI,
                                             grid.to(dev) if grid is not None else None, 0.01,
                                             cone, cap)
    o_r, o_0, o_1, o_c = omarch.march_bg(rays.numpy(), near.numpy(), far.numpy(), ROI,
                                        grid.numpy() if grid is not None else None, 0.01, cone,
                                        cap)
    assert ridx.shape[0] == o_r.shape[0] > 0
    assert np.array_equal(ridx.cpu().numpy(), o_r)
    assert np.array_equal(t0.cpu().numpy(), o_0)
    assert np.array_equal(t1.cpu().numpy(), o_1)
    assert np.array_equal(packed[:, 1].cpu().numpy(), o_c)
    assert np.array_equal(packed[:, 0].cpu().numpy(), np.cumsum(o_c) - o_c)


def test_bg_early_stop_filter(dev):
    from angelo_amd import functions as fn
    R = 400
    rays = _rays(R, seed=3)
    near = torch.full((R,), 0.1)
    far = torch.full((R,), 1e3)
    cone = _cone(64)
    seen = {}

    def sigma_fn(t0, t1, ridx):
        # a density that stops some rays early and not others
        gen = torch.Generator(device=t0.device).manual_seed(7)
        per_ray = torch.rand(R, device=t0.device, generator=gen) ** 3 * 40.0
        s = per_ray[ridx.long()] * torch.exp(-0.05 * t0)
        seen["sigma"] = s.clone()
        seen["samples"] = (ridx.clone(), t0.clone(), t1.clone())
        return s

    ridx, t0, t1, packed = fn.ray_marching_bg(rays.to(dev), near.to(dev), far.to(dev), ROI, None,
                                             0.01, cone, 0, sigma_fn, 1e-4)
    a_r, a0, a1 = (t.cpu().numpy() for t in seen["samples"])
    sig = seen["sigma"].cpu().numpy()
    counts = np.bincount(a_r, minlength=R).astype(np.int32)
    keep = omarch.visible_prefix(sig, a0, a1, counts, 1e-4)
    got = packed[:, 1].cpu().numpy()
    # the device exp may differ from the host's in the last place: a ray may
    # differ only where its transmittance sits on the threshold
    starts = np.cumsum(counts) - counts
    differ = np.nonzero(got != keep)[0]
    for r in differ:
        k = min(got[r], keep[r])
        s = sig[starts[r]:starts[r] + k + 1].astype(np.float64)
        dt = (a1 - a0)[starts[r]:starts[r] + k + 1].astype(np.float64)
        T = np.prod(np.exp(-s[:k] * dt[:k]))
        assert abs(T - 1e-4) < 1e-4 * 1e-4, (r, got[r], keep[r], T)
    assert len(differ) <= 2
    assert 0 < keep.sum() < counts.sum()       # the filter did cut some rays
    if len(differ) == 0:
        k_r, k_0, k_1 = omarch.keep_prefix(a_r, a0, a1, counts, keep)
        assert np.array_equal(ridx.cpu().numpy(), k_r)
        assert np.array_equal(t0.cpu().numpy(), k_0)
        assert np.array_equal(t1.cpu().numpy(), k_1)


def test_density_composite_fwd_bwd(dev):
    from angelo_amd import functions as fn
    g = torch.Generator().manual_seed(11)
    counts = torch.tensor([0, 1, 5, 64, 65, 150, 3, 0, 200, 17])
    R = counts.shape[0]
    ridx = torch.repeat_interleave(torch.arange(R), counts)
    n = ridx.shape[0]
    dt = torch.rand(n, generator=g) * 0.1 + 0.001
    t0 = torch.zeros(n)
    starts = torch.cumsum(counts, 0) - counts
    for r in range(R):
        a, c = int(starts[r]), int(counts[r])
        if c:
            t0[a:a + c] = 0.1 + torch.cumsum(dt[a:a + c], 0) - dt[a:a + c]
    t1 = t0 + dt
    density = torch.exp(torch.randn(n, generator=g) * 2.0)
    rgb = torch.rand(n, 3, generator=g)
    packed = torch.stack([starts, counts], -1).to(torch.int32)
    gw = torch.randn(n, generator=g)
    gop = torch.randn(R, generator=g)
    gdep = torch.randn(R, generator=g)
    grgb = torch.randn(R, 3, generator=g)

    d_dev = density.to(dev).requires_grad_(True)
    c_dev = rgb.to(dev).requires_grad_(True)
    w, op, dep, crgb = fn.composite_density(d_dev, c_dev, t0.to(dev), t1.to(dev), packed.to(dev))
    loss = ((w * gw.to(dev)).sum() + (op * gop.to(dev)).sum() + (dep * gdep.to(dev)).sum()
            + (crgb * grgb.to(dev)).sum())
    loss.backward()

    d64 = density.double().requires_grad_(True)
    c64 = rgb.double().requires_grad_(True)
    ow, _ = orender.weights_from_density(d64, t0.double(), t1.double(), ridx, R)
    mid = (t0.double() + t1.double()) / 2
    oop = orender.accumulate(ow, ridx, R)[:, 0]
    odep = orender.accumulate(ow, ridx, R, mid[:, None])[:, 0]
    orgb = orender.accumulate(ow, ridx, R, c64)
    ol = ((ow * gw.double()).sum() + (oop * gop.double()).sum() + (odep * gdep.double()).sum()
          + (orgb * grgb.double()).sum())
    ol.backward()
    # fp32 sums over <= 200 samples per ray
    assert helpers.rel_err(w, ow) < 2e-5
    assert helpers.rel_err(op, oop) < 2e-5
    assert helpers.rel_err(dep, odep) < 2e-5
    assert helpers.rel_err(crgb, orgb) < 2e-5
    assert helpers.rel_err(d_dev.grad, d64.grad) < 1e-4
    assert helpers.rel_err(c_dev.grad, c64.grad) < 2e-5
    assert float(op[0].detach()) == 0.0 and float(op[7].detach()) == 0.0


def _bg_oracle(model, dtype=torch.float64):
    geo, tex = model.geometry_bg, model.texture_bg
    ewn = geo.encoding_with_network
    gm = ewn.encoding.encoding.encoding
    lt = ohg.level_table(gm.n_levels, gm.base_resolution, gm.per_level_scale,
                         gm.log2_hashmap_size, gm.n_features_per_level)
    table = gm.table.detach().cpu().to(dtype).clone().requires_grad_(True)
    gflat = ewn.network.flat_params().detach().cpu().to(dtype).clone().requires_grad_(True)
    rflat = tex.network.flat_params().detach().cpu().to(dtype).clone().requires_grad_(True)
    gdims = onets.mlp_dims(ewn.network.dim_in, ewn.network.dim_out, ewn.network.n_hidden_layers)
    rdims = onets.mlp_dims(tex.network.dim_in, tex.network.dim_out, tex.network.n_hidden_layers)
    act = gm.active_levels

    def field(x):
        h = torch.cat([x * 2.0 - 1.0, ohg.encode_torch(x, table, lt, act, mode="f64")], -1)
        return onets.vanilla_mlp(h, onets.unflatten(gflat, gdims), "relu", "none")

    def radiance(feature, dirs):
        sh = onets.sh_encode((dirs + 1.0) / 2.0, 4)
        c = onets.vanilla_mlp(torch.cat([feature, sh], -1), onets.unflatten(rflat, rdims), "relu",
                              "none")
        rgb = torch.sigmoid(c)
        if getattr(tex, "base_color", 0):
            rgb = rgb + torch.sigmoid(feature[:, 1:4])
        return rgb

    return field, radiance, dict(table=table, geo=gflat, rad=rflat), gm, ewn.network


def test_forward_bg_matches_oracle(dev):
    import angelo_amd
    cfg = _bg_model_config()
    model = angelo_amd.make("neus", cfg)
    g = torch.Generator().manual_seed(4)
    with torch.no_grad():
        gm = model.geometry_bg.encoding_with_network.encoding.encoding.encoding
        gm.params.copy_((torch.rand(gm.params.shape, generator=g) * 2 - 1) * 0.5)
        for net in (model.geometry_bg.encoding_with_network.network, model.texture_bg.network):
            for lin in net.linears():
                lin.weight.copy_(torch.randn(lin.weight.shape, generator=g) / lin.weight.shape[1] ** 0.5)
                lin.bias.copy_(torch.randn(lin.bias.shape, generator=g) * 0.1)
    model.train()
    model = model.to(dev)
    R = 48
    rays = _rays(R, seed=6)
    near = torch.full((R,), 0.1)
    far = torch.full((R,), 3.0)   # a short span keeps the background partly transparent
    o_r, o_0, o_1, o_c = omarch.march_bg(rays.numpy(), near.numpy(), far.numpy(), ROI, None, 0.01,
                                        _cone(32))
    ridx = torch.from_numpy(o_r)
    t0, t1 = torch.from_numpy(o_0), torch.from_numpy(o_1)
    cnt = torch.from_numpy(o_c).long()
    packed = torch.stack([torch.cumsum(cnt, 0) - cnt, cnt], -1).to(torch.int32)
    bgc = torch.tensor([0.2, 0.5, 0.9])
    model.background_color = bgc.to(dev)

    leaves = {}
    for name, net in (("geo", model.geometry_bg.encoding_with_network.network),
                      ("rad", model.texture_bg.network)):
        leaf = net.flat_params().detach().clone().requires_grad_(True)
        net.flat_params = (lambda l=leaf: l)
        leaves[name] = leaf
    out = model.forward_bg_(rays.to(dev), (ridx.to(dev), t0.to(dev), t1.to(dev), packed.to(dev)))
    gg = torch.Generator().manual_seed(9)
    w_rgb = torch.randn(R, 3, generator=gg)
    w_dep = torch.randn(R, 1, generator=gg) * 1e-3
    loss = (out["comp_rgb"] * w_rgb.to(dev)).sum() + (out["depth"] * w_dep.to(dev)).sum()
    loss.backward()

    field, radiance, ol, _, _ = _bg_oracle(model)
    o = orender.forward_bg(rays.double(), ridx, t0.double(), t1.double(), R, field, radiance, 1.0,
                           -1.0, bgc.double(), mode="f64")
    oloss = (o["comp_rgb"] * w_rgb.double()).sum() + (o["depth"] * w_dep.double()).sum()
    oloss.backward()
    assert helpers.rel_err(out["comp_rgb"], o["comp_rgb"]) < 1e-4
    assert helpers.rel_err(out["opacity"], o["opacity"]) < 1e-4
    assert helpers.rel_err(out["depth"], o["depth"]) < 1e-4
    assert helpers.rel_err(out["weights"], o["weights"]) < 1e-4
    assert 0.05 < float(o["opacity"].detach().mean()) < 0.999        # the background is neither empty nor opaque
    helpers.assert_nontrivial(gm.params.grad, "bg table grad")
    assert helpers.rel_err(gm.params.grad.view(-1, 2), ol["table"].grad) < 1e-3
    assert helpers.rel_err(leaves["geo"].grad, ol["geo"].grad) < 1e-3
    assert helpers.rel_err(leaves["rad"].grad, ol["rad"].grad) < 1e-3


def test_neus_steps_with_background(dev):
    import angelo_amd
    from angelo_amd import scenes
    cfg = _bg_model_config(grid_prune=True)
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=1).to(dev)
    model.train()
    model.background_color = torch.rand(3, device=dev)
    rays, gt = scenes.config0_batch(n_rays=256, n_views=4, seed=2)
    rays, gt = rays.to(dev), gt.to(dev)
    params = [p for p in model.parameters() if p.requires_grad]
    opt = torch.optim.Adam(params, lr=1e-3)
    for step in range(3):
        model.update_step(0, step)
        out = model(rays)
        loss = ((out["comp_rgb_full"] - gt) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        assert torch.isfinite(loss)
        if step == 0:
            gm = model.geometry_bg.encoding_with_network.encoding.encoding.encoding
            helpers.assert_nontrivial(gm.params.grad, "bg table grad", min_frac=1e-4)
            for lin in model.texture_bg.network.linears():
                helpers.assert_nontrivial(lin.weight.grad, "bg radiance grad")
        opt.step()
    assert int(out["num_samples_bg"]) > 0
    assert out["comp_rgb_bg"].shape == (256, 3)
    og = model.occupancy_grid_bg
    assert og.binary.shape == (256, 256, 256)
    assert 0 < int(og.binary.sum()) <= og.num_cells
    assert bool(torch.isfinite(og.occs).all())


def test_bg_empty_cases(dev):
    """No background sample at all (far plane before the near plane): empty
    marching outputs, zero opacity / depth / colour, zero-size gradients."""
    from angelo_amd import functions as fn
    R = 7
    rays = _rays(R, seed=8).to(dev)
    near = torch.full((R,), 5.0, device=dev)
    far = torch.full((R,), 1.0, device=dev)
    ridx, t0, t1, packed = fn.ray_marching_bg(rays, near, far, ROI, None, 0.01, _cone(64), 0,
                                             lambda a, b, c: torch.ones_like(a), 1e-4)
    assert ridx.shape[0] == 0 and int(packed[:, 1].sum()) == 0
    density = torch.zeros(0, device=dev, requires_grad=True)
    rgb = torch.zeros(0, 3, device=dev, requires_grad=True)
    w, op, dep, crgb = fn.composite_density(density, rgb, t0, t1, packed)
    assert w.shape[0] == 0 and torch.equal(op, torch.zeros(R, device=dev))
    assert torch.equal(dep, torch.zeros(R, device=dev))
    assert torch.equal(crgb, torch.zeros(R, 3, device=dev))
    (op.sum() + crgb.sum()).backward()
    assert density.grad.shape == (0,) and rgb.grad.shape == (0, 3)


def test_neus_step_with_background_matches_oracle(dev):
    """NeuSModel.forward with learned_background on given foreground and
    background samples: comp_rgb_full = fg + bg (1 - opacity_fg)
    (models/neus.py:263-283), the colour loss and the gradients of the
    foreground table and of every background parameter against the FP64
    oracle (foreground + background passes composed as the reference does)."""
    import angelo_amd
    from angelo_amd import scenes
    cfg = _bg_model_config()
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=3)
    g = torch.Generator().manual_seed(12)
    with torch.no_grad():
        gm = model.geometry_bg.encoding_with_network.encoding.encoding.encoding
        gm.params.copy_((torch.rand(gm.params.shape, generator=g) * 2 - 1) * 0.5)
        for net in (model.geometry_bg.encoding_with_network.network, model.texture_bg.network):
            for lin in net.linears():
                lin.weight.copy_(torch.randn(lin.weight.shape, generator=g) / lin.weight.shape[1] ** 0.5)
                lin.bias.copy_(torch.randn(lin.bias.shape, generator=g) * 0.1)
    model.train()
    model = model.to(dev)
    model.update_step(0, 30000)
    bgc = torch.tensor([0.3, 0.6, 0.1])
    model.background_color = bgc.to(dev)
    n_rays = 48
    rays, gt = scenes.config0_batch(n_rays=n_rays, n_views=4, seed=5)
    ridx, t0, t1, packed = scenes.uniform_samples(rays, 32)
    near = torch.full((n_rays,), 1.5)
    far = torch.full((n_rays,), 4.0)
    b_r, b_0, b_1, b_c = omarch.march_bg(rays.numpy(), near.numpy(), far.numpy(), ROI, None, 0.01,
                                        _cone(32))
    bcnt = torch.from_numpy(b_c).long()
    bpacked = torch.stack([torch.cumsum(bcnt, 0) - bcnt, bcnt], -1).to(torch.int32)
    bridx, bt0, bt1 = torch.from_numpy(b_r), torch.from_numpy(b_0), torch.from_numpy(b_1)
    randn = torch.randn(ridx.shape[0], 3, generator=g)

    # oracle, FP64
    field, radiance, ol = helpers.oracle_fns(model, torch.float64, "f64")
    bfield, bradiance, bol, _, _ = _bg_oracle(model)
    inv_s = torch.exp(model.variance.variance.detach().cpu().double() * 10.0).clip(1e-6, 1e6)
    eps = float(model.geometry._finite_difference_eps)
    o = orender.neus_forward(rays.double(), ridx, t0.double(), t1.double(), n_rays, field,
                             radiance, inv_s, 1.0, eps, model.cos_anneal_ratio, randn.double())
    ob = orender.forward_bg(rays.double(), bridx, bt0.double(), bt1.double(), n_rays, bfield,
                            bradiance, 1.0, -1.0, bgc.double(), mode="f64")
    o_full = o["comp_rgb"] + ob["comp_rgb"] * (1.0 - o["opacity"])
    o_loss = ((o_full - gt.double()) ** 2).mean()
    o_loss.backward()

    # HIP path
    leaves = {}
    for name, net in (("bgeo", model.geometry_bg.encoding_with_network.network),
                      ("brad", model.texture_bg.network)):
        leaf = net.flat_params().detach().clone().requires_grad_(True)
        net.flat_params = (lambda l=leaf: l)
        leaves[name] = leaf
    out = model(rays.to(dev), samples=(ridx.to(dev), t0.to(dev), t1.to(dev), packed.to(dev)),
                randn=randn.to(dev),
                samples_bg=(bridx.to(dev), bt0.to(dev), bt1.to(dev), bpacked.to(dev)))
    loss = ((out["comp_rgb_full"] - gt.to(dev)) ** 2).mean()
    loss.backward()
    assert helpers.rel_err(out["comp_rgb_full"], o_full) < 1e-4
    assert helpers.rel_err(out["comp_rgb_bg"], ob["comp_rgb"]) < 1e-4
    assert abs(loss.item() - o_loss.item()) / o_loss.item() < 1e-4
    assert int(out["num_samples_bg"]) == bridx.shape[0]
    assert int(out["num_samples_full"]) == bridx.shape[0] + ridx.shape[0]
    assert set(["comp_rgb_bg", "opacity_bg", "depth_bg", "weights_bg", "points_bg",
                "intervals_bg", "ray_indices_bg", "rays_valid_full"]) <= set(out.keys())
    gm = model.geometry_bg.encoding_with_network.encoding.encoding.encoding
    helpers.assert_nontrivial(bol["table"].grad, "oracle bg table gradient")
    assert helpers.rel_err(gm.params.grad.view(-1, 2), bol["table"].grad) < 1e-3
    assert helpers.rel_err(leaves["bgeo"].grad, bol["geo"].grad) < 1e-3
    assert helpers.rel_err(leaves["brad"].grad, bol["rad"].grad) < 1e-3
    helpers.assert_nontrivial(ol["table"].grad, "oracle fg table gradient")
    assert helpers.rel_err(model.geometry.grid_module.params.grad.view(-1, 2),
                           ol["table"].grad) < 2e-3


def test_bg_occupancy_update_duplicate_cells_take_the_max(dev):
    """The background grid's update with a cell drawn three times (different
    jitter): that cell becomes max(old * decay, max of its three estimates),
    the same on every run; a cell drawn once follows the plain EMA rule."""
    import angelo_amd
    from angelo_amd.neus import contract_inv_unbounded
    cfg = _bg_model_config(grid_prune=True)
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=5).to(dev)
    og = model.occupancy_grid_bg
    R = og.resolution
    c0 = (R // 2) * R * R + (R // 2) * R + R // 4        # inside the contracted ball
    c1 = (R // 2) * R * R + (R // 3) * R + R // 2
    cells = torch.tensor([c0, c1, c0, c0], device=dev, dtype=torch.int32)
    jitter = torch.rand(4, 3, device=dev, generator=torch.Generator(device=dev).manual_seed(6))
    decay = 0.95
    with torch.no_grad():
        idx = cells.long()
        coords = torch.stack([idx // (R * R), (idx // R) % R, idx % R], -1).float()
        pts = contract_inv_unbounded((coords + jitter) / R, og.roi_aabb)
        est = model.geometry_bg(pts)[0] * model.render_step_size_bg
    results = []
    for _ in range(2):
        og.occs.fill_(0.0)
        og.occs[c0] = 1e-6
        og.occs[c1] = 1e-6
        model._occ_update_bg(cells, jitter, 4, 0.01, decay)
        results.append(og.occs.clone())
    assert torch.equal(results[0], results[1])
    e0 = torch.maximum(torch.tensor(1e-6, device=dev) * decay, est[[0, 2, 3]].max())
    e1 = torch.maximum(torch.tensor(1e-6, device=dev) * decay, est[1])
    assert torch.equal(results[0][c0], e0)
    assert torch.equal(results[0][c1], e1)
    assert int((results[0] != 0).sum()) == 2


def _contract_points(n, seed):
    g = torch.Generator().manual_seed(seed)
    p = torch.randn(n, 3, generator=g, dtype=torch.float64) * torch.logspace(
        -2, 3, n, dtype=torch.float64)[:, None]
    edge = torch.tensor([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.6, 0.8, 0.0],
                         [1e6, -1e6, 1e6]], dtype=torch.float64)
    return torch.cat([p, edge]).float()


def test_contract_unbounded_matches_oracle(dev):
    """The HIP unbounded-sphere contraction (VolumeDensity's
    contract_to_unisphere, models/geometry.py:19-31) against the oracle's
    float64 restatement on the same fp32 points: the unit cube, the unit
    ball's boundary, points up to 1e6 out; outputs in [0, 1]."""
    from angelo_amd import functions as fn
    from oracle import field as ofield
    p = _contract_points(20000, 31)
    for r in (1.0, 0.5, 1.5):
        got = fn.contract_unbounded(p.to(dev), r).cpu().double()
        ref = ofield.contract_unbounded(p.double(), r, mode="f64")
        assert torch.allclose(got, ref, rtol=0, atol=1e-6), float((got - ref).abs().max())
        assert float(got.min()) >= 0.0 and float(got.max()) <= 1.0
    e = torch.zeros(0, 3, device=dev)
    assert fn.contract_unbounded(e, 1.0).shape == (0, 3)


def test_contract_unbounded_backward(dev):
    """J^T g against float64 autograd through the oracle's restatement, and
    the on-axis closed form: at y = (m, 0, 0), m > 1 (y = x / r), the
    contraction's Jacobian is diag(1/m^2, (2 - 1/m)/m, (2 - 1/m)/m) / (4r)."""
    from angelo_amd import functions as fn
    from oracle import field as ofield
    p = _contract_points(20000, 32)
    g = torch.randn(p.shape, generator=torch.Generator().manual_seed(33))
    for r in (1.0, 0.5):
        x = p.to(dev).requires_grad_(True)
        fn.contract_unbounded(x, r).backward(g.to(dev))
        x64 = p.double().requires_grad_(True)
        ofield.contract_unbounded(x64, r, mode="f64").backward(g.double())
        scale = x64.grad.abs().clamp_min(1e-3)
        rel = ((x.grad.cpu().double() - x64.grad).abs() / scale).max()
        assert float(rel) < 1e-4, float(rel)
    r = 0.5
    for m in (1.5, 4.0, 100.0):
        x = torch.tensor([[m * r, 0.0, 0.0]], device=dev)
        jac = []
        for d in range(3):
            xi = x.clone().requires_grad_(True)
            fn.contract_unbounded(xi, r)[0, d].backward()
            jac.append(float(xi.grad[0, d]))
        want = [1.0 / m ** 2 / (4 * r), (2 - 1 / m) / m / (4 * r), (2 - 1 / m) / m / (4 * r)]
        for a, b in zip(jac, want):
            assert abs(a - b) <= 1e-5 * abs(b), (m, jac, want)
