"""Shared builders for parity tests: a product NeuSModel (angelo_amd) and the
oracle functions evaluated with the *same* parameters."""
import numpy as np
import torch

from oracle import field as ofield
from oracle import hashgrid as ohg
from oracle import nets as onets


def model_config(L=8, log2T=15, base=16, scale=1.5, geo_hidden=1, feature_dim=16, radius=1.0,
                 start_level=None, rad_hidden=2, grid_prune=False, step=None, cap=0, seed=0,
                 grad_type="finite_difference"):
    from angelo_amd.utils import to_config
    return to_config({
        "name": "neus", "radius": radius, "num_samples_per_ray": 64, "ray_chunk": 4096,
        "randomized": True, "grid_prune": grid_prune, "grid_prune_occ_thre": 0.001,
        "cos_anneal_end": 20000, "learned_background": False, "background_color": "white",
        "render_step_size": step, "max_samples_per_ray": cap,
        "variance": {"init_val": 0.3, "modulate": False},
        "geometry": {
            "name": "volume-sdf", "radius": radius, "feature_dim": feature_dim,
            "grad_type": grad_type, "finite_difference_eps": "progressive", "seed": seed,
            "isosurface": None,
            "xyz_encoding_config": {
                "otype": "ProgressiveBandHashGrid", "n_levels": L, "n_features_per_level": 2,
                "log2_hashmap_size": log2T, "base_resolution": base, "per_level_scale": scale,
                "include_xyz": True, "start_level": L if start_level is None else start_level,
                "start_step": 0, "update_steps": 1000},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": geo_hidden, "sphere_init": True,
                                   "sphere_init_radius": 0.5, "weight_norm": True}},
        "texture": {
            # feature (feature_dim + xyz) and the normal, as the reference's
            # input_feature_dim = feature_dim + 6 counts them
            "name": "volume-radiance", "input_feature_dim": feature_dim + 6,
            "dir_encoding_config": {"otype": "SphericalHarmonics", "degree": 4},
            "mlp_network_config": {"otype": "VanillaMLP", "activation": "ReLU",
                                   "output_activation": "none", "n_neurons": 64,
                                   "n_hidden_layers": rad_hidden},
            "color_activation": "sigmoid"},
    })


def randomize_(model, table_scale=0.05, seed=0, geo_feature_scale=0.5):
    """Replaces the near-zero tcnn-style table init by a larger one and gives
    the SDF MLP non-zero weights on the encoding (sphere init sets them to
    exactly 0, which would make every table gradient 0), so the parity tests
    exercise every term.  The SDF keeps its sphere-like xyz part."""
    g = torch.Generator().manual_seed(seed)
    grid = model.geometry.grid_module
    with torch.no_grad():
        grid.params.copy_((torch.rand(grid.params.shape, generator=g) * 2 - 1) * table_scale)
        lins = model.geometry.network.linears()
        first = lins[0]
        if hasattr(first, "weight_v"):
            v = first.weight_v
            v[:, 3:] = torch.randn(v[:, 3:].shape, generator=g) * geo_feature_scale
            first.weight_g.copy_(v.norm(dim=1, keepdim=True))
        else:
            first.weight[:, 3:] = torch.randn(first.weight[:, 3:].shape, generator=g) * geo_feature_scale
        for lin in model.texture.network.linears():
            lin.weight.copy_(torch.randn(lin.weight.shape, generator=g) / lin.weight.shape[1] ** 0.5)
            lin.bias.copy_(torch.randn(lin.bias.shape, generator=g) * 0.1)
    return model


def assert_nontrivial(t, name, min_frac=0.01):
    """A gradient check means something only if the gradient is not ~0."""
    t = t.detach().double().cpu()
    assert bool(torch.isfinite(t).all()), f"{name} has non-finite entries"
    assert float(t.abs().max()) > 0, f"{name} is identically zero"
    assert float((t != 0).double().mean()) >= min_frac or t.numel() < 100, \
        f"{name}: fewer than {min_frac:.0%} non-zero entries"


def level_table(model):
    g = model.geometry.grid_module
    return ohg.level_table(g.n_levels, g.base_resolution, g.per_level_scale, g.log2_hashmap_size,
                           g.n_features_per_level)


def oracle_fns(model, dtype=torch.float64, mode="f64", active=None):
    """Oracle field / radiance callables over leaf copies of the model's
    parameters.  Returns (field, radiance, leaves dict)."""
    lt = level_table(model)
    geo, tex = model.geometry, model.texture
    table = geo.grid_module.table.detach().cpu().to(dtype).clone().requires_grad_(True)
    gflat = geo.network.flat_params().detach().cpu().to(dtype).clone().requires_grad_(True)
    rflat = tex.network.flat_params().detach().cpu().to(dtype).clone().requires_grad_(True)
    gdims = onets.mlp_dims(geo.network.dim_in, geo.network.dim_out, geo.network.n_hidden_layers)
    rdims = onets.mlp_dims(tex.network.dim_in, tex.network.dim_out, tex.network.n_hidden_layers)
    act = active if active is not None else geo.active_levels()

    def encode(x):
        return ohg.encode_torch(x, table, lt, act, mode=mode)

    def mlp_geo(h):
        return onets.vanilla_mlp(h, onets.unflatten(gflat, gdims), "softplus", "none", 100.0)

    field = ofield.field_from(encode, mlp_geo)

    base = getattr(tex, "base_color", 0)

    def radiance(feature, dirs, normal):
        sh = onets.sh_encode((dirs + 1.0) / 2.0, 4)
        h = torch.cat([feature, sh, normal], -1)
        rgb = torch.sigmoid(onets.vanilla_mlp(h, onets.unflatten(rflat, rdims), "relu", "none"))
        if base:   # VolumeDualColor (models/texture.py:51-55)
            rgb = rgb + torch.sigmoid(feature[:, 1:4])
        return rgb

    return field, radiance, dict(table=table, geo=gflat, rad=rflat)


def rel_err(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))
