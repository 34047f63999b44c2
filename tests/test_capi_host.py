"""CPU tests of the C-ABI boundary (load, exports, host-only entry points,
argument validation) and of the host-side Python mirror of the reference
API (config access, schedules, update_step protocol, state-dict names,
loss scaling, loud failure without the device path)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "angelo_amd.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(am_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from angelo_amd import _native as nat
    lib = ctypes.CDLL(nat.lib_path())
    declared = _declared()
    assert len(declared) >= 25
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding knows every declared entry point
    assert set(declared) <= set(nat.exported_symbols())


def _declared_arities():
    """name -> (parameter count, list of 'ptr'|'int64'|'int32'|'float'|'size')
    parsed from the prototypes in the header."""
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(am_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        kinds = []
        for p in params:
            if "*" in p:
                kinds.append("ptr")
            elif p.startswith("int64_t"):
                kinds.append("int64")
            elif p.startswith("size_t"):
                kinds.append("size")
            elif p.startswith("float"):
                kinds.append("float")
            else:
                kinds.append("int32")
        out[m.group(1)] = kinds
    return out


def test_ctypes_signatures_match_header_prototypes():
    from angelo_amd import _native as nat
    decl = _declared_arities()
    assert len(decl) >= 25
    kind_of = {ctypes.c_void_p: "ptr", ctypes.c_char_p: "ptr", ctypes.c_int64: "int64",
               ctypes.c_size_t: "size", ctypes.c_float: "float", ctypes.c_int32: "int32"}
    for name, kinds in decl.items():
        _, args = nat._SIGS[name]
        got = [kind_of[a] for a in args]
        assert got == kinds, (name, got, kinds)


def test_grid_init_matches_oracle_level_table():
    from angelo_amd import descs
    from oracle import hashgrid as ohg
    for cfg in [(8, 2, 15, 16, 1.5), (16, 2, 21, 32, 1.3195079107728942), (16, 2, 19, 32, 1.3195079107728942),
                (12, 4, 14, 8, 2.0)]:
        L, F, log2T, base, s = cfg
        g = descs.grid_arrays(descs.grid_desc(L, F, log2T, base, s))
        lt = ohg.level_table(L, base, s, log2T, F)
        assert np.array_equal(g["scale"], lt.scale)
        assert np.array_equal(g["res"], lt.res)
        assert np.array_equal(g["offset"], lt.offset)
        assert np.array_equal(g["size"], lt.size)
        assert np.array_equal(g["hashed"], lt.hashed)
        assert g["n_entries"] == lt.n_entries


def test_grid_init_override_and_validation():
    from angelo_amd import _native as nat, descs
    g = descs.grid_desc(3, 2, 12, 16, 1.5, res_override=[16, 25, 40])
    assert list(g.res[:3]) == [16, 25, 40] and list(g.scale[:3]) == [15.0, 24.0, 39.0]
    with pytest.raises(nat.NativeError, match="n_levels"):
        descs.grid_desc(0, 2, 12, 16, 1.5)
    with pytest.raises(nat.NativeError, match="n_features"):
        descs.grid_desc(4, 3, 12, 16, 1.5)


def test_mlp_param_count_and_validation():
    from angelo_amd import _native as nat, descs
    assert descs.mlp_n_params(descs.mlp_desc(35, 16, 1)) == 64 * 35 + 64 + 16 * 64 + 16
    assert descs.mlp_n_params(descs.mlp_desc(38, 3, 2)) == 64 * 38 + 64 + 64 * 64 + 64 + 3 * 64 + 3
    with pytest.raises(nat.NativeError):
        descs.mlp_n_params(descs.mlp_desc(35, 16, 1, n_neurons=32))


def test_device_entry_points_reject_bad_arguments_without_launching():
    """Validation happens on the host before any launch (no GPU needed)."""
    from angelo_amd import _native as nat, descs
    lib = nat.load()
    g = descs.grid_desc(8, 2, 15, 16, 1.5)
    rc = lib.am_hashgrid_fwd(ctypes.byref(g), None, 10, None, 8, None, None)
    assert rc == -1 and b"NULL" in lib.am_last_error()
    rc = lib.am_hashgrid_fwd(ctypes.byref(g), None, 10, None, 99, None, None)
    assert rc == -1
    assert lib.am_hashgrid_fwd(ctypes.byref(g), None, 0, None, 8, None, None) == 0  # empty batch


def test_product_has_no_cpu_fallback():
    import angelo_amd
    from angelo_amd import _native as nat
    from angelo_amd.workloads import model_config
    model = angelo_amd.make("neus", model_config(8, 15, 16, 1.5, rays=8, samples_per_ray=8))
    model.update_step(0, 100)
    with pytest.raises(nat.NativeError, match="no CPU fallback"):
        model.geometry.forward_level(torch.zeros(4, 3))


def test_host_tensors_never_reach_a_kernel():
    """The occupancy update of a model still on the host (update_step at a
    multiple of 16 with grid_prune on, in training mode) and every raw
    pointer hand-off refuse host tensors before any launch: a host address
    in a kernel argument faults the GPU."""
    import angelo_amd
    from angelo_amd import _native as nat
    from angelo_amd.workloads import model_config
    model = angelo_amd.make("neus", model_config(8, 15, 16, 1.5, rays=8, samples_per_ray=8,
                                                 grid_prune=True, step=2.0 / 256))
    model.train()
    model.update_step(0, 30001)        # no occupancy update at this step
    with pytest.raises(nat.NativeError, match="device tensors"):
        model.update_step(0, 30000)
    with pytest.raises(nat.NativeError, match="device tensors"):
        nat.ptr(torch.zeros(3))
    assert nat.ptr(None) is None


def test_state_dict_names_mirror_reference():
    import angelo_amd
    from angelo_amd.workloads import model_config
    model = angelo_amd.make("neus", model_config(8, 15, 16, 1.5, rays=8, samples_per_ray=8,
                                                 geo_hidden=2))
    keys = set(model.state_dict().keys())
    for k in ["geometry.encoding.encoding.encoding.params", "geometry.network.layers.0.weight_g",
              "geometry.network.layers.0.weight_v", "geometry.network.layers.0.bias",
              "geometry.network.layers.2.weight_g", "geometry.network.layers.4.bias",
              "texture.network.layers.0.weight", "texture.network.layers.4.bias",
              "variance.variance", "scene_aabb"]:
        assert k in keys, k
    assert model.geometry.grid_module.params.numel() == 214528 * 2
    p = model.geometry.grid_module.params
    assert p.abs().max() <= 1e-4  # tcnn's uniform +-1e-4 init


def test_update_step_protocol():
    import angelo_amd
    from angelo_amd.workloads import model_config
    cfg = model_config(16, 19, 32, 1.3195079107728942, rays=8, samples_per_ray=8, start_level=4)
    cfg.geometry.xyz_encoding_config.start_step = 5000
    model = angelo_amd.make("neus", cfg)
    geo = model.geometry
    model.update_step(0, 0)
    assert geo.active_levels() == 4
    # progressive eps = 2 r / (base * scale^(level-1)) in float64 (geometry.py:304-309)
    assert geo._finite_difference_eps == 2 * 1.0 / (32 * 1.3195079107728942 ** 3)
    model.update_step(0, 7500)
    assert geo.active_levels() == 6
    model.update_step(0, 100000)
    assert geo.active_levels() == 16
    assert model.cos_anneal_ratio == 1.0
    model.update_step(0, 5000)          # the mask only gains levels (network_utils.py:66)
    assert geo.active_levels() == 16
    assert abs(model.cos_anneal_ratio - 0.25) < 1e-12


def test_scheduled_values_and_config():
    from angelo_amd.utils import scheduled_value, to_config, chunk_batch, get_activation
    assert scheduled_value([0, 0, 0.5, 5000], 2500) == 0.25
    assert scheduled_value([0, 0, 0.5, 5000], 9000) == 0.5
    assert scheduled_value([1, 0.0, 1.0, 3.0], 0, current_epoch=2) == 0.5
    assert scheduled_value(0.1, 123) == 0.1
    c = to_config({"a": {"b": 2}, "l": [{"x": 1}]})
    assert c.a.b == 2 and c.get("z", 5) == 5 and c.l[0].x == 1
    out = chunk_batch(lambda t: {"y": t * 2}, 3, False, torch.arange(10))
    assert torch.equal(out["y"], torch.arange(10) * 2)
    assert torch.equal(get_activation("scale2")(torch.tensor([3.0])), torch.tensor([1.0]))


def test_dynamic_loss_scaler():
    from angelo_amd.system import DynamicLossScaler
    p = torch.nn.Parameter(torch.ones(3))
    opt = torch.optim.SGD([p], lr=0.1)
    s = DynamicLossScaler(init_scale=1024.0, growth_interval=2)
    p.grad = torch.tensor([1.0, float("inf"), 0.0])
    assert not s.step(opt, [p]) and s.scale == 512.0 and torch.equal(p.data, torch.ones(3))
    p.grad = torch.tensor([512.0, 0.0, 0.0])
    assert s.step(opt, [p]) and abs(p.data[0].item() - 0.9) < 1e-7
    p.grad = torch.tensor([0.0, 0.0, 0.0])
    s.step(opt, [p])
    assert s.scale == 1024.0
    # checkpoint state: a restored scaler continues the growth count
    p.grad = torch.tensor([0.0, 0.0, 0.0])
    s.step(opt, [p])
    r = DynamicLossScaler()
    r.load_state_dict(s.state_dict())
    assert r.state_dict() == s.state_dict() and r.scale == 1024.0 and r._good == 1
    p.grad = torch.tensor([0.0, 0.0, 0.0])
    r.step(opt, [p])
    assert r.scale == 2048.0


def test_losses_on_cpu_tensors():
    from angelo_amd.system import NeuSTrainer
    from angelo_amd.workloads import system_config
    from angelo_amd.utils import to_config
    tr = NeuSTrainer(None, system_config(), to_config({"train_num_rays": 4, "num_samples_per_ray": 8}))
    tr.global_step = 10000
    out = {"rays_valid_full": torch.tensor([[True], [False], [True], [True]]),
           "comp_rgb_full": torch.full((4, 3), 0.5), "sdf_grad_samples": torch.tensor([[0.0, 2.0, 0.0]]),
           "opacity": torch.ones(4, 1), "sdf_samples": torch.zeros(1),
           "sdf_laplace_samples": torch.tensor([[0.2], [-0.4]])}
    loss, terms = tr.losses(out, {"rgb": torch.zeros(4, 3)})
    assert abs(terms["rgb_mse"].item() - 0.25) < 1e-7
    assert abs(terms["eikonal"].item() - 1.0) < 1e-7
    assert abs(terms["curvature"].item() - 0.3) < 1e-7
    assert abs(loss.item() - (2.5 + 0.1 + 0.15)) < 1e-6


def _trainer(**model_keys):
    from angelo_amd.system import NeuSTrainer
    from angelo_amd.workloads import system_config
    from angelo_amd.utils import to_config
    mc = {"train_num_rays": 4, "num_samples_per_ray": 8}
    mc.update(model_keys)
    tr = NeuSTrainer(None, system_config(), to_config(mc))
    tr.global_step = 10000
    return tr


def _loss_out():
    return {"rays_valid_full": torch.tensor([[True], [True]]),
            "comp_rgb_full": torch.full((2, 3), 0.5), "sdf_grad_samples": torch.tensor([[0.0, 1.0, 0.0]]),
            "opacity": torch.tensor([[0.25], [0.75]]), "sdf_samples": torch.zeros(1),
            "sdf_laplace_samples": torch.zeros(1, 1)}


def test_mask_loss_counts_only_with_dataset_masks():
    """systems/neus.py:143-145: the BCE mask term is weighted by lambda_mask
    only when the dataset has masks; the sampler's all-ones fg_mask of a
    mask-less dataset adds nothing."""
    import math
    tr = _trainer()
    tr.cfg.loss["lambda_mask"] = 0.5
    batch = {"rgb": torch.full((2, 3), 0.5), "fg_mask": torch.ones(2)}
    base, terms = tr.losses(_loss_out(), batch)
    assert "mask" not in terms
    with_mask, terms = tr.losses(_loss_out(), dict(batch, has_mask=True))
    bce = -(math.log(0.25) + math.log(0.75)) / 2     # BCE(opacity, 1)
    assert abs(terms["mask"].item() - bce) < 1e-6
    assert abs((with_mask - base).item() - 0.5 * bce) < 1e-5
    tr.has_mask = True
    assert "mask" in tr.losses(_loss_out(), batch)[1]


def test_distortion_losses_raise_when_weighted():
    """A non-zero distortion weight raises (the efficient distortion loss is
    not built); the background's counts only with the learned background,
    as at systems/neus.py:172."""
    import pytest
    batch = {"rgb": torch.zeros(2, 3)}
    tr = _trainer()
    tr.cfg.loss["lambda_distortion_bg"] = 0.1
    tr.losses(_loss_out(), batch)          # no learned background: not counted
    tr = _trainer(learned_background=True)
    tr.cfg.loss["lambda_distortion_bg"] = 0.1
    with pytest.raises(NotImplementedError):
        tr.losses(_loss_out(), batch)
    tr = _trainer()
    tr.cfg.loss["lambda_distortion"] = 0.1
    with pytest.raises(NotImplementedError):
        tr.losses(_loss_out(), batch)


def test_sparse_point_losses_closed_form():
    """systems/neus.py:180-194 on a sphere SDF |p| - 0.5 (gradient p/|p|):
    the l1 term is mean|sdf| times the mean confidence, the normal term 0
    for the true normals and 1 for perpendicular ones; lambda_normal
    defaults to lambda_sdf_l1."""
    class Sphere:
        def __call__(self, pts, with_grad=True, with_feature=False):
            r = pts.norm(dim=-1)
            return r - 0.5, pts / r[:, None]

    class M:
        geometry = Sphere()

    tr = _trainer()
    tr.model = M()
    pts = torch.tensor([[1.0, 0.0, 0.0], [0.0, 0.25, 0.0], [0.0, 0.0, -0.5]])
    w = torch.tensor([1.0, 0.5, 0.0])
    batch = {"rgb": torch.full((2, 3), 0.5), "pts": pts, "pts_weights": w,
             "pts_normal": pts * 3.0}
    base = tr.losses(_loss_out(), {"rgb": batch["rgb"]})[0]
    tr.cfg.loss["lambda_sdf_l1"] = 2.0
    loss, terms = tr.losses(_loss_out(), batch)
    l1 = (0.5 + 0.25 + 0.0) / 3 * (1.5 / 3)
    assert abs(terms["sdf_l1"].item() - l1) < 1e-7
    assert abs(terms["normal_cos"].item()) < 1e-7
    assert abs((loss - base).item() - 2.0 * l1) < 1e-6
    perp = torch.tensor([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    _, terms = tr.losses(_loss_out(), dict(batch, pts_normal=perp))
    assert abs(terms["normal_cos"].item() - 1.0) < 1e-7


def test_oracle_losses_closed_form():
    """The oracle's loss (oracle/render.py, systems/neus.py:130-159 with the
    Neuralangelo weights 10 / 0.1 / 0.5) on hand-computed inputs: MSE over
    the valid rays only, eikonal mean((|g| - 1)^2), curvature mean|laplace|."""
    from oracle import render as orender
    out = {"rays_valid": torch.tensor([[True], [False], [True]]),
           "comp_rgb_full": torch.tensor([[0.5, 0.5, 0.5], [9.0, 9.0, 9.0], [1.0, 0.0, 0.0]]),
           "grad": torch.tensor([[0.0, 3.0, 0.0], [0.6, 0.8, 0.0]]),
           "laplace": torch.tensor([[0.1], [-0.3]])}
    total, t = orender.losses(out, torch.zeros(3, 3))
    mse = (3 * 0.25 + 1.0) / 6
    assert abs(t["rgb_mse"].item() - mse) < 1e-7
    assert abs(t["eikonal"].item() - 2.0) < 1e-6          # ((3-1)^2 + 0) / 2
    assert abs(t["curvature"].item() - 0.2) < 1e-7
    assert abs(total.item() - (10 * mse + 0.1 * 2.0 + 0.5 * 0.2)) < 1e-5


REF_CONFIG = "/root/reference/configs/neuralangelo-colmap_sparse.yaml"


def _subset_equal(mine, ref, path=""):
    """Every key of `mine` is in `ref` with the same value (lists / tuples
    compared as lists, numbers as floats)."""
    if isinstance(mine, dict):
        for k, v in mine.items():
            assert k in ref, f"{path}.{k} missing from the reference config"
            _subset_equal(v, ref[k], f"{path}.{k}")
        return
    if isinstance(mine, (list, tuple)):
        assert list(mine) == list(ref), (path, mine, ref)
        return
    if isinstance(mine, (int, float)) and not isinstance(mine, bool):
        assert abs(float(mine) - float(ref)) <= 1e-12 * max(1.0, abs(float(ref))), (path, mine, ref)
        return
    assert mine == ref, (path, mine, ref)


def test_neuralangelo_config_restates_the_reference_yaml():
    """The restated Neuralangelo model and loss configuration equals the
    reference's configs/neuralangelo-colmap_sparse.yaml as load_config
    resolves it (interpolations ${model.radius} and ${add:65, 6}); skipped
    where the reference checkout is absent."""
    import os
    import pytest
    from angelo_amd.utils import load_config
    from angelo_amd.workloads import neuralangelo_configs
    if not os.path.exists(REF_CONFIG):
        pytest.skip("reference checkout not present")
    ref = load_config(REF_CONFIG, cli_args=["dataset.root_dir=/data/scene"])
    assert ref.name == "neuralangelo-colmap_sparse-scene"
    model, system = neuralangelo_configs()
    _subset_equal(model, ref.model)
    _subset_equal(system.loss, ref.system.loss)
    o = ref.system.optimizer
    assert system.optimizer.lr == o.args.lr == o.params.geometry.lr
    assert list(system.optimizer.betas) == list(o.args.betas) and system.optimizer.eps == o.args.eps
    assert system.optimizer.variance_lr == o.params.variance.lr


def test_load_config_interpolations(tmp_path):
    """utils/misc.py semantics: merge order, key.path=value overrides, key
    interpolation, nested resolver calls, typed single interpolations."""
    from angelo_amd.utils import load_config
    a = tmp_path / "a.yaml"
    b = tmp_path / "b.yaml"
    a.write_text("x: {r: 1.5, n: 3}\ny: ${add:${x.n},6}\nz: ${mul:${x.r},2}\n"
                 "w: s-${basename:${p}}\np: /a/b/c.txt\nq: ${calc_exp_lr_decay_rate:0.1,${sub:10,8}}\n"
                 "m: ???\n")
    b.write_text("x: {n: 65}\n")
    c = load_config(str(a), str(b), cli_args=["x.r=2.0"])
    assert c.y == 71 and isinstance(c.y, int)
    assert c.z == 4.0 and c.w == "s-c.txt" and abs(c.q - 0.1 ** 0.5) < 1e-15
    assert c.m == "???" and c.x.n == 65


def test_neuralangelo_model_builds_with_the_reference_shapes():
    """make('neus', <the Neuralangelo configuration>) builds the shapes the
    fused kernels take: SDF MLP 35 -> 64 -> 64 -> 65 (analytic), texture 87
    inputs with the base colour, the learned background."""
    import angelo_amd
    from angelo_amd.workloads import neuralangelo_configs
    torch.manual_seed(0)
    model = angelo_amd.make("neus", neuralangelo_configs()[0])
    geo, tex = model.geometry, model.texture
    assert [l.weight.shape[1] for l in geo.network.linears()] == [35, 64, 64]
    assert geo.network.linears()[-1].weight.shape[0] == 65 and geo.grad_type == "analytic"
    assert geo._fused_ok()
    assert tex.n_input_dims == 87 and tex.base_color == 1
    assert model.learned_background and model.geometry_bg.n_output_dims == 8
    assert geo.grid_module.log2_hashmap_size == 19


def test_serving_sample_buckets():
    """Render-stage graph buckets: base * 2^k * {1, 5/4, 3/2, 7/4}, the
    smallest holding the batch; padding stays under a quarter."""
    from angelo_amd.serving import sample_bucket
    base = 1 << 10
    assert sample_bucket(0, base) == base and sample_bucket(base, base) == base
    assert sample_bucket(base + 1, base) == base * 5 // 4
    prev = 0
    for n in range(1, 40 * base, 97):
        nb = sample_bucket(n, base)
        assert nb >= n and nb % 16 == 0
        assert n <= base or (nb - n) < nb / 4 + 1
        assert nb >= prev
        prev = nb


def test_set_flag_validates_names():
    """am_set_flag knows its variant switches and refuses other names on the
    host (no device needed)."""
    from angelo_amd import _native as nat
    lib = nat.load()
    assert lib.am_set_flag(b"no_such_flag", 1) == -1 and b"unknown flag" in lib.am_last_error()
    nat.set_flag("table_prepartition", 1)
    nat.set_flag("table_prepartition", 0)
    nat.set_flag("sdf_bwd_sync_free", 1)
    nat.set_flag("sdf_bwd_sync_free", 0)


def test_mixed_table_copy_follows_fused_optimizer_steps():
    """torch's fused AdamW updates parameters without bumping their version
    counter; the fp16 table copy the mixed-form kernels read must still
    follow the master: rebuilt on a grad-mode call, after update_step and on
    a train / eval switch (host-only: the copy is a torch conversion)."""
    from angelo_amd.network_utils import HashGrid
    g = HashGrid(3, {"n_levels": 2, "n_features_per_level": 2, "log2_hashmap_size": 8,
                     "base_resolution": 4, "per_level_scale": 1.5}, seed=3)
    g.precision = "mixed"
    opt = torch.optim.AdamW(g.parameters(), lr=0.1, fused=True)

    def fused_step():
        g.params.grad = torch.ones_like(g.params)
        v = g.params._version
        opt.step()
        assert g.params._version == v   # the premise: no version bump

    def matches():
        return torch.equal(g.kernel_table(), g.table.detach().to(torch.float16))

    with torch.no_grad():
        g.kernel_table()
    fused_step()
    assert matches()                     # grad mode on: rebuilt
    fused_step()
    g.update_step(0, 1)
    with torch.no_grad():
        assert matches()                 # update_step dropped the copy
    fused_step()
    g.eval()
    with torch.no_grad():
        assert matches()                 # mode switch dropped it
    g.load_state_dict({"params": torch.zeros_like(g.params)})
    with torch.no_grad():
        assert matches()                 # version bump of load_state_dict
