"""Encoding / MLP plugin factories (models/network_utils.py:14-185).

get_encoding(n_input_dims, config) and get_mlp(n_input_dims, n_output_dims,
config) keep the reference's names, config keys and module attributes
(n_input_dims, n_output_dims, update_step, parameter names), so state dicts
and call sites carry over.  The tcnn-backed encodings are replaced by HIP
kernels (HashGrid / ProgressiveBandHashGrid / SphericalHarmonics);
VanillaMLP keeps its nn.Linear + weight-norm parameters and runs on the MFMA
MLP kernels.
"""
import math

import torch
import torch.nn as nn

from . import _native as nat
from . import descs
from . import functions as fn
from .utils import config_to_primitive, get_activation, update_module_step


def _cfg(config, key, default=None):
    if hasattr(config, "get"):
        return config.get(key, default)
    return getattr(config, key, default)


class VanillaFrequency(nn.Module):
    """models/network_utils.py:14-37 (positional encoding; plain torch, not on
    the NeuS hot path)."""

    def __init__(self, in_channels, config):
        super().__init__()
        self.N_freqs = config["n_frequencies"]
        self.in_channels, self.n_input_dims = in_channels, in_channels
        self.funcs = [torch.sin, torch.cos]
        self.freq_bands = 2 ** torch.linspace(0, self.N_freqs - 1, self.N_freqs)
        self.n_output_dims = self.in_channels * (len(self.funcs) * self.N_freqs)
        self.n_masking_step = config.get("n_masking_step", 0)
        self.update_step(None, None)

    def forward(self, x):
        out = []
        for freq, mask in zip(self.freq_bands, self.mask):
            for func in self.funcs:
                out += [func(freq * x) * mask]
        return torch.cat(out, -1)

    def update_step(self, epoch, global_step):
        if self.n_masking_step <= 0 or global_step is None:
            self.mask = torch.ones(self.N_freqs, dtype=torch.float32)
        else:
            self.mask = (1.0 - torch.cos(math.pi * (global_step / self.n_masking_step * self.N_freqs
                                                     - torch.arange(0, self.N_freqs)).clamp(0, 1))) / 2.0


class HashGrid(nn.Module):
    """tcnn HashGrid encoding (otype 'HashGrid'): x[B,3] in [0,1] -> [B, L*F].
    `params` is the flat fp32 table [n_entries*F] (tcnn's parameter name),
    initialised U(-1e-4, 1e-4) as tcnn does."""

    def __init__(self, n_input_dims, config, seed=None):
        super().__init__()
        if n_input_dims != 3:
            raise NotImplementedError("HashGrid is implemented for 3-D inputs")
        self.n_input_dims = 3
        self.n_levels = int(config["n_levels"])
        self.n_features_per_level = int(config.get("n_features_per_level", 2))
        self.log2_hashmap_size = int(config.get("log2_hashmap_size", 19))
        self.base_resolution = int(config.get("base_resolution", 16))
        self.per_level_scale = float(config.get("per_level_scale", 2.0))
        res = config.get("level_resolutions", None)
        self.grid = descs.grid_desc(self.n_levels, self.n_features_per_level,
                                    self.log2_hashmap_size, self.base_resolution,
                                    self.per_level_scale, res)
        self.n_output_dims = self.n_levels * self.n_features_per_level
        n = int(self.grid.n_entries) * self.n_features_per_level
        gen = torch.Generator().manual_seed(int(seed)) if seed is not None else None
        init = (torch.rand(n, generator=gen) * 2.0 - 1.0) * 1e-4
        self.params = nn.Parameter(init)
        self.active_levels = self.n_levels
        # "mixed": the fused kernels read an fp16 copy of the fp32 master
        # table.  torch's fused AdamW updates parameters without bumping
        # their version counter, so the version alone cannot say when the
        # master changed: the copy is rebuilt on every call that can lead to
        # an optimizer step (grad mode on), at every update_step (the start
        # of a training step, before its no-grad occupancy update) and on
        # train / eval switches, and otherwise whenever the version or the
        # storage changes (load_state_dict, plain optimizers).
        self.precision = "fp32"
        self._half = None
        self._half_key = None

    @property
    def table(self):
        return self.params.view(-1, self.n_features_per_level)

    def kernel_table(self):
        """The table the fused kernels read: the fp32 master, or its fp16
        copy in the mixed-precision form."""
        if self.precision != "mixed":
            return self.table
        key = (self.params.data_ptr(), self.params._version)
        fresh = torch.is_grad_enabled() and self.params.requires_grad
        if self._half is None or self._half_key != key or fresh:
            with torch.no_grad():
                self._half = self.table.detach().to(torch.float16).contiguous()
            self._half_key = key
        return self._half

    def invalidate_half(self):
        """Drops the fp16 copy (the master may have changed unseen)."""
        self._half = None
        self._half_key = None

    def update_step(self, epoch, global_step):
        self.invalidate_half()

    def train(self, mode=True):
        self.invalidate_half()
        return super().train(mode)

    def level_table(self):
        return descs.grid_arrays(self.grid)

    def forward(self, x):
        shape = x.shape[:-1]
        out = fn.hashgrid_encode(x.reshape(-1, 3), self.table, self.grid, self.active_levels)
        return out.view(*shape, self.n_output_dims)


class ProgressiveBandHashGrid(nn.Module):
    """models/network_utils.py:40-66: HashGrid whose levels >= current_level
    are masked to zero; update_step advances current_level."""

    def __init__(self, in_channels, config, seed=None):
        super().__init__()
        self.n_input_dims = in_channels
        enc_cfg = dict(config)
        enc_cfg["otype"] = "HashGrid"
        self.encoding = HashGrid(in_channels, enc_cfg, seed=seed)
        self.n_output_dims = self.encoding.n_output_dims
        self.n_level = config["n_levels"]
        self.n_features_per_level = config["n_features_per_level"]
        self.start_level, self.start_step, self.update_steps = (
            config["start_level"], config["start_step"], config["update_steps"])
        self.current_level = self.start_level
        # the reference's mask only ever gains ones (:53-54,66)
        self._mask_levels = self.current_level
        self.encoding.active_levels = self._mask_levels

    @property
    def mask(self):
        m = torch.zeros(self.n_level * self.n_features_per_level)
        m[: self._mask_levels * self.n_features_per_level] = 1.0
        return m

    def forward(self, x):
        return self.encoding(x)

    def update_step(self, epoch, global_step):
        self.encoding.update_step(epoch, global_step)
        current_level = min(self.start_level + max(global_step - self.start_step, 0) // self.update_steps,
                            self.n_level)
        self.current_level = current_level
        self._mask_levels = max(self._mask_levels, current_level)
        self.encoding.active_levels = self._mask_levels


class SphericalHarmonics(nn.Module):
    """tcnn SphericalHarmonics (otype 'SphericalHarmonics'), input in [0,1]^3."""

    def __init__(self, n_input_dims, config):
        super().__init__()
        if n_input_dims != 3:
            raise NotImplementedError("SphericalHarmonics needs 3-D directions")
        self.degree = int(config["degree"])
        if not 1 <= self.degree <= 4:
            raise NotImplementedError("SphericalHarmonics degree must be in [1,4]")
        self.n_input_dims = 3
        self.n_output_dims = self.degree ** 2

    def forward(self, x):
        shape = x.shape[:-1]
        return fn.sh_encode(x.reshape(-1, 3), self.degree).view(*shape, self.n_output_dims)


class CompositeEncoding(nn.Module):
    """models/network_utils.py:69-80: optionally prepends x*xyz_scale+xyz_offset."""

    def __init__(self, encoding, include_xyz=False, xyz_scale=1.0, xyz_offset=0.0):
        super().__init__()
        self.encoding = encoding
        self.include_xyz, self.xyz_scale, self.xyz_offset = include_xyz, xyz_scale, xyz_offset
        self.n_input_dims = encoding.n_input_dims
        self.n_output_dims = int(self.include_xyz) * self.encoding.n_input_dims + self.encoding.n_output_dims

    def forward(self, x, *args):
        if not self.include_xyz:
            return self.encoding(x, *args)
        return torch.cat([x * self.xyz_scale + self.xyz_offset, self.encoding(x, *args)], dim=-1)

    def update_step(self, epoch, global_step):
        update_module_step(self.encoding, epoch, global_step)


def get_encoding(n_input_dims, config, seed=None):
    """models/network_utils.py:83-93."""
    otype = _cfg(config, "otype")
    prim = config_to_primitive(config)
    if otype == "VanillaFrequency":
        encoding = VanillaFrequency(n_input_dims, prim)
    elif otype == "ProgressiveBandHashGrid":
        encoding = ProgressiveBandHashGrid(n_input_dims, prim, seed=seed)
    elif otype == "HashGrid":
        encoding = HashGrid(n_input_dims, prim, seed=seed)
    elif otype == "SphericalHarmonics":
        encoding = SphericalHarmonics(n_input_dims, prim)
    else:
        raise NotImplementedError(f"encoding otype {otype!r}")
    return CompositeEncoding(encoding, include_xyz=_cfg(config, "include_xyz", False),
                             xyz_scale=2.0, xyz_offset=-1.0)


class VanillaMLP(nn.Module):
    """models/network_utils.py:96-140: nn.Linear layers (weight-norm and sphere
    init as configured), Softplus(beta=100) hidden activation with
    sphere_init, ReLU otherwise.  Runs on the MFMA MLP kernels; the
    weight-norm product W = g*v/||v|| is formed by torch each call (tiny)."""

    def __init__(self, dim_in, dim_out, config):
        super().__init__()
        self.n_neurons, self.n_hidden_layers = config["n_neurons"], config["n_hidden_layers"]
        self.sphere_init, self.weight_norm = config.get("sphere_init", False), config.get("weight_norm", False)
        self.sphere_init_radius = config.get("sphere_init_radius", 0.5)
        self.dim_in, self.dim_out = dim_in, dim_out
        layers = [self.make_linear(dim_in, self.n_neurons, is_first=True, is_last=False), self.make_activation()]
        for _ in range(self.n_hidden_layers - 1):
            layers += [self.make_linear(self.n_neurons, self.n_neurons, is_first=False, is_last=False),
                       self.make_activation()]
        layers += [self.make_linear(self.n_neurons, dim_out, is_first=False, is_last=True)]
        self.layers = nn.Sequential(*layers)
        self.output_activation_name = (config.get("output_activation", "none") or "none").lower()
        self.output_activation = get_activation(config.get("output_activation", None))
        hidden = "softplus" if self.sphere_init else "relu"
        kernel_out = self.output_activation_name if self.output_activation_name in ("none", "relu", "sigmoid") else "none"
        self._kernel_out_act = kernel_out
        self.desc = descs.mlp_desc(dim_in, dim_out, self.n_hidden_layers, self.n_neurons, hidden,
                                   kernel_out, 100.0)

    def set_precision(self, precision):
        """"fp32", or "mixed": the fused field kernels run this network's
        matmuls on bf16 MFMA with fp32 accumulation."""
        self.desc.precision = nat.PREC[precision]

    def make_linear(self, dim_in, dim_out, is_first, is_last):
        layer = nn.Linear(dim_in, dim_out, bias=True)
        if self.sphere_init:
            if is_last:
                torch.nn.init.constant_(layer.bias, -self.sphere_init_radius)
                torch.nn.init.normal_(layer.weight, mean=math.sqrt(math.pi) / math.sqrt(dim_in), std=0.0001)
            elif is_first:
                torch.nn.init.constant_(layer.bias, 0.0)
                torch.nn.init.constant_(layer.weight[:, 3:], 0.0)
                torch.nn.init.normal_(layer.weight[:, :3], 0.0, math.sqrt(2) / math.sqrt(dim_out))
            else:
                torch.nn.init.constant_(layer.bias, 0.0)
                torch.nn.init.normal_(layer.weight, 0.0, math.sqrt(2) / math.sqrt(dim_out))
        else:
            torch.nn.init.constant_(layer.bias, 0.0)
            torch.nn.init.kaiming_uniform_(layer.weight, nonlinearity="relu")
        if self.weight_norm:
            layer = nn.utils.weight_norm(layer)
        return layer

    def make_activation(self):
        return nn.Softplus(beta=100) if self.sphere_init else nn.ReLU(inplace=True)

    def linears(self):
        return [m for m in self.layers if isinstance(m, nn.Linear)]

    def flat_params(self):
        """[W0,b0,...] with the weight-norm product formed (autograd to g, v)."""
        parts = []
        for lin in self.linears():
            if hasattr(lin, "weight_g"):
                W = torch._weight_norm(lin.weight_v, lin.weight_g, 0)
            else:
                W = lin.weight
            parts += [W.reshape(-1), lin.bias.reshape(-1)]
        return torch.cat(parts)

    def forward(self, x):
        shape = x.shape[:-1]
        y = fn.mlp_apply(x.reshape(-1, self.dim_in).float(), self.flat_params(), self.desc)
        if self.output_activation_name not in ("none", "relu", "sigmoid"):
            y = self.output_activation(y)
        return y.view(*shape, self.dim_out)


def get_mlp(n_input_dims, n_output_dims, config):
    """models/network_utils.py:177-185 (otype 'VanillaMLP'; tcnn networks are
    not part of this build)."""
    otype = _cfg(config, "otype")
    if otype != "VanillaMLP":
        raise NotImplementedError(f"network otype {otype!r}: only VanillaMLP is provided")
    return VanillaMLP(n_input_dims, n_output_dims, config_to_primitive(config))
