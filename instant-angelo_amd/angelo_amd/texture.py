"""VolumeRadiance (models/texture.py:10-36), registered as 'volume-radiance',
and VolumeDualColor (:38-64), 'volume-dual-color': the same network plus a
base colour sigmoid(feature[..., 1:4]) when the colour activation is set.

The SH direction encoding, the radiance MLP and the colour activation run as
one fused HIP launch (am_radiance_fwd / am_radiance_bwd) when the extra
input is the normal, as NeuSModel passes it; other argument patterns use the
separate SH and MLP kernels.
"""
import torch
import torch.nn as nn

from . import register
from . import functions as fn
from .network_utils import get_encoding, get_mlp, SphericalHarmonics
from .utils import get_activation, update_module_step


@register("volume-radiance")
class VolumeRadiance(nn.Module):
    def __init__(self, config):
        super().__init__()
        self.config = config
        self.n_dir_dims = self.config.get("n_dir_dims", 3)
        self.n_output_dims = 3
        encoding = get_encoding(self.n_dir_dims, self.config.dir_encoding_config)
        self.n_input_dims = self.config.input_feature_dim + encoding.n_output_dims
        network = get_mlp(self.n_input_dims, self.n_output_dims, self.config.mlp_network_config)
        self.encoding = encoding
        self.network = network
        self.color_activation = self.config.get("color_activation", None)
        self.set_precision(self.config.get("precision", "fp32"))

    def set_precision(self, precision):
        """"fp32" or "mixed" (radiance MLP on bf16 MFMA, fp32 accumulation)."""
        if precision not in ("fp32", "mixed"):
            raise ValueError(f"precision {precision!r}")
        self.precision = precision
        self.network.set_precision(precision)

    def _fused_ok(self, features, args):
        enc = self.encoding
        return (isinstance(enc.encoding, SphericalHarmonics) and not enc.include_xyz
                and self.n_dir_dims == 3 and len(args) == 1 and args[0].shape[-1] == 3
                and (self.color_activation or "").lower() == "sigmoid"
                and self.network.output_activation_name == "none"
                and self.network.n_hidden_layers in (1, 2)
                and self.n_input_dims <= 96
                and features.shape[-1] + enc.n_output_dims + 3 == self.network.dim_in)

    base_color = 0

    def fused(self, features, normal, rays, ray_indices):
        """rgb of samples whose view directions are rays[ray_indices, 3:6]."""
        if (self.color_activation or "").lower() != "sigmoid":
            raise NotImplementedError("the fused radiance launch applies the sigmoid colour "
                                      f"activation (got {self.color_activation!r})")
        return fn.RadianceFn.apply(features.reshape(-1, features.shape[-1]),
                                   normal.reshape(-1, 3), self.network.flat_params(), rays,
                                   ray_indices, self.network.desc, self.encoding.encoding.degree,
                                   self.base_color)

    def forward(self, features, dirs, *args):
        """models/texture.py:23-30."""
        if self._fused_ok(features, args):
            n = dirs.reshape(-1, 3).shape[0]
            rays = torch.cat([torch.zeros_like(dirs.reshape(-1, 3)), dirs.reshape(-1, 3)], -1).float().contiguous()
            ridx = torch.arange(n, device=dirs.device, dtype=torch.int32)
            color = self.fused(features, args[0], rays, ridx)
            return color.view(*features.shape[:-1], self.n_output_dims)
        d = (dirs + 1.0) / 2.0
        dirs_embd = self.encoding(d.view(-1, self.n_dir_dims))
        net_in = torch.cat([features.view(-1, features.shape[-1]), dirs_embd]
                           + [a.view(-1, a.shape[-1]) for a in args], dim=-1)
        color = self.network(net_in).view(*features.shape[:-1], self.n_output_dims).float()
        if self.color_activation is not None:
            color = get_activation(self.color_activation)(color)
            if self.base_color:
                color = color + get_activation(self.color_activation)(features[..., 1:4])
        return color

    def update_step(self, epoch, global_step):
        update_module_step(self.encoding, epoch, global_step)

    def regularizations(self, out):
        return {}


@register("volume-dual-color")
class VolumeDualColor(VolumeRadiance):
    """models/texture.py:38-64: color = act(MLP(...)) + act(features[..., 1:4])
    with the colour activation (sigmoid); fused in the radiance launch."""
    base_color = 1
