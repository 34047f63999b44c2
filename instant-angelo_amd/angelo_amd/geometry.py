"""VolumeSDF (models/geometry.py:114-314), registered as 'volume-sdf'.

forward(points, with_grad, with_feature, with_laplace) keeps the reference's
signature and return convention; the finite-difference gradient, the
curvature ("laplace") evaluations and the encode + SDF MLP of all of them run
in one fused HIP launch (am_sdf_fwd / am_sdf_bwd).
"""
import math

import numpy as np
import torch
import torch.nn as nn

from . import _native as nat
from . import register
from . import functions as fn
from .network_utils import get_encoding, get_mlp, HashGrid
from .utils import get_activation, scale_anything, update_module_step


def contract_to_unisphere(x, radius, contraction_type="aabb"):
    """models/geometry.py:19-31: "aabb" maps [-r, r]^3 to [0, 1]^3;
    "un_bounded_sphere" maps the cube to [-1, 1]^3, points of norm m > 1 to
    (2 - 1/m) x/m, then [-2, 2]^3 to [0, 1]^3."""
    if contraction_type in ("aabb", None):
        return scale_anything(x, (-radius, radius), (0, 1))
    if contraction_type != "un_bounded_sphere":
        raise NotImplementedError(f"contraction {contraction_type!r}")
    x = scale_anything(x, (-radius, radius), (0, 1))
    x = x * 2 - 1
    mag = x.norm(dim=-1, keepdim=True)
    outside = mag > 1
    m = torch.where(outside, mag, torch.ones_like(mag))   # no 1/0 in the unused branch
    x = torch.where(outside, (2 - 1 / m) * (x / m), x)
    return x / 4 + 0.5


class BaseImplicitGeometry(nn.Module):
    """models/geometry.py:114-149 minus the mesh export (marching cubes stays
    on the host in the reference; SDF-grid queries are `sdf_grid`)."""

    def __init__(self, config):
        super().__init__()
        self.config = config
        self.radius = float(config.radius)
        self.contraction_type = "aabb"
        self.setup()
        if config.get("weights", None):
            self.load_state_dict(torch.load(config.weights, weights_only=True))

    def setup(self):
        raise NotImplementedError

    def regularizations(self, out):
        return {}


@register("volume-sdf")
class VolumeSDF(BaseImplicitGeometry):
    def setup(self):
        self.n_output_dims = self.config.feature_dim
        encoding = get_encoding(3, self.config.xyz_encoding_config, seed=self.config.get("seed", None))
        network = get_mlp(encoding.n_output_dims, self.n_output_dims, self.config.mlp_network_config)
        self.encoding, self.network = encoding, network
        self.grad_type = self.config.grad_type
        self.finite_difference_eps = self.config.get("finite_difference_eps", 1e-3)
        self._finite_difference_eps = None
        # 1: curvature points are re-contracted exactly as the reference does
        # (models/geometry.py:246,264-265); 0: shift in world space.
        self.curvature_recontract = int(self.config.get("curvature_recontract", 1))
        self.precision = "fp32"
        self.set_precision(self.config.get("precision", "fp32"))

    def set_precision(self, precision):
        """"fp32" or "mixed" (fp16 table rows, fp16 table-gradient pairs
        summed in fp32).  The SDF MLP stays fp32 in both: the central
        differences of the FD normal divide each evaluation's rounding by
        eps, so bf16 activations there would corrupt the normals."""
        if precision not in ("fp32", "mixed"):
            raise ValueError(f"precision {precision!r}")
        self.precision = precision
        self.network.set_precision("fp32")
        if isinstance(self.grid_module, HashGrid):
            self.grid_module.precision = precision

    # -- helpers ------------------------------------------------------------
    @property
    def grid_module(self):
        enc = self.encoding.encoding
        return enc.encoding if not isinstance(enc, HashGrid) else enc

    def active_levels(self):
        return self.grid_module.active_levels

    def _sdf_desc(self, with_grad, with_laplace):
        d = nat.SdfDesc()
        d.radius = self.radius
        eps = self._finite_difference_eps if self._finite_difference_eps is not None else 1e-3
        d.fd_eps = float(eps)
        d.with_grad = int(with_grad)
        d.with_laplace = int(with_laplace)
        d.recontract = self.curvature_recontract
        d.active_levels = int(self.active_levels())
        d.precision = nat.PREC[self.precision]
        d.grad_type = nat.GRAD_TYPE[self.grad_type]
        return d

    def _fused_ok(self):
        enc = self.encoding
        # up to 16 outputs, or up to 80 with two hidden layers (the 65 of
        # the Neuralangelo configurations)
        max_out = 80 if self.network.n_hidden_layers == 2 else 16
        return (self.grad_type in ("finite_difference", "analytic") and enc.include_xyz
                and isinstance(self.grid_module, HashGrid)
                and self.grid_module.n_features_per_level == 2
                and self.grid_module.n_levels <= 16
                and self.network.n_hidden_layers in (1, 2)
                and self.network._kernel_out_act == "none"
                and self.network.output_activation_name == "none"
                and self.n_output_dims <= max_out
                and "sdf_activation" not in self.config and "feature_activation" not in self.config)

    def fused(self, points, with_grad=True, with_laplace=False, randn=None, order=None):
        """All outputs of the fused launch: sdf [n], grad [n,3], normal [n,3],
        feature [n, feature_dim+3], laplace [n].  order: the samples'
        evaluation order (int32 permutation; None: the Morton order of a
        large batch, computed here, or the caller's order; False: the
        caller's order)."""
        if not self._fused_ok():
            raise NotImplementedError(
                "fused VolumeSDF needs grad_type finite_difference or analytic, include_xyz, a "
                "2-feature hash grid with <= 16 levels and a 1-2 hidden-layer VanillaMLP "
                "(<= 16 outputs, <= 80 with two hidden layers) without output/sdf/feature "
                "activations")
        pts = points.reshape(-1, 3)
        if with_laplace and randn is None:
            randn = torch.randn_like(pts)
        g = self.grid_module
        params = self.network.flat_params()
        # the backward's saved buffers (FD values, stored encodings) only when
        # a backward can follow: grad mode on and a parameter that takes one
        keep = torch.is_grad_enabled() and (g.table.requires_grad or params.requires_grad)
        args = (g.table, params, g.grid, self.network.desc,
                self._sdf_desc(with_grad, with_laplace), g.kernel_table())
        so = fn.SPATIAL_ORDER
        # the analytic-gradient form (7 evaluations, more arithmetic per gather)
        # measured no gain from the spatial order
        analytic = with_grad and self.grad_type == "analytic"
        if order is False:
            order = None
        elif order is None and so["on"] and pts.shape[0] >= so["min_samples"] and not analytic:
            pts = pts.float().contiguous()
            order, _ = fn.spatial_order(pts, self.radius)
        return fn.SDFFieldFn.apply(pts, randn, *args, order, keep)

    # -- reference API ------------------------------------------------------
    def forward(self, points, with_grad=True, with_feature=True, with_laplace=False,
                with_auxiliary_feature=False, randn=None):
        if with_auxiliary_feature:
            raise NotImplementedError("with_auxiliary_feature (sh-neus) is not part of this build")
        shape = points.shape[:-1]
        with torch.set_grad_enabled(self.training and torch.is_grad_enabled()):
            sdf, grad, _, feature, lap = self.fused(points, with_grad, with_laplace, randn)
        rv = [sdf.view(*shape)]
        if with_grad:
            rv.append(grad.view(*shape, 3))
        if with_feature:
            rv.append(feature.view(*shape, self.n_output_dims + 3))
        if with_laplace:
            rv.append(lap.view(*shape, 1))
        rv = [v if self.training else v.detach() for v in rv]
        return rv[0] if len(rv) == 1 else rv

    def forward_level(self, points):
        """models/geometry.py:289-294."""
        shape = points.shape[:-1]
        sdf = self.fused(points, with_grad=False)[0]
        return sdf.view(*shape)

    @torch.no_grad()
    def sdf_grid(self, resolution, bounds=None, chunk=64 ** 3):
        """SDF on the marching-cubes lattice of MarchingCubeHelper
        (models/geometry.py:36-113): torch.arange(min, max, 2/resolution) per
        axis, evaluated in `chunk`-point blocks; returns [X,Y,Z] float32."""
        r = self.radius
        bounds = bounds if bounds is not None else [[-r, r]] * 3
        intv = 2.0 / resolution
        axes = [torch.arange(lo, hi, intv) for lo, hi in bounds]
        dev = self.grid_module.params.device
        X, Y, Z = (len(a) for a in axes)
        out = torch.empty(X * Y * Z, device=dev)
        xs, ys, zs = (a.to(dev, torch.float32) for a in axes)
        for s in range(0, X * Y * Z, chunk):
            idx = torch.arange(s, min(s + chunk, X * Y * Z), device=dev)
            p = torch.stack([xs[idx // (Y * Z)], ys[(idx // Z) % Y], zs[idx % Z]], -1)
            out[s:s + len(idx)] = self.forward_level(p)
        return out.view(X, Y, Z)

    def update_step(self, epoch, global_step):
        """models/geometry.py:296-314."""
        update_module_step(self.encoding, epoch, global_step)
        update_module_step(self.network, epoch, global_step)
        if isinstance(self.finite_difference_eps, float):
            self._finite_difference_eps = self.finite_difference_eps
        elif self.finite_difference_eps == "progressive":
            hg = self.config.xyz_encoding_config
            assert hg.otype == "ProgressiveBandHashGrid"
            current_level = min(hg.start_level + max(global_step - hg.start_step, 0) // hg.update_steps,
                                hg.n_levels)
            grid_res = hg.base_resolution * hg.per_level_scale ** (current_level - 1)
            self._finite_difference_eps = 2 * self.config.radius / grid_res
        else:
            raise ValueError(f"Unknown finite_difference_eps={self.finite_difference_eps}")

    def isosurface(self):
        raise NotImplementedError("triangulation (mcubes) is host-side and out of scope; "
                                  "use sdf_grid() for the SDF lattice")


class EncodingWithNetwork(nn.Module):
    """models/network_utils.py:188-198: network(encoding(x))."""

    def __init__(self, encoding, network):
        super().__init__()
        self.encoding, self.network = encoding, network

    def forward(self, x):
        return self.network(self.encoding(x))

    def update_step(self, epoch, global_step):
        update_module_step(self.encoding, epoch, global_step)
        update_module_step(self.network, epoch, global_step)


@register("volume-density")
class VolumeDensity(BaseImplicitGeometry):
    """models/geometry.py:153-177, the learned background's field: the
    unbounded-sphere contraction (elementwise torch), the hash-grid encode
    and the MLP on their HIP kernels, then density = act(out[..., 0] + bias)
    (trunc_exp) and feature = out."""

    def setup(self):
        self.n_input_dims = self.config.get("n_input_dims", 3)
        self.n_output_dims = self.config.feature_dim
        encoding = get_encoding(self.n_input_dims, self.config.xyz_encoding_config,
                                seed=self.config.get("seed", None))
        network = get_mlp(encoding.n_output_dims, self.n_output_dims, self.config.mlp_network_config)
        self.encoding_with_network = EncodingWithNetwork(encoding, network)
        self.contraction_type = "un_bounded_sphere"

    def _density(self, out):
        density = out[..., 0]
        if "density_activation" in self.config:
            density = get_activation(self.config.density_activation)(
                density + float(self.config.density_bias))
        return density

    def _contract(self, points):
        """contract_to_unisphere (models/geometry.py:19-31); the unbounded
        sphere of the background runs as one HIP launch."""
        if self.contraction_type == "un_bounded_sphere":
            return fn.contract_unbounded(points, self.radius)
        return contract_to_unisphere(points, self.radius, self.contraction_type)

    def forward(self, points):
        points = self._contract(points)
        out = self.encoding_with_network(points.view(-1, self.n_input_dims)).view(
            *points.shape[:-1], self.n_output_dims).float()
        density, feature = self._density(out), out
        if "feature_activation" in self.config:
            feature = get_activation(self.config.feature_activation)(feature)
        return density, feature

    def forward_level(self, points):
        points = self._contract(points)
        out = self.encoding_with_network(points.reshape(-1, self.n_input_dims)).reshape(
            *points.shape[:-1], self.n_output_dims)
        return -self._density(out)

    def update_step(self, epoch, global_step):
        update_module_step(self.encoding_with_network, epoch, global_step)
