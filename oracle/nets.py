"""Network oracles -- test infrastructure only.

vanilla_mlp   VanillaMLP (models/network_utils.py:96-140): nn.Linear layers,
              hidden activation Softplus(beta=100) when sphere_init else ReLU
              (:136-140), output activation (:107,112), weight-norm
              W = g * v / ||v|| per output row (nn.utils.weight_norm, dim 0,
              :132-133), sphere init (:117-127).  Computed with torch's own
              functional ops, i.e. the same library routines the reference
              calls; pinned against nn.Linear/nn.Softplus modules assembled the
              way make_linear does (tests/test_oracle_nets.py).
sh_encode     tcnn SphericalHarmonics, input in [0,1]^3 mapped to 2*d-1
              (models/texture.py:17,24-25): real SH up to degree 4 written
              from the closed forms; pinned against the svox2 constants the
              reference itself carries (models/texture.py:177-252) on unit
              vectors.
"""
import math

import torch
import torch.nn.functional as F


def act_fn(name, beta=100.0):
    name = (name or "none").lower()
    if name == "relu":
        return F.relu
    if name == "softplus":
        return lambda z: F.softplus(z, beta=beta)
    if name == "sigmoid":
        return torch.sigmoid
    if name == "none":
        return lambda z: z
    raise ValueError(name)


def vanilla_mlp(x, layers, hidden_act="relu", output_act="none", beta=100.0):
    """layers: [(W [out,in], b [out]), ...]."""
    h = x
    for i, (W, b) in enumerate(layers):
        h = F.linear(h, W, b)
        if i < len(layers) - 1:
            h = act_fn(hidden_act, beta)(h)
    return act_fn(output_act, beta)(h)


def weight_norm_weight(g, v):
    """nn.utils.weight_norm(dim=0): W = g * v / ||v||_2 over all dims but 0."""
    return g * v / v.norm(dim=1, keepdim=True)


def flatten(layers):
    return torch.cat([t.reshape(-1) for W, b in layers for t in (W, b)])


def unflatten(flat, dims):
    """dims: [(out, in), ...] -> [(W, b), ...] views of flat."""
    out, o = [], 0
    for n_out, n_in in dims:
        W = flat[o:o + n_out * n_in].view(n_out, n_in)
        o += n_out * n_in
        b = flat[o:o + n_out]
        o += n_out
        out.append((W, b))
    return out


def mlp_dims(n_in, n_out, n_hidden_layers, n_neurons=64):
    dims = [(n_neurons, n_in)]
    dims += [(n_neurons, n_neurons)] * (n_hidden_layers - 1)
    dims += [(n_out, n_neurons)]
    return dims


def sphere_init_layers(n_in, n_out, n_hidden_layers, n_neurons=64, radius=0.5, gen=None,
                       dtype=torch.float32):
    """Sphere initialisation of VanillaMLP (models/network_utils.py:117-127)."""
    layers = []
    dims = mlp_dims(n_in, n_out, n_hidden_layers, n_neurons)
    for i, (o, n) in enumerate(dims):
        W = torch.zeros(o, n, dtype=dtype)
        b = torch.zeros(o, dtype=dtype)
        if i == len(dims) - 1:
            W.normal_(math.sqrt(math.pi) / math.sqrt(n), 0.0001, generator=gen)
            b.fill_(-radius)
        elif i == 0:
            W[:, :3].normal_(0.0, math.sqrt(2) / math.sqrt(o), generator=gen)
        else:
            W.normal_(0.0, math.sqrt(2) / math.sqrt(o), generator=gen)
        layers.append((W, b))
    return layers


# ---- spherical harmonics -------------------------------------------------
_SQPI = math.sqrt(math.pi)
# Real SH normalisations, written from the closed forms sqrt(.)/(k*sqrt(pi)).
SH_K = {
    "c0": 1.0 / (2 * _SQPI),
    "c1": math.sqrt(3.0) / (2 * _SQPI),
    "c2a": math.sqrt(15.0) / (2 * _SQPI),
    "c2b": math.sqrt(5.0) / (4 * _SQPI),
    "c2c": math.sqrt(15.0) / (4 * _SQPI),
    "c3a": math.sqrt(35.0 / 2.0) / (4 * _SQPI),
    "c3b": math.sqrt(105.0) / (2 * _SQPI),
    "c3c": math.sqrt(21.0 / 2.0) / (4 * _SQPI),
    "c3d": math.sqrt(7.0) / (4 * _SQPI),
    "c3e": math.sqrt(105.0) / (4 * _SQPI),
}


def sh_encode(d01, degree=4):
    """Real SH bases of v = 2*d01 - 1 (band-limited to `degree` bands,
    degree^2 outputs), in the ordering and signs of tcnn / svox2."""
    K = SH_K
    x, y, z = (d01 * 2 - 1).unbind(-1)
    out = [torch.full_like(x, K["c0"])]
    if degree > 1:
        out += [-K["c1"] * y, K["c1"] * z, -K["c1"] * x]
    if degree > 2:
        xx, yy, zz = x * x, y * y, z * z
        out += [K["c2a"] * x * y, -K["c2a"] * y * z, K["c2b"] * (3 * zz - 1),
                -K["c2a"] * x * z, K["c2c"] * (xx - yy)]
    if degree > 3:
        xx, yy, zz = x * x, y * y, z * z
        out += [-K["c3a"] * y * (3 * xx - yy), K["c3b"] * x * y * z,
                -K["c3c"] * y * (5 * zz - 1), K["c3d"] * z * (5 * zz - 3),
                -K["c3c"] * x * (5 * zz - 1), K["c3e"] * z * (xx - yy),
                -K["c3a"] * x * (xx - 3 * yy)]
    if degree > 4:
        raise NotImplementedError("degree > 4")
    return torch.stack(out, dim=-1)
