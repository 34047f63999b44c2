"""Pins of the network oracles: SH bases, VanillaMLP semantics."""
import math

import torch
import torch.nn as nn

from oracle import nets as onets


def _svox2_bases(d):
    """The reference's own SH closed form (models/texture.py:177-252,
    svox2_eval_sh_bases) for unit directions, degree 4 (16 bases)."""
    C0 = 0.28209479177387814
    C1 = 0.4886025119029199
    C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
          0.5462742152960396]
    C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
          -0.4570457994644658, 1.445305721320277, -0.5900435899266435]
    x, y, z = d.unbind(-1)
    xx, yy, zz = x * x, y * y, z * z
    xy, yz, xz = x * y, y * z, x * z
    return torch.stack([
        torch.full_like(x, C0), -C1 * y, C1 * z, -C1 * x,
        C2[0] * xy, C2[1] * yz, C2[2] * (2.0 * zz - xx - yy), C2[3] * xz, C2[4] * (xx - yy),
        C3[0] * y * (3 * xx - yy), C3[1] * xy * z, C3[2] * y * (4 * zz - xx - yy),
        C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy),
        C3[5] * z * (xx - yy), C3[6] * x * (xx - 3 * yy)], -1)


def test_sh_matches_reference_closed_form_on_unit_vectors():
    g = torch.Generator().manual_seed(0)
    d = torch.nn.functional.normalize(torch.randn(1000, 3, generator=g, dtype=torch.float64), dim=-1)
    ours = onets.sh_encode((d + 1) / 2, 4)
    torch.testing.assert_close(ours, _svox2_bases(d), rtol=0, atol=1e-12)


def test_sh_orthonormal_on_sphere():
    # Fibonacci sphere quadrature: int Y_i Y_j dOmega = delta_ij
    n = 20000
    k = torch.arange(n, dtype=torch.float64) + 0.5
    z = 1 - 2 * k / n
    phi = math.pi * (1 + 5 ** 0.5) * k
    r = (1 - z * z).sqrt()
    d = torch.stack([r * phi.cos(), r * phi.sin(), z], -1)
    Y = onets.sh_encode((d + 1) / 2, 4)
    G = Y.T @ Y * (4 * math.pi / n)
    torch.testing.assert_close(G, torch.eye(16, dtype=torch.float64), rtol=0, atol=2e-3)


def test_vanilla_mlp_matches_reference_modules():
    """VanillaMLP as the reference assembles it: nn.Linear + weight_norm,
    Softplus(beta=100) with sphere init; the oracle's functional form with
    W = g v / ||v|| gives the same outputs."""
    torch.manual_seed(0)
    dims = onets.mlp_dims(35, 16, 2)
    mods = []
    for i, (o, n) in enumerate(dims):
        lin = nn.utils.weight_norm(nn.Linear(n, o).double())
        mods.append(lin)
        if i < len(dims) - 1:
            mods.append(nn.Softplus(beta=100))
    seq = nn.Sequential(*mods)
    x = torch.randn(50, 35, dtype=torch.float64)
    ref = seq(x)
    layers = [(onets.weight_norm_weight(m.weight_g, m.weight_v), m.bias) for m in seq
              if isinstance(m, nn.Linear)]
    out = onets.vanilla_mlp(x, layers, "softplus", "none", 100.0)
    torch.testing.assert_close(out, ref, rtol=1e-12, atol=1e-12)


def test_sphere_init_structure():
    """Sphere init (models/network_utils.py:117-127): first layer zero beyond
    the xyz columns with std sqrt(2/out) on them, last layer mean
    sqrt(pi)/sqrt(in) and bias -r; averaged over many draws the SDF of the
    one-hidden-layer ReLU limit is |x| - r."""
    g = torch.Generator().manual_seed(1)
    x = torch.rand(64, 3, generator=g, dtype=torch.float64) * 2 - 1
    h = torch.cat([x, torch.zeros(64, 32, dtype=torch.float64)], -1)
    acc = torch.zeros(64, dtype=torch.float64)
    draws = 400
    for _ in range(draws):
        layers = onets.sphere_init_layers(35, 16, 1, 64, 0.5, gen=g, dtype=torch.float64)
        W0, b0 = layers[0]
        assert (W0[:, 3:] == 0).all() and (b0 == 0).all()
        assert (layers[-1][1] == -0.5).all()
        acc += onets.vanilla_mlp(h, layers, "relu")[:, 0]
    torch.testing.assert_close(acc / draws, x.norm(dim=-1) - 0.5, rtol=0, atol=0.08)


def test_flatten_roundtrip():
    dims = onets.mlp_dims(10, 3, 2)
    layers = [(torch.randn(o, i), torch.randn(o)) for o, i in dims]
    back = onets.unflatten(onets.flatten(layers), dims)
    for (a, b), (c, d) in zip(layers, back):
        assert torch.equal(a, c) and torch.equal(b, d)
