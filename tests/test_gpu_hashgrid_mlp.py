"""HIP hash-grid encode / deterministic backward and MFMA MLP vs the oracle."""
import numpy as np
import pytest
import torch

from oracle import hashgrid as ohg
from oracle import nets as onets

import helpers

pytestmark = pytest.mark.gpu

CONFIGS = {
    # name: (L, F, log2T, base, scale)
    "c0": (8, 2, 15, 16, 1.5),
    "c1_small": (16, 2, 16, 32, 1.3195079107728942),
    # full Neuralangelo table: 24.7 M rows, 25-bit sort keys
    "c1_full": (16, 2, 21, 32, 1.3195079107728942),
}


def _points(n, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(n, 3, generator=g)
    # edge cases: exact 0 and 1, cell boundaries of level 0
    extra = torch.tensor([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.5, 0.25, 1.0], [1.0, 0.0, 0.5]])
    return torch.cat([x, extra], 0)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_hashgrid_fwd_bitexact(dev, name):
    from angelo_amd import descs
    from angelo_amd.functions import hashgrid_encode
    L, F, log2T, base, s = CONFIGS[name]
    g = descs.grid_desc(L, F, log2T, base, s)
    lt = ohg.level_table(L, base, s, log2T, F)
    assert g.n_entries == lt.n_entries
    x = _points(3000, 1)
    gen = torch.Generator().manual_seed(2)
    table = (torch.rand(lt.n_entries, F, generator=gen) * 2 - 1)
    for active in (L, L // 2):
        out = hashgrid_encode(x.to(dev), table.to(dev), g, active).cpu().numpy()
        ref = ohg.encode_exact(x.numpy(), table.numpy(), lt, active)
        assert np.array_equal(out, ref), np.abs(out - ref).max()


@pytest.mark.parametrize("name,clustered", [(n, False) for n in CONFIGS] + [("c1_full", True)])
def test_hashgrid_bwd_matches_autograd_and_is_deterministic(dev, name, clustered):
    """clustered: 6000 points in a cube of side 0.004, so the coarse levels'
    few rows collect tens of thousands of gradient pairs (a reduction bucket
    split over several work units, combined in unit order)."""
    from angelo_amd import descs
    from angelo_amd.functions import hashgrid_encode
    L, F, log2T, base, s = CONFIGS[name]
    g = descs.grid_desc(L, F, log2T, base, s)
    lt = ohg.level_table(L, base, s, log2T, F)
    if clustered:
        x = 0.3 + 0.004 * torch.rand(6000, 3, generator=torch.Generator().manual_seed(9))
    else:
        x = _points(2000, 3)
    gen = torch.Generator().manual_seed(4)
    table = torch.rand(lt.n_entries, F, generator=gen) * 2 - 1
    dy = torch.randn(x.shape[0], L * F, generator=gen)
    active = L - 1
    # oracle gradients in float64
    x64 = x.double().requires_grad_(True)
    t64 = table.double().requires_grad_(True)
    ref = ohg.encode_torch(x64, t64, lt, active)
    (ref * dy.double()).sum().backward()

    res = []
    for _ in range(2):
        xd = x.to(dev).requires_grad_(True)
        td = table.to(dev).requires_grad_(True)
        out = hashgrid_encode(xd, td, g, active)
        (out * dy.to(dev)).sum().backward()
        res.append((td.grad.cpu(), xd.grad.cpu()))
    dtab, dx = res[0]
    assert torch.equal(res[0][0], res[1][0]), "table gradient not run-to-run identical"
    assert torch.equal(res[0][1], res[1][1])
    if clustered:
        # rows summing thousands of terms, and position gradients of fine
        # levels cancelling in fp32: the bar is torch's own fp32 autograd
        # error against the float64 oracle
        x32 = x.clone().requires_grad_(True)
        t32 = table.clone().requires_grad_(True)
        (ohg.encode_torch(x32, t32, lt, active) * dy).sum().backward()
        for got, ref32, ref64 in ((dtab, t32.grad, t64.grad), (dx, x32.grad, x64.grad)):
            e32 = float((ref32.double() - ref64).abs().max())
            e_hip = float((got.double() - ref64).abs().max())
            assert e_hip <= 8 * e32 + 1e-7, (e_hip, e32)
    else:
        torch.testing.assert_close(dtab.double(), t64.grad, rtol=1e-5, atol=1e-6)
        torch.testing.assert_close(dx.double(), x64.grad, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("case", [
    dict(n_in=35, n_out=16, nh=1, act="softplus", out="none"),
    dict(n_in=19, n_out=16, nh=2, act="softplus", out="none"),
    dict(n_in=38, n_out=3, nh=2, act="relu", out="sigmoid"),
    dict(n_in=64, n_out=40, nh=3, act="relu", out="none"),
])
def test_mlp_fwd_bwd(dev, case):
    from angelo_amd import descs
    from angelo_amd.functions import mlp_apply
    gen = torch.Generator().manual_seed(5)
    n = 1000 + 7
    dims = onets.mlp_dims(case["n_in"], case["n_out"], case["nh"])
    layers = [(torch.randn(o, i, generator=gen) / i ** 0.5, torch.randn(o, generator=gen) * 0.1)
              for o, i in dims]
    flat = onets.flatten(layers)
    x = torch.randn(n, case["n_in"], generator=gen)
    dy = torch.randn(n, case["n_out"], generator=gen)
    m = descs.mlp_desc(case["n_in"], case["n_out"], case["nh"], 64, case["act"], case["out"])
    assert descs.mlp_n_params(m) == flat.numel()

    x64 = x.double().requires_grad_(True)
    f64 = flat.double().requires_grad_(True)
    y64 = onets.vanilla_mlp(x64, onets.unflatten(f64, dims), case["act"], case["out"])
    (y64 * dy.double()).sum().backward()

    xd = x.to(dev).requires_grad_(True)
    fd = flat.to(dev).requires_grad_(True)
    y = mlp_apply(xd, fd, m)
    (y * dy.to(dev)).sum().backward()
    torch.testing.assert_close(y.detach().cpu().double(), y64.detach(), rtol=1e-4, atol=1e-5)
    torch.testing.assert_close(xd.grad.cpu().double(), x64.grad, rtol=1e-4, atol=1e-4)
    torch.testing.assert_close(fd.grad.cpu().double(), f64.grad, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("name", ["c0", "c1_full"])
def test_hashgrid_spatial_order(dev, monkeypatch, name):
    """The encode of a large batch runs in Morton order: outputs and dL/dx
    bit-identical with the caller's order, the table gradient equal up to
    the order of its fp32 sums, and repeatable bit for bit."""
    from angelo_amd import descs, functions
    from angelo_amd.functions import hashgrid_encode
    L, F, log2T, base, scale = CONFIGS[name]
    g = descs.grid_desc(L, F, log2T, base, scale)
    gen = torch.Generator().manual_seed(21)
    table = ((torch.rand(int(g.n_entries), F, generator=gen) * 2 - 1) * 0.1).to(dev)
    x = _points(100000, 22).to(dev)
    dy = torch.randn(x.shape[0], L * F, generator=gen).to(dev)

    def run(on):
        monkeypatch.setitem(functions.SPATIAL_ORDER, "encode", on)
        xd = x.clone().requires_grad_(True)
        td = table.clone().requires_grad_(True)
        out = hashgrid_encode(xd, td, g, L)
        (out * dy).sum().backward()
        return out.detach(), xd.grad, td.grad

    o0, dx0, dt0 = run(False)
    o1, dx1, dt1 = run(True)
    o2, dx2, dt2 = run(True)
    assert torch.equal(o0, o1) and torch.equal(dx0, dx1)
    assert torch.equal(dt1, dt2)
    helpers.assert_nontrivial(dt0, "table gradient")
    assert helpers.rel_err(dt1, dt0) < 1e-5
