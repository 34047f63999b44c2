"""Pins of the VolumeSDF and NeuS-rendering oracles."""
import math

import torch
import torch.nn.functional as F

from oracle import field as ofield
from oracle import render as orender

D = torch.float64


def test_contract_modes():
    x = torch.tensor([-1.5, 0.0, 0.7, 1.5], dtype=torch.float32)
    c32 = ofield.contract(x, 1.5, "f32")
    c64 = ofield.contract(x.double(), 1.5, "f64")
    torch.testing.assert_close(c32.double(), c64, rtol=0, atol=1e-7)
    assert c64[0] == 0 and c64[-1] == 1
    # fp32 semantics: multiply by the fp32 reciprocal of 3 (torch's CUDA scalar divide)
    inv = torch.tensor(1.0, dtype=torch.float32) / torch.tensor(3.0, dtype=torch.float32)
    assert torch.equal(c32, (x + 1.5) * inv)


def test_fd_gradient_of_linear_field_is_exact():
    a = torch.tensor([0.3, -1.2, 0.5], dtype=D)
    r = 1.0

    def field(x01):  # linear in the world position P = x01*2r - r
        P = x01 * 2 * r - r
        return torch.cat([(P @ a)[:, None], torch.zeros(x01.shape[0], 15, dtype=D)], -1)

    P = torch.rand(40, 3, dtype=D) * 1.6 - 0.8
    o = ofield.volume_sdf(P, field, r, 1e-3, True, False, mode="f64")
    torch.testing.assert_close(o["grad"], a.expand(40, 3), rtol=1e-8, atol=1e-8)
    torch.testing.assert_close(o["normal"], (a / a.norm()).expand(40, 3))
    torch.testing.assert_close(o["feature"][:, 16:], ofield.contract(P, r, "f64") * 2 - 1)


def _sphere_field(r_scene, R):
    def field(x01):
        P = x01 * 2 * r_scene - r_scene
        return torch.cat([(P.norm(dim=-1) - R)[:, None], torch.zeros(x01.shape[0], 15, dtype=D)], -1)
    return field


def test_curvature_world_shift_closed_form():
    """recontract off: the tangent shift moves along the sphere surface, the
    normal turns by eps/|P| -> laplace = eps/(pi |P|) to first order."""
    R = 0.5
    P = torch.tensor([[0.6, 0.0, 0.0], [0.0, 0.0, 0.45]], dtype=D)
    rnd = torch.tensor([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0]], dtype=D)
    eps = 1e-3
    o = ofield.volume_sdf(P, _sphere_field(1.0, R), 1.0, eps, True, True, rnd, recontract=False,
                          mode="f64")
    ref = eps / (math.pi * P.norm(dim=-1))
    torch.testing.assert_close(o["laplace"][:, 0], ref, rtol=2e-3, atol=0)


def test_curvature_recontract_quirk_hand_case():
    """recontract on (the reference): the shifted point is the contracted point
    plus the tangent; its FD points are contracted again.  For P = (0.6,0,0),
    r = 1: x01 = (0.8, 0.5, 0.5), normal x, randn y, tangent x cross y = +z,
    so the shifted point is (0.8, 0.5, 0.5+eps), taken as a world position
    and contracted again: ((0.8+1)/2, (0.5+1)/2, (0.5+eps+1)/2)."""
    eps = 1e-3
    P = torch.tensor([[0.6, 0.0, 0.0]], dtype=D)
    rnd = torch.tensor([[0.0, 1.0, 0.0]], dtype=D)
    seen = []

    def field(x01):
        seen.append(x01.clone())
        return _sphere_field(1.0, 0.5)(x01)

    ofield.volume_sdf(P, field, 1.0, eps, True, True, rnd, recontract=True, mode="f64")
    shifted = seen[2]  # base, FD, curvature FD
    centre = shifted.view(6, 3).mean(0)
    torch.testing.assert_close(centre, torch.tensor([0.9, 0.75, (1.5 + eps) / 2], dtype=D),
                               rtol=0, atol=1e-12)


def test_volume_sdf_gradcheck():
    g = torch.Generator().manual_seed(0)
    W1 = torch.randn(8, 3, generator=g, dtype=D)
    W2 = torch.randn(16, 8, generator=g, dtype=D) * 0.3
    P = (torch.rand(3, 3, generator=g, dtype=D) * 1.4 - 0.7)
    rnd = torch.randn(3, 3, generator=g, dtype=D)

    def fn(w1, w2):
        field = lambda x: torch.tanh(x @ w1.T) @ w2.T
        o = ofield.volume_sdf(P, field, 1.0, 1e-2, True, True, rnd, mode="f64")
        return o["sdf"].sum() + o["grad"].sum() + o["laplace"].sum() + o["feature"].sum()

    assert torch.autograd.gradcheck(fn, (W1.requires_grad_(True), W2.requires_grad_(True)),
                                    eps=1e-6, atol=1e-4)


def test_analytic_gradient_closed_forms():
    """grad_type 'analytic': a field linear in the world position gives its
    coefficient exactly; for a quadratic field central differences are exact,
    so both modes agree to rounding."""
    a = torch.tensor([0.3, -1.2, 0.5], dtype=D)
    r = 1.0

    def linear(x01):
        P = x01 * 2 * r - r
        return torch.cat([(P @ a)[:, None], torch.zeros(x01.shape[0], 15, dtype=D)], -1)

    P = torch.rand(40, 3, dtype=D) * 1.6 - 0.8
    o = ofield.volume_sdf(P, linear, r, 1e-3, True, False, mode="f64", grad_type="analytic")
    torch.testing.assert_close(o["grad"], a.expand(40, 3), rtol=1e-12, atol=1e-12)
    c = torch.tensor([0.7, -0.4, 1.1], dtype=D)

    def quadratic(x01):
        P = x01 * 2 * r - r
        return torch.cat([((P * P) @ c + P @ a)[:, None],
                          torch.zeros(x01.shape[0], 15, dtype=D)], -1)

    oa = ofield.volume_sdf(P, quadratic, r, 1e-2, True, False, mode="f64", grad_type="analytic")
    of = ofield.volume_sdf(P, quadratic, r, 1e-2, True, False, mode="f64")
    torch.testing.assert_close(oa["grad"], 2 * P * c + a, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(oa["grad"], of["grad"], rtol=1e-9, atol=1e-9)


def test_analytic_volume_sdf_second_order_gradcheck():
    """Losses on the analytic grad (eikonal, normals, curvature tangent)
    differentiate to second order: gradcheck of the parameter gradients."""
    g = torch.Generator().manual_seed(1)
    W1 = torch.randn(8, 3, generator=g, dtype=D)
    W2 = torch.randn(16, 8, generator=g, dtype=D) * 0.3
    P = (torch.rand(3, 3, generator=g, dtype=D) * 1.4 - 0.7)
    rnd = torch.randn(3, 3, generator=g, dtype=D)

    def fn(w1, w2):
        field = lambda x: torch.tanh(x @ w1.T) @ w2.T
        o = ofield.volume_sdf(P, field, 1.0, 1e-2, True, True, rnd, mode="f64",
                              grad_type="analytic")
        eik = ((o["grad"].norm(dim=-1) - 1) ** 2).sum()
        return o["sdf"].sum() + eik + o["normal"].sum() + o["laplace"].sum()

    assert torch.autograd.gradcheck(fn, (W1.requires_grad_(True), W2.requires_grad_(True)),
                                    eps=1e-6, atol=1e-4)


def test_alpha_hand_case():
    """sdf 0 at a head-on crossing (cos = -1, ratio 1): prev = +d/2, next = -d/2."""
    s, d = 20.0, 0.1
    a = orender.get_alpha(torch.zeros(1, dtype=D), torch.tensor([[0.0, 0.0, 1.0]], dtype=D),
                          torch.tensor([[0.0, 0.0, -1.0]], dtype=D), torch.tensor([d], dtype=D),
                          torch.tensor([s], dtype=D), 1.0)
    sig = lambda v: 1 / (1 + math.exp(-v))
    ref = (sig(s * d / 2) - sig(-s * d / 2) + 1e-5) / (sig(s * d / 2) + 1e-5)
    assert abs(a.item() - ref) < 1e-15
    # leaving the surface (cos > 0, ratio 1): iter_cos = 0 -> alpha = 1e-5/(sig(0)+1e-5)
    a2 = orender.get_alpha(torch.zeros(1, dtype=D), torch.tensor([[0.0, 0.0, 1.0]], dtype=D),
                           torch.tensor([[0.0, 0.0, 1.0]], dtype=D), torch.tensor([d], dtype=D),
                           torch.tensor([s], dtype=D), 1.0)
    assert abs(a2.item() - 1e-5 / (0.5 + 1e-5)) < 1e-15


def test_transmittance_closed_forms():
    ridx = torch.tensor([0, 0, 0, 2, 3, 3], dtype=torch.int32)
    alpha = torch.tensor([0.5, 0.25, 0.1, 0.3, 1.0, 0.7], dtype=D)
    T = orender.transmittance(alpha, ridx, 4)
    torch.testing.assert_close(T, torch.tensor([1.0, 0.5, 0.375, 1.0, 1.0, 0.0], dtype=D))
    w = T * alpha
    op = orender.accumulate(w, ridx, 4)[:, 0]
    assert abs(op[0].item() - (1 - 0.5 * 0.75 * 0.9)) < 1e-15
    assert op[1] == 0 and abs(op[2] - 0.3) < 1e-15 and op[3] == 1.0


def test_accumulate_matches_loops():
    g = torch.Generator().manual_seed(2)
    ridx = torch.sort(torch.randint(0, 10, (100,), generator=g)).values
    w = torch.rand(100, generator=g, dtype=D)
    v = torch.rand(100, 3, generator=g, dtype=D)
    acc = orender.accumulate(w, ridx, 10, v)
    ref = torch.zeros(10, 3, dtype=D)
    for i in range(100):
        ref[ridx[i]] += w[i] * v[i]
    torch.testing.assert_close(acc, ref)


def test_composite_gradcheck():
    g = torch.Generator().manual_seed(3)
    ridx = torch.tensor([0, 0, 0, 1, 1, 3], dtype=torch.int32)
    n = 6
    sdf = (torch.randn(n, generator=g, dtype=D) * 0.1).requires_grad_(True)
    nrm = torch.nn.functional.normalize(torch.randn(n, 3, generator=g, dtype=D), dim=-1).requires_grad_(True)
    rgb = torch.rand(n, 3, generator=g, dtype=D).requires_grad_(True)
    inv_s = torch.tensor([15.0], dtype=D).requires_grad_(True)
    dirs = torch.nn.functional.normalize(torch.randn(n, 3, generator=g, dtype=D), dim=-1)
    t0 = torch.rand(n, generator=g, dtype=D)
    t1 = t0 + 0.05

    def fn(s, nn_, c, i):
        o = orender.composite(s, nn_, c, dirs, t0, t1, ridx, 4, i, 0.4)
        return o["opacity"].sum() + (o["comp_rgb"] * 1.3).sum() + o["depth"].sum()

    assert torch.autograd.gradcheck(fn, (sdf, nrm, rgb, inv_s), eps=1e-6, atol=1e-5)


def test_density_weights_closed_forms():
    n_rays, per = 5, 7
    ridx = torch.arange(n_rays).repeat_interleave(per)
    dt = 0.05
    t0 = torch.arange(per, dtype=torch.float64).repeat(n_rays) * dt
    t1 = t0 + dt
    sig = torch.linspace(0.5, 20.0, n_rays, dtype=torch.float64).repeat_interleave(per)
    w, T = orender.weights_from_density(sig, t0, t1, ridx, n_rays)
    j = torch.arange(per, dtype=torch.float64).repeat(n_rays)
    assert torch.allclose(T, torch.exp(-sig * j * dt), rtol=1e-12)
    tot = orender.accumulate(w, ridx, n_rays)[:, 0]
    s_ray = sig.view(n_rays, per)[:, 0] * dt * per
    assert torch.allclose(tot, 1 - torch.exp(-s_ray), rtol=1e-12)
    # sequential product form (nerfacc's per-ray loop), ragged rays
    ridx2 = torch.tensor([0, 0, 0, 2, 2, 3])
    a = torch.tensor([0.0, 0.1, 0.3, 0.0, 0.2, 1.0], dtype=torch.float64)
    b = a + torch.tensor([0.1, 0.2, 0.1, 0.2, 0.5, 0.3], dtype=torch.float64)
    sg = torch.tensor([1.0, 3.0, 0.5, 7.0, 2.0, 4.0], dtype=torch.float64)
    w2, _ = orender.weights_from_density(sg, a, b, ridx2, 4)
    ref = []
    for r in range(4):
        T_ = 1.0
        for k in (ridx2 == r).nonzero()[:, 0].tolist():
            alpha = 1 - math.exp(-float(sg[k]) * float(b[k] - a[k]))
            ref.append(T_ * alpha)
            T_ *= 1 - alpha
    assert torch.allclose(w2, torch.tensor(ref, dtype=torch.float64), rtol=1e-12)


def test_contract_unbounded_hand_cases_and_gradcheck():
    x = torch.tensor([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [3.0, 0.0, 0.0], [0.0, -100.0, 0.0]],
                     dtype=torch.float64)
    y = ofield.contract_unbounded(x, 1.0, mode="f64")
    assert torch.allclose(y[0], torch.full((3,), 0.5, dtype=torch.float64))
    assert torch.allclose(y[1], torch.tensor([0.625, 0.5, 0.5], dtype=torch.float64))
    # |x| = 3 > 1: (2 - 1/3) / 4 + 1/2
    assert torch.allclose(y[2, 0], torch.tensor((2 - 1 / 3) / 4 + 0.5, dtype=torch.float64))
    # x = (0, -100, 0): (2 - 1/100) * (-1) / 4 + 1/2
    assert torch.allclose(y[3], torch.tensor([0.5, 0.0025, 0.5], dtype=torch.float64))


def test_forward_bg_gradcheck():
    torch.manual_seed(0)
    n_rays = 3
    rays = torch.cat([torch.rand(n_rays, 3) - 0.5, F.normalize(torch.randn(n_rays, 3), dim=-1)],
                     -1).double()
    ridx = torch.tensor([0, 0, 1, 1, 1, 2])
    t0 = torch.tensor([0.1, 0.3, 0.2, 0.5, 2.0, 0.4], dtype=torch.float64)
    t1 = t0 + torch.tensor([0.2, 0.3, 0.3, 1.5, 3.0, 0.1], dtype=torch.float64)
    W = torch.randn(4, 3, dtype=torch.float64, requires_grad=True)

    def radiance(feature, dirs):
        return torch.sigmoid(feature[:, 1:4] + dirs)

    def f(Wv):
        out = orender.forward_bg(rays, ridx, t0, t1, n_rays, lambda x: x @ Wv.t(), radiance, 1.0,
                                 mode="f64")
        return out["comp_rgb"], out["depth"], out["opacity"]
    assert torch.autograd.gradcheck(f, (W,), eps=1e-6, atol=1e-6)


def test_losses_closed_forms():
    """The oracle's loss glue (systems/neus.py:130-159) on a hand case whose
    terms have closed forms: the colour MSE over valid rays only (the invalid
    ray's colour would change it), the eikonal mean of (|g| - 1)^2 over
    samples and the mean |laplace|, weighted 10 / 0.1 / 0.5."""
    out = {"rays_valid": torch.tensor([[True], [False], [True]]),
           "comp_rgb_full": torch.tensor([[0.5] * 3, [9.0] * 3, [0.5] * 3], dtype=D),
           "grad": torch.tensor([[0.0, 2.0, 0.0], [3.0, 4.0, 0.0]], dtype=D),
           "laplace": torch.tensor([0.2, -0.4], dtype=D)}
    total, terms = orender.losses(out, torch.zeros(3, 3, dtype=D))
    assert math.isclose(terms["rgb_mse"].item(), 0.25, rel_tol=1e-15)
    assert math.isclose(terms["eikonal"].item(), (1.0 + 16.0) / 2, rel_tol=1e-15)
    assert math.isclose(terms["curvature"].item(), 0.3, rel_tol=1e-15)
    assert math.isclose(total.item(), 0.25 * 10 + 8.5 * 0.1 + 0.3 * 0.5, rel_tol=1e-15)
