"""Host-side pieces of the learned background (no GPU): the unbounded-sphere
contraction of VolumeDensity against the oracle's restatement of
models/geometry.py:22-28, its inverse used by the background occupancy-grid
update (a round trip), and the background settings of NeuSModel
(models/neus.py:53-59,63-68)."""
import math

import torch

from oracle import field as ofield


def test_contraction_matches_oracle():
    from angelo_amd.geometry import contract_to_unisphere
    g = torch.Generator().manual_seed(0)
    p = torch.randn(5000, 3, generator=g, dtype=torch.float64) * torch.logspace(
        -2, 3, 5000, dtype=torch.float64)[:, None]
    for r in (1.0, 0.5):
        a = contract_to_unisphere(p, r, "un_bounded_sphere")
        b = ofield.contract_unbounded(p, r, mode="f64")
        assert torch.allclose(a, b, rtol=0, atol=1e-15)
        assert float(a.min()) >= 0.0 and float(a.max()) <= 1.0


def test_contract_inverse_round_trip():
    from angelo_amd.neus import contract_inv_unbounded
    g = torch.Generator().manual_seed(1)
    x = torch.rand(20000, 3, generator=g, dtype=torch.float64)
    x = x[(x - 0.5).norm(dim=1) < 0.499]
    for roi in ([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0], [-0.5, -0.5, -0.5, 0.5, 0.5, 0.5]):
        w = contract_inv_unbounded(x, roi)
        y = ofield.contract_unbounded(w, roi[3], mode="f64")
        assert torch.allclose(y, x, rtol=0, atol=1e-12)


def test_background_model_settings():
    import angelo_amd
    import test_gpu_background as tb
    cfg = tb._bg_model_config(grid_prune=True)
    m = angelo_amd.make("neus", cfg)
    assert m.learned_background
    assert type(m.geometry_bg).__name__ == "VolumeDensity"
    assert type(m.texture_bg).__name__ == "VolumeDualColor"
    assert m.near_plane_bg == 0.1 and m.far_plane_bg == 1e3 and m.render_step_size_bg == 0.01
    # (1 + cone)^num_samples_per_ray_bg == far plane
    assert abs((1 + m.cone_angle_bg) ** cfg.num_samples_per_ray_bg - 1e3) < 1e-6
    assert m.occupancy_grid_bg.resolution == 256
    # 8 field outputs: density pre-activation, base colour (1:4), features
    assert m.geometry_bg.encoding_with_network.network.desc.n_out == 8
