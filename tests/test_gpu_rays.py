"""On-device training-ray sampling (systems/neus.py:35-118) vs the oracle:
rays, target colours and masks bit-exact for the same draws; per-view
directions; four-channel images; out-of-range draws clamped and counted;
the sampler's draw paths (per-ray views, one view per batch, foreground
ratio)."""
import numpy as np
import pytest
import torch

from oracle import rays as orays

import test_oracle_rays as tor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C,per_view,apply_mask", [(3, False, False), (4, True, True),
                                                    (3, False, True)])
def test_sample_rays_bit_exact(dev, C, per_view, apply_mask):
    from angelo_amd.sampling import TrainRaySampler
    images, masks, c2w, dirs, index, ys, xs = tor._scene(3, n=5, H=17, W=23, C=C)
    if per_view:
        rng = np.random.default_rng(4)
        dirs = (dirs[None] * rng.uniform(0.5, 1.5, size=(5, 1, 1, 3))).astype(np.float32)
    s = TrainRaySampler(torch.from_numpy(images), torch.from_numpy(c2w), torch.from_numpy(dirs),
                        torch.from_numpy(masks), device=dev, apply_mask=apply_mask)
    bg = torch.tensor([0.1, 0.6, 0.3])
    i32 = lambda a: torch.from_numpy(np.asarray(a, np.int32)).to(dev)
    rays, rgb, fg = s.gather(i32(index), i32(ys), i32(xs), bg)
    o_rays, o_rgb, o_fg = orays.sample_rays(images, c2w, dirs, index, ys, xs, masks, bg.numpy(),
                                            apply_mask)
    assert np.array_equal(rays.cpu().numpy(), o_rays)
    assert np.array_equal(rgb.cpu().numpy(), o_rgb)
    assert np.array_equal(fg.cpu().numpy(), o_fg)
    assert s.out_of_range() == 0


def test_out_of_range_draws_are_clamped_and_counted(dev):
    from angelo_amd.sampling import TrainRaySampler
    images, masks, c2w, dirs, index, ys, xs = tor._scene(5, n=3, H=6, W=9)
    s = TrainRaySampler(torch.from_numpy(images), torch.from_numpy(c2w), torch.from_numpy(dirs),
                        None, device=dev)
    index = np.array([0, 3, -1, 2, 1], np.int32)
    ys = np.array([0, 1, 2, 6, 5], np.int32)
    xs = np.array([8, 9, 0, 1, -4], np.int32)
    i32 = lambda a: torch.from_numpy(a).to(dev)
    rays, rgb, fg = s.gather(i32(index), i32(ys), i32(xs), torch.ones(3))
    assert s.out_of_range() == 4
    ci, cy, cx = np.clip(index, 0, 2), np.clip(ys, 0, 5), np.clip(xs, 0, 8)
    o_rays, o_rgb, o_fg = orays.sample_rays(images, c2w, dirs, ci, cy, cx)
    assert np.array_equal(rays.cpu().numpy(), o_rays)
    assert np.array_equal(rgb.cpu().numpy(), o_rgb)
    assert np.array_equal(fg.cpu().numpy(), o_fg)


def test_sampler_draw_paths(dev):
    from angelo_amd.sampling import TrainRaySampler
    images, masks, c2w, dirs, _, _, _ = tor._scene(6, n=4, H=8, W=10)
    fg_idx = torch.nonzero(torch.from_numpy(masks) > 0.5)
    bg_idx = torch.nonzero(torch.from_numpy(masks) <= 0.5)
    g = torch.Generator(device=dev).manual_seed(0)

    class M:
        background_color = None
    model = M()
    for kw, check in [
        (dict(), None),
        (dict(batch_image_sampling=False), "one_view"),
        (dict(sample_foreground_ratio=0.5, fg_indices=fg_idx, bg_indices=bg_idx,
              background_color="random", apply_mask=True), "fg_ratio"),
    ]:
        s = TrainRaySampler(torch.from_numpy(images), torch.from_numpy(c2w),
                            torch.from_numpy(dirs), torch.from_numpy(masks), device=dev, **kw)
        batch = s.sample(200, generator=g, model=model)
        index, y, x, bg = (t.cpu() for t in s.last_draw)
        o_rays, o_rgb, o_fg = orays.sample_rays(images, c2w, dirs, index.numpy(), y.numpy(),
                                                x.numpy(), masks, bg.numpy(), s.apply_mask)
        assert np.array_equal(batch["rays"].cpu().numpy(), o_rays)
        assert np.array_equal(batch["rgb"].cpu().numpy(), o_rgb)
        assert model.background_color is not None
        if check == "one_view":
            assert int(index.unique().numel()) == 1
        if check == "fg_ratio":
            assert float(batch["fg_mask"][:160].min()) == 1.0
            assert float(batch["fg_mask"][160:].max()) == 0.0
        assert s.out_of_range() == 0
