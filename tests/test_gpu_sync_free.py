"""A training step of the plain NeuS workload (8 levels x 2^15 rows, 1024
rays x 64 uniform samples inside the unit sphere) without a single host
read: with the worst-case workspace the SDF backward skips its pair-count
read (its sort covers the bound), so forward, losses, backward and AdamW
run with torch's sync debug mode set to "error" and the library records no
host wait.  Loss and every gradient equal the counted path's bit for bit."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _step(model, trainer, opt, batch, sync_free):
    from angelo_amd import _native as nat, functions as fn
    fn.set_sdf_backward(sync_free=sync_free)
    nat.timer_enable(True)
    opt.zero_grad(set_to_none=True)
    torch.cuda.synchronize()
    torch.cuda.set_sync_debug_mode("error" if sync_free else "default")
    try:
        out = model(batch["rays"], samples=batch["samples"], randn=batch["randn"])
        loss, _ = trainer.losses(out, batch)
        loss.backward()
        grads = [p.grad.clone() for p in model.parameters() if p.grad is not None]
    finally:
        torch.cuda.set_sync_debug_mode("default")
        fn.set_sdf_backward(sync_free=False)
    torch.cuda.synchronize()
    syncs = nat.timer_units("host_syncs")
    nat.timer_enable(False)
    nat.sdf_bwd_check()
    return loss.detach().clone(), grads, syncs


def test_training_step_without_host_reads(dev):
    import angelo_amd
    from angelo_amd import scenes
    from angelo_amd.system import NeuSTrainer
    from angelo_amd.workloads import make_optimizer, model_config, system_config
    import helpers
    print("sync-free training step: 1024 rays x 64 samples", flush=True)
    cfg = model_config(8, 15, 16, 1.5, rays=1024, samples_per_ray=64)
    torch.manual_seed(51)
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=51).to(dev)
    model.train()
    model.update_step(0, 30000)
    trainer = NeuSTrainer(model, system_config(), cfg)
    trainer.global_step = 30000
    opt = make_optimizer(model, system_config())
    rays, gt = scenes.config0_batch(n_rays=1024, n_views=8, seed=52)
    rays, gt = rays.to(dev), gt.to(dev)
    samples = scenes.uniform_samples(rays, 64)
    g = torch.Generator(device=dev).manual_seed(53)
    randn = torch.randn(samples[0].shape[0], 3, device=dev, generator=g)
    batch = dict(rays=rays, rgb=gt, samples=samples, randn=randn)
    loss_c, grads_c, syncs_c = _step(model, trainer, opt, batch, False)
    loss_f, grads_f, syncs_f = _step(model, trainer, opt, batch, True)
    assert syncs_c >= 1          # the counted path reads its pair count
    assert syncs_f == 0
    assert torch.equal(loss_c, loss_f)
    assert len(grads_c) == len(grads_f) > 0
    for a, b in zip(grads_c, grads_f):
        assert torch.equal(a, b)
    # and the optimizer step itself runs without a host read
    torch.cuda.set_sync_debug_mode("error")
    try:
        opt.step()
    finally:
        torch.cuda.set_sync_debug_mode("default")
    torch.cuda.synchronize()


@pytest.mark.parametrize("grad_type", ["finite_difference", "analytic"])
@pytest.mark.parametrize("laplace", [False, True])
@pytest.mark.parametrize("part", [0, 1])
def test_sync_free_field_backward_is_identical(dev, grad_type, laplace, part):
    """Field level: pair arrays sized for the worst-case pair count (sentinel
    tail, sort over the bound) and arrays of the exact count give
    bit-identical table and MLP gradients, in both pair layouts, and the
    emit pass stays within the arrays (the consistency flags stay clear)."""
    from angelo_amd import _native as nat, functions as fn
    from test_gpu_pipeline import _field_model
    import helpers
    model = _field_model(dev, L=16, log2T=19, base=32, scale=1.3195079107728942,
                         grad_type=grad_type)
    geo = model.geometry
    g = torch.Generator().manual_seed(61)
    n = 20000
    pts = ((torch.rand(n, 3, generator=g) * 2 - 1) * 0.9).to(dev)
    randn = torch.randn(n, 3, generator=g).to(dev)
    grads = {}
    nat.set_flag("table_prepartition", part)
    try:
        for sync_free in (True, False):
            fn.set_sdf_backward(sync_free=sync_free)
            model.zero_grad()
            sdf, grad, normal, feature, lap = geo.fused(pts, True, laplace, randn)
            loss = sdf.sum() + grad.square().sum() + feature.sum()
            if laplace:
                loss = loss + lap.sum()
            loss.backward()
            grads[sync_free] = [p.grad.clone() for p in geo.parameters()]
    finally:
        fn.set_sdf_backward(sync_free=False)
        nat.set_flag("table_prepartition", 0)
    nat.sdf_bwd_check()
    helpers.assert_nontrivial(grads[True][0], "table gradient")
    for a, b in zip(grads[True], grads[False]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("grad_type", ["finite_difference", "analytic"])
@pytest.mark.parametrize("laplace", [False, True])
def test_atomic_table_gradient_matches_sorted(dev, grad_type, laplace):
    """The atomic table-gradient form (each merged pair added into its row
    with float atomics) sums the same terms as the sorted reduce in another
    order: equal within fp32 reassociation error, MLP gradients identical
    (they do not pass through the table reduce)."""
    from angelo_amd import _native as nat
    from test_gpu_pipeline import _field_model
    import helpers
    model = _field_model(dev, L=16, log2T=19, base=32, scale=1.3195079107728942,
                         grad_type=grad_type)
    geo = model.geometry
    g = torch.Generator().manual_seed(71)
    n = 20000
    pts = ((torch.rand(n, 3, generator=g) * 2 - 1) * 0.9).to(dev)
    randn = torch.randn(n, 3, generator=g).to(dev)
    grads = {}
    try:
        for atomic in (1, 0):
            nat.set_flag("table_atomic", atomic)
            model.zero_grad()
            sdf, grad, normal, feature, lap = geo.fused(pts, True, laplace, randn)
            loss = sdf.sum() + grad.square().sum() + feature.sum()
            if laplace:
                loss = loss + lap.sum()
            loss.backward()
            grads[atomic] = [p.grad.clone() for p in geo.parameters()]
    finally:
        nat.set_flag("table_atomic", 0)
    table_a, table_s = grads[1][0], grads[0][0]
    helpers.assert_nontrivial(table_s, "table gradient")
    # fp32 sums of up to a few hundred terms per row in another order
    scale = table_s.abs().max()
    torch.testing.assert_close(table_a, table_s, rtol=1e-4, atol=1e-6 * float(scale))
    for a, b in zip(grads[1][1:], grads[0][1:]):
        assert torch.equal(a, b)
