"""Serving loop: continuous batching of renders and SDF-lattice queries.

* Every request's result is bit-identical alone and in a mix (per-ray and
  per-point computations do not depend on the rest of the batch).
* Graph replay gives the same results as eager execution, bit for bit.
* Admission by footprint never holds more than the HBM budget (beyond a
  single request that exceeds it on its own) and serves every request.
* A served render equals the eval-mode NeuSModel.forward of the same rays
  within the fp32 tolerance, and the served SDF lattice equals
  VolumeSDF.sdf_grid.
"""
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu


def _model(dev, grid_prune=False):
    import angelo_amd
    cfg = helpers.model_config(L=8, log2T=15, base=16, scale=1.5, grid_prune=grid_prune,
                               step=2.0 / 256 if grid_prune else None)
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=81)
    model = model.to(dev)
    model.update_step(0, 30000)
    model.eval()
    return model


def _requests(seed=82):
    from angelo_amd import serving
    from angelo_amd.scenes import look_at_c2w
    g = torch.Generator().manual_seed(seed)
    reqs = []
    for i, size in enumerate((32, 48, 40, 24)):
        v = torch.randn(3, generator=g)
        reqs.append(serving.Request("render", 0.001 * i, c2w=look_at_c2w(v / v.norm() * 2.5),
                                    W=size, H=size, fov=40.0))
    for j, res in enumerate((24, 40)):
        reqs.append(serving.Request("sdf_grid", 0.0015 + 0.001 * j, resolution=res))
    return reqs


def _clone_reqs(reqs):
    from angelo_amd import serving
    return [serving.Request(r.kind, r.arrival, **r.params) for r in reqs]


def _serve(model, reqs, **kw):
    from angelo_amd import serving
    srv = serving.FieldServer(model, rays_per_step=1500, sdf_chunks_per_step=1,
                              sample_bucket=1 << 14, **kw)
    srv.run(reqs, realtime=False)
    assert all(r.finished is not None for r in reqs)
    return srv


@pytest.mark.parametrize("use_graphs", [False, True])
def test_results_identical_alone_and_in_mix(dev, use_graphs):
    model = _model(dev)
    mix = _clone_reqs(_requests())
    _serve(model, mix, use_graphs=use_graphs)
    for i, r in enumerate(_clone_reqs(_requests())):
        _serve(model, [r], use_graphs=use_graphs)
        m = mix[i]   # the same request (same kind and parameters) in the mix
        assert m.kind == r.kind and m.n_units() == r.n_units()
        for k in r.result:
            assert torch.equal(r.result[k], m.result[k]), (r.kind, k)


def test_graph_replay_matches_eager(dev):
    model = _model(dev, grid_prune=True)
    og = model.occupancy_grid
    from angelo_amd import scenes
    og.binary.copy_(scenes.prewarm_occupancy(scenes.sdf_union, og.resolution, 1.0, 1.5, dev))
    a = _clone_reqs(_requests(83))
    b = _clone_reqs(_requests(83))
    _serve(model, a, use_graphs=False)
    srv = _serve(model, b, use_graphs=True)
    assert srv._render_graphs or srv._sdf_graph is not None
    for x, y in zip(a, b):
        for k in x.result:
            assert torch.equal(x.result[k], y.result[k]), (x.kind, k)


def test_admission_by_footprint(dev):
    from angelo_amd import serving
    model = _model(dev)
    reqs = _clone_reqs(_requests(84))
    budget = max(r.footprint_bytes() for r in reqs) + 1
    srv = serving.FieldServer(model, rays_per_step=1000, sdf_chunks_per_step=1,
                              hbm_budget_bytes=budget, sample_bucket=1 << 14, use_graphs=False)
    peak = 0
    orig = srv._admit

    def admit(now):
        nonlocal peak
        orig(now)
        peak = max(peak, srv.in_flight_bytes)

    srv._admit = admit
    srv.run(reqs, realtime=False)
    assert all(r.finished is not None for r in reqs)
    assert peak <= budget
    assert srv.in_flight_bytes == 0
    # FIFO admission: requests were admitted in arrival order
    adm = [r.admitted for r in sorted(reqs, key=lambda r: r.arrival)]
    assert adm == sorted(adm)


def test_served_render_and_lattice_match_model(dev):
    from angelo_amd import serving
    model = _model(dev)
    reqs = _clone_reqs(_requests(85))
    _serve(model, reqs, use_graphs=False)
    r = next(q for q in reqs if q.kind == "render")
    srv = serving.FieldServer(model, rays_per_step=1500, sdf_chunks_per_step=1,
                              sample_bucket=1 << 14, use_graphs=False)
    srv._start(serving.Request("render", 0.0, **r.params), 0.0)
    rays = srv.active[0].rays
    with torch.no_grad():
        out = model(rays)
    torch.testing.assert_close(r.result["rgb"], out["comp_rgb_full"], rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(r.result["opacity"], out["opacity"][:, 0], rtol=1e-5, atol=1e-5)
    g = next(q for q in reqs if q.kind == "sdf_grid")
    ref = model.geometry.sdf_grid(g.params["resolution"])
    assert torch.equal(g.result["sdf"].view_as(ref), ref)


def test_served_render_with_learned_background(dev):
    """A model with the learned background: the served render equals the
    eval forward's comp_rgb_full (foreground over the background pass)."""
    import angelo_amd
    from angelo_amd import serving
    import test_gpu_background as tb
    cfg = tb._bg_model_config()
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=83).to(dev)
    model.update_step(0, 30000)
    model.eval()
    reqs = _clone_reqs(_requests(86))[:2]
    _serve(model, reqs, use_graphs=False)
    for r in reqs:
        srv = serving.FieldServer(model, rays_per_step=1500, sdf_chunks_per_step=1,
                                  sample_bucket=1 << 14, use_graphs=False)
        srv._start(serving.Request("render", 0.0, **r.params), 0.0)
        rays = srv.active[0].rays
        with torch.no_grad():
            out = model(rays)
        assert int(out["num_samples_bg"]) > 0
        torch.testing.assert_close(r.result["rgb"], out["comp_rgb_full"], rtol=1e-5, atol=1e-5)


def test_latency_stats_and_poisson_mix():
    from angelo_amd import serving
    reqs = serving.poisson_requests(40, 10.0, seed=1)
    assert [r.arrival for r in reqs] == sorted(r.arrival for r in reqs)
    kinds = {r.kind for r in reqs}
    assert kinds == {"render", "sdf_grid"}
    for i, r in enumerate(reqs):
        r.finished = r.arrival + 0.01 * (i + 1)
    st = serving.latency_stats(reqs)
    assert st["n"] == 40 and st["p50_s"] <= st["p99_s"]


def test_full_size_render_and_lattice_alone_and_in_mix_with_graphs(dev):
    """The serving workload's sizes on a field of the Neuralangelo training
    workload's shape (16 levels, 2^21 rows per level, base 32, occupancy
    grid 128^3 pre-warmed from the analytic scene, step 2/1024, at most 512
    samples per ray): one 1024^2 render and one 256^3 SDF lattice, served
    with graph replay alone and in a mix with other requests, give
    bit-identical results; both match the eager server."""
    import angelo_amd
    from angelo_amd import scenes, serving
    from angelo_amd.scenes import look_at_c2w
    from angelo_amd.workloads import NEURALANGELO_SCALE, model_config
    print("serving at full size: 1024^2 render + 256^3 lattice, graphs", flush=True)
    cfg = model_config(16, 21, 32, NEURALANGELO_SCALE, rays=8192, samples_per_ray=512,
                       step=2.0 / 1024, cap=512, grid_prune=True)
    torch.manual_seed(31)
    model = helpers.randomize_(angelo_amd.make("neus", cfg), seed=31).to(dev)
    model.eval()
    og = model.occupancy_grid
    og.binary.copy_(scenes.prewarm_occupancy(scenes.sdf_union, og.resolution, 1.0, 1.5, dev))
    eye = torch.tensor([1.3, -1.7, 1.2])
    big = [serving.Request("render", 0.0, c2w=look_at_c2w(eye / eye.norm() * 2.5), W=1024,
                           H=1024, fov=40.0),
           serving.Request("sdf_grid", 0.0, resolution=256)]
    others = _clone_reqs(_requests(87))

    def serve(reqs, use_graphs):
        srv = serving.FieldServer(model, rays_per_step=1 << 17, sdf_chunks_per_step=8,
                                  use_graphs=use_graphs)
        srv.run(reqs, realtime=False)
        assert all(r.finished is not None for r in reqs)
        return srv

    alone = _clone_reqs(big)
    srv = serve(alone, True)
    assert srv._render_graphs and srv._sdf_graph is not None
    mix = _clone_reqs(others[:3]) + _clone_reqs(big) + _clone_reqs(others[3:])
    serve(mix, True)
    eager = _clone_reqs(big)
    serve(eager, False)
    for a, m, e in zip(alone, mix[3:5], eager):
        assert a.kind == m.kind == e.kind
        for k in a.result:
            assert torch.equal(a.result[k], m.result[k]), (a.kind, k)
            assert torch.equal(a.result[k], e.result[k]), (a.kind, k)
    assert float(alone[0].result["opacity"].max()) > 0.5     # the render sees the scene
