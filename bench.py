#!/usr/bin/env python
"""Benchmark: NeuS / Neuralangelo training step on one MI355X.

A step is one pass of the hot path over one batch: occupancy update (every
16th step), ray-AABB + occupancy marching, fused VolumeSDF (base + 6
central-difference + 6 curvature evaluations per sample), fused radiance,
NeuS alpha / transmittance / accumulation, rgb-MSE + eikonal + curvature
losses, backward to every parameter, AdamW.  Inputs are device-resident
before timing; the timed region is bracketed by device synchronisation.

Prints one JSON line (metric, value, ms_per_step, roofline of the dominant
kernel from live HIP-event timers, and the CPU oracle baseline).

    python bench.py [--steps K] [--warmup W]
                    [--workload neuralangelo-fp32|neuralangelo-mixed|sphere-fp32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "instant-angelo_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
TIMED_KERNELS = ["geo_fwd", "geo_bwd", "pair_count", "pair_emit", "pair_atomic", "radix_sort",
                 "seg_reduce",
                 "rad_fwd", "rad_bwd", "composite_fwd", "composite_bwd", "march_count",
                 "march_fill", "occ_sdf", "sample_rays", "hg_fwd", "hg_emit", "march_bg_count",
                 "march_bg_fill", "march_bg_visible", "composite_bg_fwd", "composite_bg_bwd"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, spec, n_samples, n_rays, n_pairs, mixed=False):
    """Algorithmic HBM bytes of one launch (per-step averages).

    geo_fwd: per encode evaluation of L levels, F=2 features, b_t B table
      entries (4 fp32, 2 in the mixed form's fp16 rows): 12 B position +
      8*L*F*b_t B corner reads; 13 evaluations per sample (base + 6 FD + 6
      curvature) + outputs sdf, grad, normal, feature (19), laplace and the
      12 FD values (39 floats) + the stored encodings (13 x L*F floats) and
      the curvature evaluations' position derivatives (6 x 3*L*F floats).
    geo_bwd (store form): reads the stored encodings and derivatives, the 12
      FD values, position and curvature direction (24 B) and the upstream
      gradients (27 floats); writes the encode gradients (13 x L*F floats).
    pair_count: position, direction, FD values.  pair_emit: the same + the
      encode gradients + 12 B (8 B mixed) per (row, value) pair written.
    radix_sort: each pair (4 B row + 8 B value; 4 B value in the mixed form)
      read and written once.
    seg_reduce: each pair read once + one 8 B row written per pair (bound).
    composite: 40 B per sample + 32 B per ray.
    Analytic-gradient mode: 7 evaluations (base with position derivatives +
    6 curvature), 7 derivative slots stored, 3 more encode-gradient slots.
    A workload held at its first band (start_level) gathers that many levels.
    """
    L, F = spec["levels"], 2
    # per-sample outputs: sdf, grad, normal, feature (feature_dim + 3), laplace
    # and the 12 FD / curvature values
    n_outf = spec.get("feature_dim", 16) + 3 + 20
    La = spec.get("start_level") or L   # levels gathered (the band's active levels)
    bt = 2 if mixed else 4
    pair = 8 if mixed else 12
    n_ev, n_dx, n_genc = evals_per_sample(spec), 6, evals_per_sample(spec)
    if spec.get("grad_type") == "analytic":
        n_dx, n_genc = 7, n_ev + 3
    per_eval = 12 + 8 * La * F * bt
    store = (n_ev * L * F + n_dx * 3 * L * F) * 4
    genc = n_genc * L * F * 4
    if kernel == "geo_fwd":
        return n_samples * (n_ev * per_eval + n_outf * 4 + store)
    if kernel == "geo_bwd":
        return n_samples * (store + 12 * 4 + 24 + (n_outf - 12) * 4 + genc)
    if kernel == "pair_count":
        return n_samples * (24 + 12 * 4)
    if kernel == "pair_emit":
        return n_samples * (24 + 12 * 4 + genc) + n_pairs * pair
    if kernel == "radix_sort":
        return n_pairs * pair * 2
    if kernel == "seg_reduce":
        return n_pairs * (pair + 8)
    if kernel in ("composite_fwd", "composite_bwd"):
        return n_samples * 40 + n_rays * 32
    return None


def evals_per_sample(spec):
    """Encode + SDF-MLP evaluations per sample: base + 6 central
    differences + 6 curvature, or base + 6 curvature in analytic mode."""
    return 7 if spec.get("grad_type") == "analytic" else 13


def profiled_traffic(workload):
    """HBM bytes per step of each timed kernel, from the committed summary of
    two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this same bench
    command, each scaled by the counter's reading for a 1 GiB copy.  None when
    no summary exists for the workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("<!-- traffic-json: "):
                return json.loads(line[len("<!-- traffic-json: "):].rsplit("-->", 1)[0])
    return None


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec_name, workload, train_rays=64, render_rays=2048, warmup=3, steps=10):
    """The oracle (torch CPU, fp32, CUDA-semantics restatement) on the host
    cores, by the stated protocol on a bounded sample of the same workload:
    training (NeuSModel.forward_ + losses + backward on train_rays rays of a
    step's batch) with `warmup` untimed and the median of `steps` timed
    steps, samples/s; rendering (the eval forward, chunk_batch semantics: one
    chunk of render_rays rays of the same batch, no gradients), rays/s.
    samples/s and rays/s are per sample and per ray, so the bounded sample
    scales linearly to the workload's batch."""
    import numpy as np
    import torch
    from oracle import march as omarch, render as orender
    from oracle import field as ofield, hashgrid as ohg, nets as onets
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    model = workload.model
    spec = workload.spec
    if spec.get("learned_background"):
        return {"skipped": "the CPU baseline covers the foreground step (no background oracle "
                           "pass in this measurement)"}
    grad_type = spec.get("grad_type", "finite_difference")
    r = float(model.geometry.radius)
    n_all = max(train_rays, render_rays)
    b = workload.batches[0] if workload.uniform else \
        workload.sampler.sample(spec["rays"], torch.Generator(device=workload.device).manual_seed(1))
    rays_all = b["rays"][:n_all].detach().cpu()
    gt_all = b["rgb"][:n_all].detach().cpu()
    box = [-r, -r, -r, r, r, r]

    def samples_of(rays):
        if workload.uniform:
            from angelo_amd import scenes
            ridx, t0, t1, _ = scenes.uniform_samples(rays, spec["samples_per_ray"])
            return ridx, t0, t1
        og = model.occupancy_grid
        tn, tf = omarch.ray_aabb(rays.numpy(), box)
        ri, a0, a1, _ = omarch.ray_marching(rays.numpy(), tn, tf, box, og.binary.cpu().numpy(),
                                            model.render_step_size, model.max_samples_per_ray)
        return torch.from_numpy(ri), torch.from_numpy(a0), torch.from_numpy(a1)

    geo, tex = model.geometry, model.texture
    g = geo.grid_module
    lt = ohg.level_table(g.n_levels, g.base_resolution, g.per_level_scale, g.log2_hashmap_size)
    table = g.table.detach().cpu().clone().requires_grad_(True)
    gflat = geo.network.flat_params().detach().cpu().clone().requires_grad_(True)
    rflat = tex.network.flat_params().detach().cpu().clone().requires_grad_(True)
    gd = onets.mlp_dims(geo.network.dim_in, geo.network.dim_out, geo.network.n_hidden_layers)
    rd = onets.mlp_dims(tex.network.dim_in, tex.network.dim_out, tex.network.n_hidden_layers)
    field = ofield.field_from(lambda x: ohg.encode_torch(x, table, lt, g.active_levels, "f32"),
                              lambda h: onets.vanilla_mlp(h, onets.unflatten(gflat, gd), "softplus"))
    base = getattr(tex, "base_color", 0)

    def radiance(feat, dirs, nrm):
        sh = onets.sh_encode((dirs + 1.0) / 2.0, 4)
        c = torch.sigmoid(onets.vanilla_mlp(torch.cat([feat, sh, nrm], -1),
                                            onets.unflatten(rflat, rd)))
        return c + torch.sigmoid(feat[:, 1:4]) if base else c

    var = model.variance.variance.detach().cpu().clone().requires_grad_(True)
    eps = float(geo._finite_difference_eps)

    def forward(rays, ridx, t0, t1, randn):
        inv_s = torch.exp(var * 10.0).clip(1e-6, 1e6)
        return orender.neus_forward(rays, ridx, t0, t1, rays.shape[0], field, radiance, inv_s, r,
                                    eps, model.cos_anneal_ratio, randn, grad_type=grad_type)

    # training
    rays, gt = rays_all[:train_rays], gt_all[:train_rays]
    ridx, t0, t1 = samples_of(rays)
    n = int(ridx.shape[0])
    randn = torch.randn(n, 3)
    times = []
    for i in range(warmup + steps):
        t = time.perf_counter()
        loss, _ = orender.losses(forward(rays, ridx, t0, t1, randn), gt)
        loss.backward()
        if i >= warmup:
            times.append(time.perf_counter() - t)
    med = sorted(times)[len(times) // 2]
    # rendering: one chunk, no gradients
    rays = rays_all[:render_rays]
    ridx_r, t0_r, t1_r = samples_of(rays)
    n_r = int(ridx_r.shape[0])
    with torch.no_grad():
        t = time.perf_counter()
        forward(rays, ridx_r, t0_r, t1_r, torch.randn(n_r, 3))
        t_render = time.perf_counter() - t
    return {"value": n / med, "unit": "samples/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model_name(),
            "sample": f"training: {train_rays} rays of a step's batch ({n} samples), oracle "
                      f"fwd + loss + bwd in torch CPU fp32, {warmup} warm-up steps then the "
                      f"median of {steps} ({med:.2f} s/step); samples/s is per sample, so it "
                      f"scales linearly to the workload's {spec['rays']} rays",
            "render_rays_per_s": round(rays.shape[0] / t_render, 1),
            "render_sample": f"eval forward of one chunk of {rays.shape[0]} rays "
                             f"({n_r} samples, chunk_batch chunk size 2048) in "
                             f"{t_render:.2f} s, no gradients"}


def serving_bench(wl, n_requests, load=0.6, seed=5):
    """Continuous-batching serving of a Poisson mix of renders (256^2, 512^2,
    1024^2 from seeded poses) and SDF lattices (128^3, 256^3, 512^3 in 64^3
    chunks) on the workload's field, resident on the device, with graph
    replay, in the fp32 and the mixed-precision form.  Per form: a
    saturation run (every request offered at once, offered load above
    capacity: capacity rays/s and points/s, and the latency of queued
    requests), then Poisson arrivals at `load` times that capacity (in
    requests/s) for the latency at a stated load below capacity."""
    import torch
    from angelo_amd import serving
    model = wl.model
    model.eval()
    out = {}
    for prec in ("fp32", "mixed"):
        model.set_precision(prec)
        srv = serving.FieldServer(model, rays_per_step=1 << 17, sdf_chunks_per_step=8,
                                  use_graphs=True)
        warm = serving.poisson_requests(8, 1000.0, seed=seed + 100)
        srv.run(warm, realtime=False)       # warm-up: graphs, workspaces, direction tables
        res = {}
        rate = None
        for mode in ("saturation", "offered"):
            srv.rays_rendered = srv.points_queried = 0
            if mode == "saturation":
                reqs = serving.poisson_requests(n_requests, 1000.0, seed=seed)
                wall = srv.run(reqs, realtime=False)
                rate = load * len(reqs) / wall
            else:
                reqs = serving.poisson_requests(max(n_requests // 2, 1), rate, seed=seed + 1)
                wall = srv.run(reqs, realtime=True)
            st = serving.latency_stats(reqs)
            res[mode] = {"p50_ms": round(st["p50_s"] * 1e3, 2),
                         "p99_ms": round(st["p99_s"] * 1e3, 2),
                         "mean_ms": round(st["mean_s"] * 1e3, 2), "requests": st["n"],
                         "rays_per_s": round(srv.rays_rendered / wall, 1),
                         "sdf_points_per_s": round(srv.points_queried / wall, 1),
                         "wall_s": round(wall, 3)}
            if mode == "offered":
                res[mode]["arrival_rate_hz"] = round(rate, 3)
            log(f"[bench] serving {prec} {mode}: {json.dumps(res[mode])}")
        torch.cuda.synchronize()
        sat = res["saturation"]
        out[prec] = {"capacity_rays_per_s": sat["rays_per_s"],
                     "capacity_sdf_points_per_s": sat["sdf_points_per_s"],
                     "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"], "requests": sat["requests"],
                     "saturation": sat, "offered": res["offered"], "graph_replay": srv.use_graphs,
                     "captured_graphs": len(srv._render_graphs) + (srv._sdf_graph is not None)}
    model.set_precision("fp32")
    return out


def time_steps(wl, steps, warmup):
    """W untimed steps, then K timed steps bracketed by device syncs and HIP
    events on the current stream, with the library's kernel timers on."""
    import torch
    from angelo_amd import _native as nat
    for i in range(warmup):
        wl.step(i)
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({warmup} steps); timing {steps} steps")
    nat.timer_enable(True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    total_samples = 0
    bg_before = wl.bg_samples
    torch.cuda.synchronize()
    keys = ("num_device_alloc", "num_device_free", "num_alloc_retries", "num_sync_all_streams")
    mem0 = torch.cuda.memory_stats()
    t0 = time.perf_counter()
    start.record()
    step_wall = []
    for i in range(steps):
        ts = time.perf_counter()
        total_samples += wl.step(warmup + i)
        step_wall.append(time.perf_counter() - ts)
    wl.drain()    # a march started for the step after the window is timed too
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mem1 = torch.cuda.memory_stats()
    # device allocations and frees by torch's caching allocator inside the
    # timed steps (0 in steady state), its out-of-memory retries, and the
    # peak of reserved device memory
    alloc = {k: int(mem1.get(k, 0) - mem0.get(k, 0)) for k in keys}
    alloc["peak_reserved_gb"] = round(mem1.get("reserved_bytes.all.peak", 0) / 1e9, 2)
    nat.sdf_bwd_check()   # the pair passes of every timed backward agreed
    kern = {}
    for k in TIMED_KERNELS:
        ms, cnt = nat.timer_read(k)
        if cnt:
            kern[k] = {"total_ms": ms, "launches": cnt, "avg_ms": ms / cnt}
    pairs = nat.timer_units("table_pairs") / steps
    host_reads = nat.timer_units("host_syncs") / steps
    nat.timer_enable(False)
    return {"gpu_ms": start.elapsed_time(stop), "wall": wall, "step_wall": step_wall,
            "kern": kern, "samples": total_samples, "pairs_per_step": pairs,
            "sdf_bwd_host_reads": host_reads, "allocator": alloc,
            "bg_samples": wl.bg_samples - bg_before}


def level_sweep(wl, levels=(1, 2, 4, 8, 12, 16), reps=3):
    """The fused SDF forward (training form: FD + curvature evaluations,
    stored encodings) over one step's marched samples with 1..L active
    levels, `reps` launches per setting, then `reps` launches of the same
    forward without the saved buffers (for HBM counters per setting under a
    profiler: the training-form launches are the first `reps` of each
    setting): per setting the average time from HIP events on the current
    stream."""
    import torch
    from angelo_amd import functions as fn
    dev = wl.device
    model, geo = wl.model, wl.model.geometry
    b = wl.sampler.sample(wl.spec["rays"], wl.sample_gen, model)
    rays = b["rays"].float().contiguous()
    with torch.no_grad():
        ridx, t0, t1, _ = model.march(rays)
    pts = fn.sample_points(rays, ridx, t0, t1)[0].detach()
    randn = torch.randn(pts.shape[0], 3, device=dev,
                        generator=torch.Generator(device=dev).manual_seed(5))
    keep = geo.grid_module.active_levels
    out = {"samples": int(pts.shape[0]), "launches_per_setting": reps, "ms": {}}
    try:
        for k in levels:
            if k > geo.grid_module.n_levels:
                continue
            geo.grid_module.active_levels = k
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(reps):
                geo.fused(pts, True, True, randn)
            ev[1].record()
            torch.cuda.synchronize()
            out["ms"][k] = round(ev[0].elapsed_time(ev[1]) / reps, 4)
            # the same forward without the backward's saved buffers (no grad
            # mode: no stored encodings, no FD values)
            torch.cuda.synchronize()
            with torch.no_grad():
                ev[0].record()
                for _ in range(reps):
                    geo.fused(pts, True, True, randn)
                ev[1].record()
            torch.cuda.synchronize()
            out.setdefault("ms_no_store", {})[k] = round(ev[0].elapsed_time(ev[1]) / reps, 4)
            log(f"[bench] level sweep: {k} active levels, {out['ms'][k]:.3f} ms per forward, "
                f"{out['ms_no_store'][k]:.3f} ms without the saved buffers")
    finally:
        geo.grid_module.active_levels = keep
    return out


def mixed_bench(dev, name, steps, warmup, fp32_value, rays=None, pipeline=False):
    """The mixed-precision form of the same step (fp16 table rows, bf16 MFMA
    radiance, fp16 table-gradient pairs with dynamic loss scaling, Kahan
    compositing, fp64 loss reductions): samples/s, its ratio to the fp32
    form, per-kernel times and the loss scaler's state."""
    from angelo_amd.workloads import Workload, WORKLOADS
    if name not in WORKLOADS:
        return None
    log(f"[bench] building workload {name}")
    wl = Workload(name, dev, rays=rays, pipeline=pipeline)
    t = time_steps(wl, steps, warmup)
    value = t["samples"] / (t["gpu_ms"] / 1e3)
    out = {"workload": name, "value": round(value, 1), "unit": "samples/s",
           "ms_per_step": round(t["gpu_ms"] / steps, 4),
           "samples_per_step": round(t["samples"] / steps, 1),
           "ratio_to_fp32": round(value / fp32_value, 4),
           "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in t["kern"].items()},
           "table_pairs_per_step": round(t["pairs_per_step"], 1),
           "final_loss_scale": wl.scaler.scale if wl.scaler else None,
           "skipped_steps": wl.skipped_steps}
    log(f"[bench] mixed: {out['ms_per_step']:.3f} ms/step, {value / 1e6:.2f} M samples/s, "
        f"{out['ratio_to_fp32']:.3f}x fp32")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="neuralangelo-fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=64,
                    help="rays of the CPU baseline's training steps (1024: the full protocol)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision form of the workload")
    ap.add_argument("--rays", type=int, default=None)
    ap.add_argument("--table-atomic", action="store_true",
                    help="SDF table gradient by float atomics (run-to-run summation order "
                         "varies) instead of the deterministic sort + reduce")
    ap.add_argument("--level-sweep", action="store_true",
                    help="only time the fused SDF forward over one step's samples with 1, 2, "
                         "4, 8, 12, 16 active levels (3 launches each; for per-level HBM "
                         "counters under rocprofv3) and print that as the JSON line")
    ap.add_argument("--sync-free", action="store_true",
                    help="SDF backward without its pair-count host read (pair arrays and "
                         "sort sized for the worst-case pair count)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="march each step's rays at the start of the step (by default the "
                         "next step's rays are drawn and marched on a second stream while "
                         "this step's backward runs, with the same step sequence; off with "
                         "--sync-free, whose worst-case pair arrays already fill most of "
                         "the device memory)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests in the serving measurement (0: skip it)")
    ap.add_argument("--serve-load", type=float, default=0.6,
                    help="Poisson arrival rate of the offered-load run, as a fraction of the "
                         "saturation run's requests/s")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy three times (the "
                         "reference kernel for scaling HBM counters of a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    import torch
    from angelo_amd import _native as nat
    from angelo_amd.workloads import Workload
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    nat.load()
    if args.calibrate_copy:
        src = torch.empty(1 << 28, device=dev, dtype=torch.float32).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 3 x 1 GiB device-to-device")
        del src, dst
    if args.sync_free:
        from angelo_amd import functions as fn
        fn.set_sdf_backward(sync_free=True)
    if args.table_atomic:
        nat.set_flag("table_atomic", 1)
    log(f"[bench] building workload {args.workload}")
    pipeline = not (args.no_pipeline or args.sync_free)
    wl = Workload(args.workload, dev, rays=args.rays, pipeline=pipeline)
    spec = wl.spec
    if args.level_sweep:
        if wl.uniform:
            print("--level-sweep needs a marched workload (neuralangelo-*)", file=sys.stderr)
            sys.exit(2)
        for i in range(args.warmup):
            wl.step(i)
        sweep = level_sweep(wl)
        print(json.dumps({"level_sweep": sweep, "workload": args.workload}), flush=True)
        return
    t = time_steps(wl, args.steps, args.warmup)
    gpu_ms, wall, step_wall, kern = t["gpu_ms"], t["wall"], t["step_wall"], t["kern"]
    total_samples, pairs_per_step = t["samples"], t["pairs_per_step"]
    host_reads = t["sdf_bwd_host_reads"]
    alloc = t["allocator"]
    ms_per_step = gpu_ms / args.steps
    samples_per_step = total_samples / args.steps
    bg_per_step = t["bg_samples"] / args.steps
    fg_per_step = samples_per_step - bg_per_step   # the byte models count foreground samples
    rays = spec["rays"]
    value = total_samples / (gpu_ms / 1e3)
    log(f"[bench] {ms_per_step:.3f} ms/step, {samples_per_step:.0f} samples/step, wall {wall:.2f}s")
    sw = sorted(step_wall)
    log(f"[bench] caching allocator in the timed steps: {alloc}")
    log(f"[bench] host time per step: min {1e3 * sw[0]:.2f} ms, median "
        f"{1e3 * sw[len(sw) // 2]:.2f} ms, max {1e3 * sw[-1]:.2f} ms")
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"[bench]   {k:14s} {v['avg_ms']:.3f} ms x {v['launches']}  "
            f"({100 * v['total_ms'] / gpu_ms:.1f}% of step time)")
    # roofline of the dominant kernel, and of every kernel with a byte model
    # (launch averages: one launch per step for each of these kernels)
    roof, per_kernel = None, {}
    for k, v in kern.items():
        if bg_per_step and k in ("radix_sort", "seg_reduce"):
            continue   # shared with the background encode's backward: no per-step byte model
        nbytes = algorithmic_bytes(k, spec, fg_per_step, rays, pairs_per_step, wl.mixed)
        if nbytes is None:
            continue
        # the byte models are per step: divide by the kernel's time per step
        # (its launch time where it runs once a step, as the fused kernels do)
        ach = nbytes / (v["total_ms"] / args.steps / 1e3) / 1e9
        per_kernel[k] = {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                         "bytes": int(nbytes), "avg_ms": round(v["avg_ms"], 4)}
    # measured HBM traffic (profiled run of this command; per launch, scaled
    # by samples/step when this run's sample count differs from the profile's)
    prof = profiled_traffic(args.workload)
    if prof:
        ratio = samples_per_step / max(prof["samples_per_step"], 1.0)
        for k, pk in per_kernel.items():
            t = prof["kernels"].get(k)
            if t:
                pk["traffic"] = int((t["fetch_bytes"] + t["write_bytes"]) * ratio)
    if per_kernel:
        # the kernel with the most time per step among those with a byte model
        dom = max(per_kernel, key=lambda k: kern[k]["total_ms"])
        if dom in per_kernel:
            pk = per_kernel[dom]
            roof = {"kernel": dom, "bound": "hbm", "achieved": pk["achieved"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": pk["frac"], "traffic": pk.get("traffic"),
                    "algorithmic_bytes": pk["bytes"], "avg_ms": pk["avg_ms"]}
            if pk.get("traffic"):
                roof["traffic_source"] = f"profiles/hbm_traffic_{args.workload}.md"
                roof["traffic_gbs"] = round(pk["traffic"] / (pk["avg_ms"] / 1e3) / 1e9, 1)
    mixed = None
    mixed_name = args.workload.replace("-fp32", "-mixed") if "-fp32" in args.workload else None
    if not args.no_mixed and not wl.mixed and mixed_name:
        del t
        mixed = mixed_bench(dev, mixed_name, args.steps, args.warmup, value, rays=args.rays,
                            pipeline=pipeline)
    serve = None
    if args.serve_requests > 0:
        log(f"[bench] serving: {args.serve_requests} requests at saturation, then at "
            f"{args.serve_load} of capacity ...")
        serve = serving_bench(wl, args.serve_requests, args.serve_load)
        log(f"[bench] serving: {json.dumps(serve)}")
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        try:
            cpu = cpu_baseline(args.workload, wl, train_rays=args.cpu_rays)
        except Exception as e:  # noqa: BLE001 - reported, not fatal for the GPU number
            cpu = {"error": repr(e)}
    line = {
        "metric": "train samples/s (fwd+bwd)", "value": round(value, 1), "unit": "samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 table / bf16 MFMA, fp32 accumulate" if wl.mixed else "fp32",
        "data": "synthetic (analytic scene, seeded cameras)",
        "config": {"workload": args.workload, "rays_per_step": rays,
                   "samples_per_step": round(samples_per_step, 1),
                   "background_samples_per_step": round(bg_per_step, 1), "levels": spec["levels"],
                   "log2_hashmap_size": spec["log2T"], "evals_per_sample": evals_per_sample(spec),
                   "grad_type": spec.get("grad_type", "finite_difference"),
                   "learned_background": bool(spec.get("learned_background", False)),
                   "active_levels": int(spec.get("start_level") or spec["levels"]),
                   "optimizer": "AdamW (in the timed step)"},
        "rays_per_s": round(rays * args.steps / (gpu_ms / 1e3), 1),
        "loss_scaling": ({"final_scale": wl.scaler.scale, "skipped_steps": wl.skipped_steps}
                         if wl.mixed else None),
        "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kern.items()},
        "kernels_roofline": per_kernel,
        "table_pairs_per_step": round(pairs_per_step, 1),
        "sdf_bwd_sync_free": bool(args.sync_free),
        "march_pipelined": bool(wl.pipeline),
        "table_gradient": "atomic" if args.table_atomic else "deterministic (sort + reduce)",
        "sdf_bwd_host_reads_per_step": host_reads,
        "allocator_in_timed_steps": alloc,
        "roofline": roof,
        "mixed": mixed,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
