"""Serving loop over a resident NeuS field: continuous batching of novel-view
renders and SDF-grid (marching-cubes lattice) queries.

The field (hash table, MLPs, variance) and the occupancy grid stay on the
device.  Requests arrive at arbitrary times; each is split into work units
(ray chunks of a render, 64^3-point chunks of an SDF lattice).  Every step
packs work units of all admitted requests into one render batch of
`rays_per_step` ray slots and one SDF batch of `sdf_chunks_per_step` lattice
chunks, runs them, and scatters the results back.  A request is admitted when
its result buffers fit the HBM budget next to those already admitted
(admission by footprint); a finished request frees its share.

Graph replay (use_graphs=True): the SDF batch is one captured graph over a
fixed point buffer; the render batch marches eagerly (the sample count is
data-dependent), then replays the graph captured for the sample-count bucket
(`sample_bucket`: base * 2^k * {1, 5/4, 3/2, 7/4}, the smallest that holds
the count) over
sample positions, fused SDF field (base + FD normal), fused radiance and
compositing, with padded samples outside every ray's range.  The captured
regions hold only kernel launches on buffers that live as long as the
graph: the fused field runs in the caller's order there (the spatial
order's library sort clears its state with memset nodes, which must not be
captured), the forward keeps no backward buffers (no grad mode), and the
table / weights are those of the warm-up run before capture.  Models with
the learned background render eagerly.
Per-ray and per-point results do not depend on what else is in the batch
(every kernel computes a ray or a point from its own inputs in a fixed
order), so a request renders the same alone and in a mix.

Reference counterparts: the eval-mode NeuSModel.forward with chunk_batch
(models/neus.py:285-293, models/utils.py:14-51) and the SDF lattice of
MarchingCubeHelper (models/geometry.py:36-113).
"""
import math
import time
from collections import deque

import torch

from . import _native as nat
from . import functions as fn
from .ray_utils import get_ray_directions, get_rays

SDF_CHUNK = 64 ** 3


class Request:
    """A render (pose c2w [3,4], W, H, fov) or an SDF lattice (resolution)."""

    _next_id = 0

    def __init__(self, kind, arrival, **params):
        if kind not in ("render", "sdf_grid"):
            raise ValueError(kind)
        self.kind, self.arrival, self.params = kind, float(arrival), params
        self.id = Request._next_id
        Request._next_id += 1
        self.admitted = None
        self.finished = None
        self.result = None
        self._cursor = 0
        self._total = 0
        self._done = 0

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    def n_units(self):
        if self.kind == "render":
            return self.params["W"] * self.params["H"]
        return _lattice_size(self.params["resolution"], self.params.get("radius", 1.0))

    def footprint_bytes(self):
        """Device bytes held while the request is in flight: the result
        buffers (render: rays 24 B + rgb/depth/opacity 20 B per pixel; SDF
        lattice: 4 B per point)."""
        n = self.n_units()
        return n * (24 + 20) if self.kind == "render" else n * 4


def _lattice_axes(resolution, radius):
    """torch.arange(-r, r, 2/resolution) per axis, as MarchingCubeHelper
    builds its lattice (models/geometry.py:42,62-64)."""
    intv = 2.0 / resolution
    a = torch.arange(-radius, radius, intv)
    return a, a, a


def _lattice_size(resolution, radius):
    a, b, c = _lattice_axes(resolution, radius)
    return len(a) * len(b) * len(c)


def sample_bucket(n, base):
    """Sample-count bucket of a render batch: the smallest of base * 2^k *
    {1, 5/4, 3/2, 7/4} that holds n (a captured graph per bucket; padding
    stays under a quarter of the batch)."""
    nb = int(base)
    while True:
        for q in (4, 5, 6, 7):
            cand = nb * q // 4
            if cand >= n:
                return cand
        nb *= 2


class FieldServer:
    def __init__(self, model, rays_per_step=1 << 16, sdf_chunks_per_step=4,
                 hbm_budget_bytes=32 << 30, sample_bucket=1 << 18, use_graphs=False,
                 background=(1.0, 1.0, 1.0)):
        self.model = model
        self.dev = model.geometry.grid_module.params.device
        self.R = int(rays_per_step)
        self.P = int(sdf_chunks_per_step) * SDF_CHUNK
        self.budget = int(hbm_budget_bytes)
        self.bucket = int(sample_bucket)
        self.use_graphs = use_graphs
        self.bg = torch.tensor(background, device=self.dev, dtype=torch.float32)
        self.aabb = model.scene_aabb.tolist()
        self.pending = deque()
        self.active = []
        self._clock = time.perf_counter   # run() rebases it on its start time
        self.done = []
        self.in_flight_bytes = 0
        self._dirs = {}
        self._render_graphs = {}
        self._sdf_graph = None
        # the tables the captured kernels read (the mixed form's fp16 copy
        # belongs to the grid, which drops it when the master may have
        # changed): held here for as long as the graphs live.  Graphs serve
        # the weights of their capture; a server is rebuilt after training.
        self._pinned = []
        # static buffers of the captured steps
        f32 = dict(device=self.dev, dtype=torch.float32)
        self.s_rays = torch.zeros(self.R, 6, **f32)
        self.s_pts = torch.zeros(self.P, 3, **f32)
        self.steps = 0
        self.rays_rendered = 0
        self.points_queried = 0

    # ---- request handling -------------------------------------------------
    def submit(self, req):
        self.pending.append(req)

    def _admit(self, now):
        while self.pending:
            req = self.pending[0]
            fp = req.footprint_bytes()
            if self.active and self.in_flight_bytes + fp > self.budget:
                break  # FIFO: later requests wait behind the head
            self.pending.popleft()
            self._start(req, now)
            self.in_flight_bytes += fp

    def _start(self, req, now):
        req.admitted = now
        req._total = req.n_units()
        if req.kind == "render":
            p = req.params
            W, H, fov = p["W"], p["H"], p.get("fov", 40.0)
            key = (W, H, fov)
            if key not in self._dirs:
                f = 0.5 * W / math.tan(0.5 * math.radians(fov))
                self._dirs[key] = get_ray_directions(W, H, f, f, W / 2, H / 2).to(self.dev)
            c2w = p["c2w"].to(self.dev, torch.float32)
            o, d = get_rays(self._dirs[key], c2w)
            d = torch.nn.functional.normalize(d, p=2, dim=-1)
            req.rays = torch.cat([o, d], -1).contiguous()
            n = req._total
            req.result = {"rgb": torch.empty(n, 3, device=self.dev),
                          "depth": torch.empty(n, device=self.dev),
                          "opacity": torch.empty(n, device=self.dev)}
        else:
            r = float(req.params.get("radius", 1.0))
            req.axes = [a.to(self.dev, torch.float32)
                        for a in _lattice_axes(req.params["resolution"], r)]
            req.result = {"sdf": torch.empty(req._total, device=self.dev)}
        self.active.append(req)

    def _finish(self, req, now):
        req.finished = now
        self.in_flight_bytes -= req.footprint_bytes()
        if req.kind == "render":
            del req.rays
        self.done.append(req)

    # ---- render batch -----------------------------------------------------
    def _render_stage(self, rays, ridx, t0, t1, packed):
        """Post-march render of the batch (captured per sample bucket)."""
        m = self.model
        pts, mid, dists = fn.sample_points(rays, ridx, t0, t1)
        # the caller's order: the spatial order's sort may not run in a graph
        sdf, _, normal, feature, _ = m.geometry.fused(pts, with_grad=True, with_laplace=False,
                                                      order=False if self.use_graphs else None)
        rgb = m.texture.fused(feature, normal, rays, ridx)
        cd = nat.CompositeDesc()
        cd.cos_anneal_ratio = float(m.cos_anneal_ratio)
        cd.kahan = int(m.kahan)
        _, opacity, depth, comp_rgb, _ = fn.composite(sdf, normal, rgb, m.inv_s().reshape(1), rays,
                                                      ridx, t0, t1, packed, cd)
        if getattr(m, "learned_background", False):
            # the background pass of the eval forward (models/neus.py:141-203,
            # 263-270) with this server's background colour behind it
            saved = m.background_color
            m.background_color = self.bg
            try:
                bg_rgb = m.forward_bg_(rays)["comp_rgb"]
            finally:
                m.background_color = saved
        else:
            bg_rgb = self.bg[None, :]
        full = comp_rgb + bg_rgb * (1.0 - opacity[:, None])
        return full, depth, opacity

    def _render_batch(self, rays):
        m = self.model
        t_near, t_far = fn.ray_aabb(rays, self.aabb)
        grid = m.occupancy_grid.binary if m.config.get("grid_prune", False) else None
        ridx, t0, t1, packed = fn.ray_marching(rays, t_near, t_far, self.aabb, grid,
                                               m.render_step_size, None, m.max_samples_per_ray)
        n = int(ridx.shape[0])
        nb = sample_bucket(n, self.bucket)
        if not self.use_graphs or getattr(m, "learned_background", False):
            return self._render_stage(rays, ridx, t0, t1, packed)
        g = self._render_graphs.get(nb)
        if g is None:
            g = self._capture_render(nb)
        s = g["static"]
        # the static buffers hold the batch: n samples of R rays
        assert rays.shape == s["rays"].shape and packed.shape == s["packed"].shape
        assert n <= s["ridx"].shape[0] and ridx.dtype == s["ridx"].dtype == torch.int32
        s["rays"].copy_(rays)
        s["ridx"].zero_()
        s["t0"].zero_()
        s["t1"].zero_()
        s["ridx"][:n].copy_(ridx)
        s["t0"][:n].copy_(t0)
        s["t1"][:n].copy_(t1)
        s["packed"].copy_(packed)
        g["graph"].replay()
        return g["out"]

    def _capture_render(self, nb):
        dev = self.dev
        st = {"rays": self.s_rays, "ridx": torch.zeros(nb, device=dev, dtype=torch.int32),
              "t0": torch.zeros(nb, device=dev), "t1": torch.zeros(nb, device=dev),
              "packed": torch.zeros(self.R, 2, device=dev, dtype=torch.int32)}
        # warm up once on the side stream (workspace sizes, attribute calls)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._render_stage(st["rays"], st["ridx"], st["t0"], st["t1"], st["packed"])
        torch.cuda.current_stream().wait_stream(side)
        self._pinned.append(self.model.geometry.grid_module.kernel_table())
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            out = self._render_stage(st["rays"], st["ridx"], st["t0"], st["t1"], st["packed"])
        rec = {"graph": graph, "static": st, "out": out}
        self._render_graphs[nb] = rec
        return rec

    # ---- SDF batch --------------------------------------------------------
    def _sdf_stage(self):
        # lattice points are in lattice order already (coherent gathers), and
        # the spatial order's sort may not run in a graph
        return self.model.geometry.fused(self.s_pts, with_grad=False, order=False)[0]

    def _sdf_batch(self, n):
        if not self.use_graphs:
            return self._sdf_stage()[:n]
        if self._sdf_graph is None:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._sdf_stage()
            torch.cuda.current_stream().wait_stream(side)
            self._pinned.append(self.model.geometry.grid_module.kernel_table())
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                out = self._sdf_stage()
            self._sdf_graph = (graph, out)
        graph, out = self._sdf_graph
        graph.replay()
        return out[:n]

    # ---- one step ---------------------------------------------------------
    @torch.no_grad()
    def step(self, now=None):
        """Admits what fits, runs one render batch and one SDF batch over the
        admitted requests' next work units, scatters the results.  Returns
        False when there was nothing to do."""
        now = self._clock() if now is None else now
        self._admit(now)
        if not self.active:
            return False
        renders = [r for r in self.active if r.kind == "render" and r._cursor < r._total]
        grids = [r for r in self.active if r.kind == "sdf_grid" and r._cursor < r._total]
        # render: fill the ray slots in admission order
        if renders:
            pieces, fill = [], 0
            for r in renders:
                take = min(self.R - fill, r._total - r._cursor)
                if take <= 0:
                    break
                pieces.append((r, r._cursor, take, fill))
                self.s_rays[fill:fill + take].copy_(r.rays[r._cursor:r._cursor + take])
                r._cursor += take
                fill += take
            if fill < self.R:   # idle slots: rays that miss the scene box
                self.s_rays[fill:, 0:3] = 1e3
                self.s_rays[fill:, 3:6] = torch.tensor([0.0, 0.0, 1.0], device=self.dev)
            rgb, depth, opacity = self._render_batch(self.s_rays)
            for r, c0, take, f0 in pieces:
                r.result["rgb"][c0:c0 + take].copy_(rgb[f0:f0 + take])
                r.result["depth"][c0:c0 + take].copy_(depth[f0:f0 + take])
                r.result["opacity"][c0:c0 + take].copy_(opacity[f0:f0 + take])
                r._done += take
            self.rays_rendered += fill
        # SDF lattice: fill the point slots chunk by chunk
        if grids:
            pieces, fill = [], 0
            for r in grids:
                while fill < self.P and r._cursor < r._total:
                    take = min(SDF_CHUNK, self.P - fill, r._total - r._cursor)
                    X, Y, Z = (len(a) for a in r.axes)
                    idx = torch.arange(r._cursor, r._cursor + take, device=self.dev)
                    xs, ys, zs = r.axes
                    self.s_pts[fill:fill + take] = torch.stack(
                        [xs[idx // (Y * Z)], ys[(idx // Z) % Y], zs[idx % Z]], -1)
                    pieces.append((r, r._cursor, take, fill))
                    r._cursor += take
                    fill += take
                if fill >= self.P:
                    break
            sdf = self._sdf_batch(fill)
            for r, c0, take, f0 in pieces:
                r.result["sdf"][c0:c0 + take].copy_(sdf[f0:f0 + take])
                r._done += take
            self.points_queried += fill
        torch.cuda.current_stream().synchronize()
        t = self._clock()   # the same time base as `now` and the arrivals
        for r in list(self.active):
            if r._done >= r._total:
                self.active.remove(r)
                self._finish(r, t)
        self.steps += 1
        return True

    def run(self, requests, realtime=True):
        """Serves `requests` (each with an arrival time in seconds from the
        start).  realtime: requests are submitted when their arrival time has
        passed on the wall clock; otherwise all are submitted at once."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        self._clock = lambda: time.perf_counter() - t0
        i = 0
        try:
            while i < len(reqs) or self.pending or self.active:
                now = self._clock()
                while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                    self.submit(reqs[i])
                    i += 1
                if not self.step(now=self._clock()):
                    if i < len(reqs):
                        time.sleep(max(0.0, min(0.002, reqs[i].arrival - now)))
        finally:
            self._clock = time.perf_counter
        return time.perf_counter() - t0


def latency_stats(requests):
    lat = sorted(r.latency for r in requests if r.latency is not None)
    if not lat:
        return {}

    def q(p):
        k = min(len(lat) - 1, max(0, int(math.ceil(p * len(lat))) - 1))
        return lat[k]

    return {"p50_s": q(0.50), "p99_s": q(0.99), "mean_s": sum(lat) / len(lat), "n": len(lat)}


def poisson_requests(n, rate_hz, seed=0, render_sizes=(256, 512, 1024),
                     sdf_resolutions=(128, 256, 512), sdf_fraction=0.25, radius=2.5):
    """A seeded mix: exponential inter-arrival times at `rate_hz`, renders at
    the given square sizes from seeded poses on a sphere of `radius` looking
    at the origin, SDF lattices at the given resolutions."""
    from .scenes import look_at_c2w
    g = torch.Generator().manual_seed(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += float(-math.log(1.0 - torch.rand(1, generator=g).item()) / rate_hz)
        if torch.rand(1, generator=g).item() < sdf_fraction:
            res = sdf_resolutions[int(torch.randint(len(sdf_resolutions), (1,), generator=g))]
            out.append(Request("sdf_grid", t, resolution=res))
        else:
            size = render_sizes[int(torch.randint(len(render_sizes), (1,), generator=g))]
            v = torch.randn(3, generator=g)
            eye = v / v.norm() * radius
            out.append(Request("render", t, c2w=look_at_c2w(eye), W=size, H=size, fov=40.0))
    return out
