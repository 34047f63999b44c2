"""torch.autograd.Function wrappers over the HIP kernels (C-ABI in
include/angelo_amd.h).  Each forward and backward is one or a few native
launches on the current stream; there is no CPU or eager-torch fallback.
"""
import math

import torch

from . import _native as nat


def _f32c(t):
    return t.contiguous() if t.dtype == torch.float32 else t.float().contiguous()


class HashGridFn(torch.autograd.Function):
    """tcnn HashGrid encode behind ProgressiveBandHashGrid
    (models/network_utils.py:45-47,56-59): x[n,3] in [0,1] -> [n, L*F]."""

    @staticmethod
    def forward(ctx, x, table, grid, active_levels):
        nat.require_cuda(x, table)
        lib = nat.load()
        x = _f32c(x)
        n = x.shape[0]
        out = torch.empty(n, grid.n_levels * grid.n_features, device=x.device, dtype=torch.float32)
        so = SPATIAL_ORDER
        order = None
        if so["encode"] and n >= so["min_samples"]:
            # points of [0, 1]^3 in Morton order (the keys take [-r, r]^3)
            order = spatial_order(x - 0.5, 0.5)[0]
            nat.check(lib.am_hashgrid_fwd_ordered(nat.sptr(grid), nat.ptr(order), nat.ptr(x), n,
                                                  nat.ptr(table), int(active_levels),
                                                  nat.ptr(out), nat.stream()),
                      "am_hashgrid_fwd_ordered")
        else:
            nat.check(lib.am_hashgrid_fwd(nat.sptr(grid), nat.ptr(x), n, nat.ptr(table),
                                          int(active_levels), nat.ptr(out), nat.stream()),
                      "am_hashgrid_fwd")
        ctx.save_for_backward(x, table, order if order is not None else x.new_empty(0))
        ctx.grid, ctx.active = grid, int(active_levels)
        return out

    @staticmethod
    def backward(ctx, g):
        x, table, order = ctx.saved_tensors
        lib = nat.load()
        g = _f32c(g)
        n = x.shape[0]
        grid = ctx.grid
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        dtable = torch.empty_like(table) if ctx.needs_input_grad[1] else None
        ws = None
        if dtable is not None:
            ws = nat.held_workspace("hashgrid_bwd", lib.am_hashgrid_bwd_workspace_size(nat.sptr(grid), n, ctx.active),
                               x.device)
        rest = (nat.ptr(x), n, nat.ptr(table), nat.ptr(g), ctx.active, nat.ptr(dtable),
                nat.ptr(dx), nat.ptr(ws), ws.numel() if ws is not None else 0, nat.stream())
        if order.numel():
            nat.check(lib.am_hashgrid_bwd_ordered(nat.sptr(grid), nat.ptr(order), *rest),
                      "am_hashgrid_bwd_ordered")
        else:
            nat.check(lib.am_hashgrid_bwd(nat.sptr(grid), *rest), "am_hashgrid_bwd")
        return dx, dtable, None, None


class MLPFn(torch.autograd.Function):
    """VanillaMLP forward/backward (models/network_utils.py:96-140) on a flat
    parameter vector [W0,b0,...] (weight-norm applied by the caller)."""

    @staticmethod
    def forward(ctx, x, params, desc):
        nat.require_cuda(x, params)
        lib = nat.load()
        x = _f32c(x)
        params = _f32c(params)
        n = x.shape[0]
        y = torch.empty(n, desc.n_out, device=x.device, dtype=torch.float32)
        nat.check(lib.am_mlp_fwd(nat.sptr(desc), nat.ptr(x), n, nat.ptr(params), nat.ptr(y),
                                 nat.stream()), "am_mlp_fwd")
        ctx.save_for_backward(x, params)
        ctx.desc = desc
        return y

    @staticmethod
    def backward(ctx, g):
        x, params = ctx.saved_tensors
        lib = nat.load()
        g = _f32c(g)
        n = x.shape[0]
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        dparams = torch.empty_like(params)
        ws = nat.held_workspace("mlp_bwd", lib.am_mlp_bwd_workspace_size(nat.sptr(ctx.desc), n), x.device)
        nat.check(lib.am_mlp_bwd(nat.sptr(ctx.desc), nat.ptr(x), n, nat.ptr(params), nat.ptr(g),
                                 nat.ptr(dx), nat.ptr(dparams), nat.ptr(ws), ws.numel(),
                                 nat.stream()), "am_mlp_bwd")
        return dx, dparams, None


def hashgrid_encode(x, table, grid, active_levels):
    return HashGridFn.apply(x, table, grid, active_levels)


def mlp_apply(x, params, desc):
    return MLPFn.apply(x, params, desc)


def sh_encode(dirs01, degree):
    """tcnn SphericalHarmonics forward (models/texture.py:25); directions carry
    no gradient on the NeuS path."""
    nat.require_cuda(dirs01)
    d = _f32c(dirs01)
    out = torch.empty(d.shape[0], degree * degree, device=d.device, dtype=torch.float32)
    nat.check(nat.load().am_sh_fwd(nat.ptr(d), d.shape[0], int(degree), nat.ptr(out),
                                   nat.stream()), "am_sh_fwd")
    return out


# Backward form of the fused SDF field.  "store": the forward keeps every
# evaluation's encoding (and the curvature evaluations' position
# derivatives), so the backward reads them back in coalesced rows instead of
# gathering the table again; "recompute": the backward re-gathers (no extra
# memory); "auto": store while the store fits store_limit_bytes.
# sync_free: the backward sizes its pair arrays for the worst-case pair
# count (library flag "sdf_bwd_sync_free") in a workspace held at that
# bound, so it makes no host read of the count (its sort then covers the
# bound; same gradients).
SDF_BACKWARD = {"mode": "auto", "store_limit_bytes": 24 << 30, "sync_free": False}

# The SDF backward's workspace is sized by its actual table-gradient pairs
# (known only after its count pass), held and grown across steps.
SDF_BWD_WORKSPACE = nat.GrowingWorkspace()


def set_sdf_backward(mode="auto", store_limit_bytes=None, sync_free=None):
    if mode not in ("auto", "store", "recompute"):
        raise ValueError(f"unknown SDF backward form {mode!r}")
    SDF_BACKWARD["mode"] = mode
    if store_limit_bytes is not None:
        SDF_BACKWARD["store_limit_bytes"] = int(store_limit_bytes)
    if sync_free is not None:
        nat.set_flag("sdf_bwd_sync_free", 1 if sync_free else 0)
        SDF_BACKWARD["sync_free"] = bool(sync_free)


def _sdf_store_floats(lib, grid, mdesc, sdesc, n):
    mode = SDF_BACKWARD["mode"]
    if mode == "recompute":
        return 0
    floats = int(lib.am_sdf_store_floats(nat.sptr(grid), nat.sptr(mdesc), nat.sptr(sdesc), n))
    if floats < 0:
        nat.check(-1, "am_sdf_store_floats")
    if mode == "auto" and floats * 4 > SDF_BACKWARD["store_limit_bytes"]:
        return 0
    return floats


# The fused SDF field runs its samples in spatial (Morton) order when there
# are at least min_samples of them; the kernels read inputs and write outputs
# (and read upstream gradients) in the caller's order through the permutation.
# Cells of 2^-bits of the box per axis; within a cell the caller's (ray)
# order is kept, so neighbouring samples of a ray still merge their table-
# gradient pairs.  5 bits measured best on the Neuralangelo step (10 bits:
# the same encode gain, 5 % more pairs).
# "encode": the same for the stand-alone hash-grid encode (off: on the
# learned background's samples it measured 0.3 ms slower per step).
SPATIAL_ORDER = {"on": True, "encode": False, "min_samples": 1 << 16, "bits": 5}


def spatial_order(pts, radius, bits=None):
    """Stable sort of the samples by their Morton key (`bits` per axis):
    (perm, inverse), int32."""
    lib = nat.load()
    n = pts.shape[0]
    bits = int(SPATIAL_ORDER["bits"] if bits is None else bits)
    perm = torch.empty(n, device=pts.device, dtype=torch.int32)
    inv = torch.empty_like(perm)
    ws = nat.held_workspace("morton_order", lib.am_morton_order_workspace_size(n, bits), pts.device)
    nat.check(lib.am_morton_order(nat.ptr(pts), n, float(radius), bits, nat.ptr(perm),
                                  nat.ptr(inv), nat.ptr(ws), ws.numel(), nat.stream()),
              "am_morton_order")
    return perm, inv


class SDFFieldFn(torch.autograd.Function):
    """Fused VolumeSDF.forward (models/geometry.py:195-287): outputs sdf [n],
    grad [n,3], normal [n,3] = F.normalize(grad), feature [n, n_out+3],
    laplace [n] (grad/normal/laplace empty when not requested)."""

    @staticmethod
    def forward(ctx, pts, randn, table, params, grid, mdesc, sdesc, table_k=None, order=None,
                keep=True):
        """table: the fp32 master (its gradient is returned); table_k: the
        rows the kernels read (the master, or its fp16 copy in the mixed
        form); order: int32 evaluation order of the samples (None: theirs),
        inputs and outputs stay in the caller's order; keep: the caller will
        run the backward, so the forward writes the FD values and stored
        encodings it reads (decided by the caller before apply: inside
        forward grad mode is always off, and needs_input_grad is set for a
        Parameter table even under no_grad)."""
        if table_k is None:
            table_k = table
        nat.require_cuda(pts, table, params, table_k)
        lib = nat.load()
        pts = _f32c(pts)
        n = pts.shape[0]
        dev = pts.device
        f32 = dict(device=dev, dtype=torch.float32)
        sdf = torch.empty(n, **f32)
        feature = torch.empty(n, mdesc.n_out + 3, **f32)
        grad = torch.empty(n, 3, **f32) if sdesc.with_grad else torch.empty(0, 3, **f32)
        normal = torch.empty(n, 3, **f32) if sdesc.with_grad else torch.empty(0, 3, **f32)
        lap = torch.empty(n, **f32) if sdesc.with_laplace else torch.empty(0, **f32)
        rnd = _f32c(randn) if (sdesc.with_laplace and randn is not None) else None
        keep = bool(keep)
        fdv = torch.empty(n, 12, **f32) if (sdesc.with_grad and keep) else None
        st_floats = _sdf_store_floats(lib, grid, mdesc, sdesc, n) if (keep and n > 0) else 0
        store = torch.empty(st_floats, **f32) if st_floats > 0 else None
        outs = (nat.ptr(table_k), nat.ptr(params), nat.ptr(sdf),
                nat.ptr(grad) if sdesc.with_grad else None,
                nat.ptr(normal) if sdesc.with_grad else None, nat.ptr(feature),
                nat.ptr(lap) if sdesc.with_laplace else None, nat.ptr(fdv), nat.ptr(store),
                nat.stream())
        if order is None:
            nat.check(lib.am_sdf_fwd(nat.sptr(grid), nat.sptr(mdesc), nat.sptr(sdesc),
                                     nat.ptr(pts), nat.ptr(rnd), n, *outs), "am_sdf_fwd")
        else:
            nat.check(lib.am_sdf_fwd_ordered(nat.sptr(grid), nat.sptr(mdesc), nat.sptr(sdesc),
                                             nat.ptr(order), nat.ptr(pts), nat.ptr(rnd), n,
                                             *outs), "am_sdf_fwd_ordered")
        empty = pts.new_empty(0)
        ctx.save_for_backward(pts, rnd if rnd is not None else empty, table, params,
                              fdv if fdv is not None else empty,
                              store if store is not None else empty, table_k,
                              order if order is not None else empty)
        ctx.descs = (grid, mdesc, sdesc)
        return sdf, grad, normal, feature, lap

    @staticmethod
    def backward(ctx, g_sdf, g_grad, g_normal, g_feature, g_lap):
        pts, rnd, table, params, fdv, store, table_k, order = ctx.saved_tensors
        grid, mdesc, sdesc = ctx.descs
        lib = nat.load()
        n = pts.shape[0]

        def opt(t, numel):
            return _f32c(t) if (t is not None and t.numel() == numel and numel > 0) else None

        gs = opt(g_sdf, n)
        gg = opt(g_grad, 3 * n) if sdesc.with_grad else None
        gn = opt(g_normal, 3 * n) if sdesc.with_grad else None
        gf = opt(g_feature, n * (mdesc.n_out + 3))
        gl = opt(g_lap, n) if sdesc.with_laplace else None
        dtable = torch.empty_like(table)
        dparams = torch.empty_like(params)
        rest = (nat.ptr(rnd) if rnd.numel() else None, n, nat.ptr(table_k), nat.ptr(params),
                nat.ptr(fdv) if fdv.numel() else None, nat.ptr(store) if store.numel() else None,
                nat.ptr(gs), nat.ptr(gg), nat.ptr(gn), nat.ptr(gf), nat.ptr(gl), nat.ptr(dtable),
                nat.ptr(dparams))
        ds = (nat.sptr(grid), nat.sptr(mdesc), nat.sptr(sdesc))
        need = int(lib.am_sdf_bwd_workspace_size(*ds, n)) if SDF_BACKWARD["sync_free"] else 0
        if order.numel():
            SDF_BWD_WORKSPACE.run(pts.device, lambda wp, wb: lib.am_sdf_bwd_ordered(
                *ds, nat.ptr(order), nat.ptr(pts), *rest, wp, wb, nat.stream()),
                "am_sdf_bwd_ordered", need)
        else:
            SDF_BWD_WORKSPACE.run(pts.device, lambda wp, wb: lib.am_sdf_bwd(
                *ds, nat.ptr(pts), *rest, wp, wb, nat.stream()), "am_sdf_bwd", need)
        return None, None, dtable, dparams, None, None, None, None, None, None


class RadianceFn(torch.autograd.Function):
    """Fused VolumeRadiance.forward (models/texture.py:23-30) with sigmoid colour:
    rgb = sigmoid(MLP(cat(feature, SH((d+1)/2), normal))), d = rays[ridx, 3:6];
    base_color (VolumeDualColor, :38-55) adds sigmoid(feature[:, 1:4])."""

    @staticmethod
    def forward(ctx, feature, normal, params, rays, ray_indices, mdesc, sh_degree, base_color=0):
        nat.require_cuda(feature, normal, params, rays, ray_indices)
        lib = nat.load()
        feature = _f32c(feature)
        normal = _f32c(normal)
        n = feature.shape[0]
        rgb = torch.empty(n, mdesc.n_out, device=feature.device, dtype=torch.float32)
        nat.check(lib.am_radiance_fwd(nat.sptr(mdesc), int(sh_degree), int(base_color),
                                      nat.ptr(feature),
                                      feature.shape[1], nat.ptr(rays), nat.ptr(ray_indices),
                                      nat.ptr(normal), n, nat.ptr(params), nat.ptr(rgb),
                                      nat.stream()), "am_radiance_fwd")
        ctx.save_for_backward(feature, normal, params, rays, ray_indices)
        ctx.mdesc, ctx.deg, ctx.base = mdesc, int(sh_degree), int(base_color)
        return rgb

    @staticmethod
    def backward(ctx, g_rgb):
        feature, normal, params, rays, ridx = ctx.saved_tensors
        lib = nat.load()
        n = feature.shape[0]
        g_rgb = _f32c(g_rgb)
        # the kernel writes every column of both (the input always ends with
        # the three normal columns)
        gf = torch.empty_like(feature) if ctx.needs_input_grad[0] else None
        gn = torch.empty_like(normal) if ctx.needs_input_grad[1] else None
        dparams = torch.empty_like(params)
        ws = nat.held_workspace("radiance_bwd", lib.am_radiance_bwd_workspace_size(nat.sptr(ctx.mdesc), n),
                           feature.device)
        nat.check(lib.am_radiance_bwd(nat.sptr(ctx.mdesc), ctx.deg, ctx.base, nat.ptr(feature),
                                      feature.shape[1], nat.ptr(rays), nat.ptr(ridx),
                                      nat.ptr(normal), n, nat.ptr(params), nat.ptr(g_rgb),
                                      nat.ptr(gf), nat.ptr(gn), nat.ptr(dparams), nat.ptr(ws),
                                      ws.numel(), nat.stream()), "am_radiance_bwd")
        return gf, gn, dparams, None, None, None, None, None


class CompositeFn(torch.autograd.Function):
    """NeuS alpha (models/neus.py:117-139) + render_weight_from_alpha +
    accumulate_along_rays (:234-241).  Outputs weights [n], opacity [R],
    depth [R], comp_rgb [R,3], comp_normal [R,3]."""

    @staticmethod
    def forward(ctx, sdf, normal, rgb, inv_s, rays, ray_indices, t_starts, t_ends, packed, cdesc):
        nat.require_cuda(sdf, normal, rgb, inv_s, rays, packed)
        lib = nat.load()
        n = sdf.shape[0]
        R = packed.shape[0]
        dev = sdf.device
        f32 = dict(device=dev, dtype=torch.float32)
        weights = torch.empty(n, **f32)
        trans = torch.empty(n, **f32)
        opacity = torch.empty(R, **f32)
        depth = torch.empty(R, **f32)
        comp_rgb = torch.empty(R, 3, **f32)
        comp_normal = torch.empty(R, 3, **f32)
        cn_raw = torch.empty(R, 3, **f32)
        sdf, normal, rgb = _f32c(sdf), _f32c(normal), _f32c(rgb)
        inv_s = _f32c(inv_s.reshape(1))
        nat.check(lib.am_composite_fwd(nat.sptr(cdesc), nat.ptr(sdf), nat.ptr(normal), nat.ptr(rays),
                                       nat.ptr(ray_indices), nat.ptr(t_starts), nat.ptr(t_ends),
                                       nat.ptr(rgb), nat.ptr(inv_s), nat.ptr(packed), R, n,
                                       nat.ptr(weights), nat.ptr(trans), nat.ptr(opacity),
                                       nat.ptr(depth), nat.ptr(comp_rgb), nat.ptr(comp_normal),
                                       nat.ptr(cn_raw), nat.stream()), "am_composite_fwd")
        ctx.save_for_backward(sdf, normal, rgb, inv_s, rays, ray_indices, t_starts, t_ends,
                              packed, trans, opacity, cn_raw)
        ctx.cdesc = cdesc
        return weights, opacity, depth, comp_rgb, comp_normal

    @staticmethod
    def backward(ctx, g_w, g_op, g_dep, g_rgb, g_cn):
        (sdf, normal, rgb, inv_s, rays, ridx, t0, t1, packed, trans, opacity,
         cn_raw) = ctx.saved_tensors
        lib = nat.load()
        n, R = sdf.shape[0], packed.shape[0]

        def opt(t):
            return _f32c(t) if t is not None else None

        gs = torch.empty_like(sdf)
        gn = torch.empty_like(normal)
        gr = torch.empty_like(rgb)
        gi = torch.empty(1, device=sdf.device, dtype=torch.float32)
        ws = nat.held_workspace("composite_bwd", lib.am_composite_bwd_workspace_size(R, n), sdf.device)
        nat.check(lib.am_composite_bwd(nat.sptr(ctx.cdesc), nat.ptr(sdf), nat.ptr(normal),
                                       nat.ptr(rays), nat.ptr(ridx), nat.ptr(t0), nat.ptr(t1),
                                       nat.ptr(rgb), nat.ptr(inv_s), nat.ptr(packed), R, n,
                                       nat.ptr(trans), nat.ptr(opacity), nat.ptr(cn_raw),
                                       nat.ptr(opt(g_w)), nat.ptr(opt(g_op)), nat.ptr(opt(g_dep)),
                                       nat.ptr(opt(g_rgb)), nat.ptr(opt(g_cn)), nat.ptr(gs),
                                       nat.ptr(gn), nat.ptr(gr), nat.ptr(gi), nat.ptr(ws),
                                       ws.numel(), nat.stream()), "am_composite_bwd")
        return gs, gn, gr, gi, None, None, None, None, None, None


def composite(sdf, normal, rgb, inv_s, rays, ray_indices, t_starts, t_ends, packed, cdesc):
    return CompositeFn.apply(sdf, normal, rgb, inv_s, rays, ray_indices, t_starts, t_ends, packed,
                             cdesc)


# ---- non-differentiable launches ------------------------------------------
def ray_aabb(rays, aabb):
    """nerfacc ray_aabb_intersect (t_min clipped to >= 0, 1e10 on a miss)."""
    import ctypes
    nat.require_cuda(rays)
    R = rays.shape[0]
    t_near = torch.empty(R, device=rays.device, dtype=torch.float32)
    t_far = torch.empty_like(t_near)
    box = (ctypes.c_float * 6)(*[float(v) for v in aabb])
    nat.check(nat.load().am_ray_aabb(nat.ptr(rays), R, ctypes.cast(box, ctypes.c_void_p),
                                     nat.ptr(t_near), nat.ptr(t_far), nat.stream()), "am_ray_aabb")
    return t_near, t_far


# Marching stages the count pass's samples (one march per ray); off: count
# and fill both march.  Without a per-ray cap the staging area is sized by
# the most samples a unit-direction ray can take in the box (its diagonal
# over the step); a batch in which some ray reaches that bound marches again
# in the unstaged form, so the samples never depend on it.
MARCH_STAGED = {"on": True}


def _implicit_stage_cap(roi, step):
    diag = math.sqrt(sum((float(roi[3 + d]) - float(roi[d])) ** 2 for d in range(3)))
    return int(math.ceil(diag / float(step) * 1.001)) + 4


def ray_marching(rays, t_near, t_far, roi, grid_binary, step, jitter=None, cap=0):
    """nerfacc ray_marching (models/neus.py:209-220): count -> scan -> fill.
    Returns (ray_indices int32 [N], t_starts [N], t_ends [N], packed [R,2]);
    one host read of N, as in the reference."""
    import ctypes
    lib = nat.load()
    nat.require_cuda(rays, t_near, t_far)
    R = rays.shape[0]
    dev = rays.device
    roi_h = (ctypes.c_float * 6)(*[float(v) for v in roi])
    roi_p = ctypes.cast(roi_h, ctypes.c_void_p)
    res = 0
    gptr = None
    if grid_binary is not None:
        res = int(grid_binary.shape[-1])
        gptr = nat.ptr(grid_binary)
    packed = torch.empty(R, 2, device=dev, dtype=torch.int32)
    total = torch.empty(1, device=dev, dtype=torch.int32)
    jit = _f32c(jitter) if jitter is not None else None
    stage_cap = int(cap) if cap > 0 else (_implicit_stage_cap(roi, step) if R > 0 else 0)
    if stage_cap > 0 and MARCH_STAGED["on"]:
        # one march per ray: the count pass stages the samples, the fill copies
        ws = nat.held_workspace("march_staged", lib.am_march_staged_workspace_size(R, stage_cap), dev)
        nat.check(lib.am_march_count_staged(nat.ptr(rays), R, nat.ptr(t_near), nat.ptr(t_far),
                                            nat.ptr(jit), roi_p, gptr, res, float(step),
                                            stage_cap, nat.ptr(packed), nat.ptr(total),
                                            nat.ptr(ws), ws.numel(), nat.stream()),
                  "am_march_count_staged")
        if cap > 0:
            N = int(total.item())
        else:   # the total and the largest count in one host read
            N, most = torch.stack([total[0], packed[:, 1].max()]).tolist()
            N = int(N)
            if most >= stage_cap:
                N = -1   # some ray may be longer than the bound: march unstaged
        if N >= 0:
            ridx = torch.empty(N, device=dev, dtype=torch.int32)
            t0 = torch.empty(N, device=dev, dtype=torch.float32)
            t1 = torch.empty(N, device=dev, dtype=torch.float32)
            if N > 0:
                nat.check(lib.am_march_fill_staged(R, stage_cap, nat.ptr(packed), nat.ptr(ws),
                                                   ws.numel(), nat.ptr(ridx), nat.ptr(t0),
                                                   nat.ptr(t1), nat.stream()),
                          "am_march_fill_staged")
            return ridx, t0, t1, packed
    ws = nat.held_workspace("march", lib.am_march_workspace_size(R), dev)
    nat.check(lib.am_march_count(nat.ptr(rays), R, nat.ptr(t_near), nat.ptr(t_far), nat.ptr(jit),
                                 roi_p, gptr, res, float(step), int(cap), nat.ptr(packed),
                                 nat.ptr(total), nat.ptr(ws), ws.numel(), nat.stream()),
              "am_march_count")
    N = int(total.item())
    ridx = torch.empty(N, device=dev, dtype=torch.int32)
    t0 = torch.empty(N, device=dev, dtype=torch.float32)
    t1 = torch.empty(N, device=dev, dtype=torch.float32)
    if N > 0:   # no ray has a sample: packed already holds zero counts
        nat.check(lib.am_march_fill(nat.ptr(rays), R, nat.ptr(t_near), nat.ptr(t_far),
                                    nat.ptr(jit), roi_p, gptr, res, float(step), int(cap),
                                    nat.ptr(packed), nat.ptr(ridx), nat.ptr(t0), nat.ptr(t1),
                                    nat.stream()), "am_march_fill")
    return ridx, t0, t1, packed


def ray_marching_bg(rays, t_near, t_far, roi, grid_binary, step, cone_angle, cap=0,
                    sigma_fn=None, early_stop_eps=1e-4):
    """nerfacc ray_marching for the background (models/neus.py:141-171):
    unbounded-sphere contraction over roi, cone-angle steps, [t_near, t_far]
    per ray given.  With sigma_fn(t_starts, t_ends, ray_indices) -> density,
    the samples behind transmittance < early_stop_eps are dropped (each ray
    keeps a prefix: am_march_bg_visible, then the fill pass again with the
    kept counts).  Returns (ray_indices int32, t_starts, t_ends, packed)."""
    cand = ray_marching_bg_candidates(rays, t_near, t_far, roi, grid_binary, step, cone_angle,
                                      cap)
    return ray_marching_bg_filter(cand, sigma_fn, early_stop_eps)


def _march_bg_ctx(rays, t_near, t_far, roi, grid_binary, step, cone_angle, cap):
    import ctypes
    nat.require_cuda(rays, t_near, t_far)
    roi_h = (ctypes.c_float * 6)(*[float(v) for v in roi])
    res, gptr = 0, None
    if grid_binary is not None:
        res = int(grid_binary.shape[-1])
        gptr = nat.ptr(grid_binary)
    return dict(rays=rays, t_near=_f32c(t_near), t_far=_f32c(t_far), roi_h=roi_h,
                roi_p=ctypes.cast(roi_h, ctypes.c_void_p), grid=grid_binary, gptr=gptr, res=res,
                step=float(step), cone=float(cone_angle), cap=int(cap))


def _march_bg_fill(c, pk, n):
    lib = nat.load()
    dev = c["rays"].device
    ridx = torch.empty(n, device=dev, dtype=torch.int32)
    t0 = torch.empty(n, device=dev, dtype=torch.float32)
    t1 = torch.empty(n, device=dev, dtype=torch.float32)
    if n > 0:
        nat.check(lib.am_march_bg_fill(nat.ptr(c["rays"]), c["rays"].shape[0], nat.ptr(c["t_near"]),
                                       nat.ptr(c["t_far"]), c["roi_p"], c["gptr"], c["res"],
                                       c["step"], c["cone"], c["cap"], nat.ptr(pk), nat.ptr(ridx),
                                       nat.ptr(t0), nat.ptr(t1), nat.stream()), "am_march_bg_fill")
    return ridx, t0, t1


def ray_marching_bg_candidates(rays, t_near, t_far, roi, grid_binary, step, cone_angle, cap=0):
    """The background marching before the early-stop filter (count, one
    host read of the sample count, fill): it reads the occupancy grid only,
    not the weights, so a pipelined training loop runs it ahead.  Returns the
    state ray_marching_bg_filter takes."""
    lib = nat.load()
    c = _march_bg_ctx(rays, t_near, t_far, roi, grid_binary, step, cone_angle, cap)
    R = rays.shape[0]
    dev = rays.device
    packed = torch.empty(R, 2, device=dev, dtype=torch.int32)
    total = torch.empty(1, device=dev, dtype=torch.int32)
    ws = nat.held_workspace("march_bg", lib.am_march_workspace_size(R), dev)
    nat.check(lib.am_march_bg_count(nat.ptr(rays), R, nat.ptr(c["t_near"]), nat.ptr(c["t_far"]),
                                    c["roi_p"], c["gptr"], c["res"], c["step"], c["cone"],
                                    c["cap"], nat.ptr(packed), nat.ptr(total), nat.ptr(ws),
                                    ws.numel(), nat.stream()), "am_march_bg_count")
    N = int(total.item())
    ridx, t0, t1 = _march_bg_fill(c, packed, N)
    c.update(ridx=ridx, t0=t0, t1=t1, packed=packed, n=N)
    return c


def ray_marching_bg_filter(c, sigma_fn=None, early_stop_eps=1e-4):
    """The early-stop filter of ray_marching_bg on candidates from
    ray_marching_bg_candidates (on the stream they were made on, or after
    it): each ray keeps the prefix before transmittance < early_stop_eps."""
    ridx, t0, t1, packed, N = c["ridx"], c["t0"], c["t1"], c["packed"], c["n"]
    if sigma_fn is None or early_stop_eps <= 0.0 or N == 0:
        return ridx, t0, t1, packed
    lib = nat.load()
    R = packed.shape[0]
    dev = packed.device
    sigma = _f32c(sigma_fn(t0, t1, ridx).reshape(-1))
    if sigma.shape[0] != N:
        raise ValueError(f"sigma_fn returned {sigma.shape[0]} values for {N} samples")
    kept = torch.empty_like(packed)
    total = torch.empty(1, device=dev, dtype=torch.int32)
    ws = nat.held_workspace("march_bg", lib.am_march_workspace_size(R), dev)
    nat.check(lib.am_march_bg_visible(nat.ptr(sigma), nat.ptr(t0), nat.ptr(t1), nat.ptr(packed), R,
                                      float(early_stop_eps), nat.ptr(kept), nat.ptr(total),
                                      nat.ptr(ws), ws.numel(), nat.stream()), "am_march_bg_visible")
    K = int(total.item())
    if K == N:
        return ridx, t0, t1, packed
    ridx, t0, t1 = _march_bg_fill(c, kept, K)
    return ridx, t0, t1, kept


class DensityCompositeFn(torch.autograd.Function):
    """nerfacc render_weight_from_density + accumulate_along_rays
    (models/neus.py:174-178) in one launch each way: weights, opacity,
    depth (midpoints), comp_rgb; gradients to density and rgb."""

    @staticmethod
    def forward(ctx, density, rgb, t_starts, t_ends, packed):
        nat.require_cuda(density, rgb, t_starts, t_ends, packed)
        lib = nat.load()
        density, rgb = _f32c(density.reshape(-1)), _f32c(rgb.reshape(-1, 3))
        n, R = density.shape[0], packed.shape[0]
        dev = density.device
        w = torch.empty(n, device=dev, dtype=torch.float32)
        trans = torch.empty(n, device=dev, dtype=torch.float32)
        op = torch.empty(R, device=dev, dtype=torch.float32)
        dep = torch.empty(R, device=dev, dtype=torch.float32)
        crgb = torch.empty(R, 3, device=dev, dtype=torch.float32)
        nat.check(lib.am_composite_density_fwd(nat.ptr(density), nat.ptr(t_starts), nat.ptr(t_ends),
                                                nat.ptr(rgb), nat.ptr(packed), R, n, nat.ptr(w),
                                                nat.ptr(trans), nat.ptr(op), nat.ptr(dep),
                                                nat.ptr(crgb), nat.stream()),
                  "am_composite_density_fwd")
        ctx.save_for_backward(density, rgb, t_starts, t_ends, packed, w, trans)
        return w, op, dep, crgb

    @staticmethod
    def backward(ctx, g_w, g_op, g_dep, g_rgb):
        density, rgb, t_starts, t_ends, packed, w, trans = ctx.saved_tensors
        lib = nat.load()
        n, R = density.shape[0], packed.shape[0]
        gd = torch.empty_like(density)
        gr = torch.empty_like(rgb)
        c = lambda t: _f32c(t) if t is not None else None
        g_w, g_op, g_dep, g_rgb = c(g_w), c(g_op), c(g_dep), c(g_rgb)
        nat.check(lib.am_composite_density_bwd(nat.ptr(density), nat.ptr(t_starts), nat.ptr(t_ends),
                                                nat.ptr(rgb), nat.ptr(packed), R, n, nat.ptr(w),
                                                nat.ptr(trans), nat.ptr(g_w), nat.ptr(g_op),
                                                nat.ptr(g_dep), nat.ptr(g_rgb), nat.ptr(gd),
                                                nat.ptr(gr), nat.stream()),
                  "am_composite_density_bwd")
        return gd, gr, None, None, None


def composite_density(density, rgb, t_starts, t_ends, packed):
    return DensityCompositeFn.apply(density, rgb, t_starts, t_ends, packed)


class ContractUnboundedFn(torch.autograd.Function):
    """contract_to_unisphere(x, r, UN_BOUNDED_SPHERE) (models/geometry.py:
    19-31) as VolumeDensity.forward applies it: one HIP launch forward, one
    backward (J^T g at the saved points)."""

    @staticmethod
    def forward(ctx, x, radius):
        nat.require_cuda(x)
        shape = x.shape
        xf = _f32c(x).reshape(-1, 3)
        out = torch.empty_like(xf)
        nat.check(nat.load().am_contract_unbounded(nat.ptr(xf), xf.shape[0], float(radius),
                                                   nat.ptr(out), nat.stream()),
                  "am_contract_unbounded")
        ctx.save_for_backward(xf)
        ctx.radius, ctx.shape, ctx.dtype = float(radius), shape, x.dtype
        return out.view(shape)

    @staticmethod
    def backward(ctx, g):
        (xf,) = ctx.saved_tensors
        gf = _f32c(g).reshape(-1, 3)
        gx = torch.empty_like(xf)
        nat.check(nat.load().am_contract_unbounded_bwd(nat.ptr(xf), nat.ptr(gf), xf.shape[0],
                                                       ctx.radius, nat.ptr(gx), nat.stream()),
                  "am_contract_unbounded_bwd")
        return gx.view(ctx.shape).to(ctx.dtype), None


def contract_unbounded(x, radius):
    if x.shape[-1] != 3:
        raise ValueError(f"points must be [..., 3] (got {tuple(x.shape)})")
    return ContractUnboundedFn.apply(x, radius)


def sample_points(rays, ray_indices, t_starts, t_ends):
    """models/neus.py:222-227: positions, midpoints, dists."""
    n = ray_indices.shape[0]
    dev = rays.device
    pts = torch.empty(n, 3, device=dev, dtype=torch.float32)
    mid = torch.empty(n, device=dev, dtype=torch.float32)
    dists = torch.empty(n, device=dev, dtype=torch.float32)
    nat.check(nat.load().am_sample_points(nat.ptr(rays), nat.ptr(ray_indices), nat.ptr(t_starts),
                                          nat.ptr(t_ends), n, nat.ptr(pts), nat.ptr(mid),
                                          nat.ptr(dists), nat.stream()), "am_sample_points")
    return pts, mid, dists
