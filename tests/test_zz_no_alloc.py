"""am_composite_bwd_noalloc (the compositing backward without a workspace,
its scratch a static device array of the library) against am_composite_bwd
with a caller workspace: bit-identical gradients on two calls in a row, the
second with fewer rays and samples (the static scratch keeps the first
call's contents), and the size limit rejected before any launch."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(R, spr, seed, dev):
    g = torch.Generator().manual_seed(seed)
    counts = torch.randint(0, spr + 1, (R,), generator=g)
    counts[0] = 0                                   # a ray without samples
    n = int(counts.sum())
    ridx = torch.repeat_interleave(torch.arange(R), counts).to(torch.int32)
    starts = torch.cumsum(counts, 0) - counts
    packed = torch.stack([starts, counts], -1).to(torch.int32)
    d = torch.randn(R, 3, generator=g)
    rays = torch.cat([torch.randn(R, 3, generator=g), d / d.norm(dim=-1, keepdim=True)], -1)
    t0 = torch.rand(n, generator=g)
    t1 = t0 + 0.01
    nrm = torch.randn(n, 3, generator=g)
    f32 = lambda t: t.to(dev, torch.float32).contiguous()
    return dict(R=R, n=n, rays=f32(rays), ridx=ridx.to(dev), t0=f32(t0), t1=f32(t1),
                sdf=f32(torch.randn(n, generator=g) * 0.05),
                normal=f32(nrm / nrm.norm(dim=-1, keepdim=True)),
                rgb=f32(torch.rand(n, 3, generator=g)), packed=packed.to(dev),
                inv_s=f32(torch.tensor([30.0])),
                g_op=f32(torch.randn(R, generator=g)), g_dep=f32(torch.randn(R, generator=g)),
                g_rgb=f32(torch.randn(R, 3, generator=g)), g_cn=f32(torch.randn(R, 3, generator=g)),
                g_w=f32(torch.randn(n, generator=g)))


def _run(p, noalloc, dev):
    from angelo_amd import _native as nat
    lib = nat.load()
    cd = nat.CompositeDesc()
    cd.cos_anneal_ratio = 0.4
    cd.kahan = 1
    R, n = p["R"], p["n"]
    f32 = dict(device=dev, dtype=torch.float32)
    w, trans = torch.empty(n, **f32), torch.empty(n, **f32)
    op, dep = torch.empty(R, **f32), torch.empty(R, **f32)
    crgb, cn, cn_raw = (torch.empty(R, 3, **f32) for _ in range(3))
    P = nat.ptr
    nat.check(lib.am_composite_fwd(nat.sptr(cd), P(p["sdf"]), P(p["normal"]), P(p["rays"]),
                                   P(p["ridx"]), P(p["t0"]), P(p["t1"]), P(p["rgb"]),
                                   P(p["inv_s"]), P(p["packed"]), R, n, P(w), P(trans), P(op),
                                   P(dep), P(crgb), P(cn), P(cn_raw), nat.stream()),
              "am_composite_fwd")
    gs, gn, gr = torch.full((n,), 7.0, **f32), torch.full((n, 3), 7.0, **f32), \
        torch.full((n, 3), 7.0, **f32)
    gi = torch.full((1,), 7.0, **f32)
    common = (nat.sptr(cd), P(p["sdf"]), P(p["normal"]), P(p["rays"]), P(p["ridx"]), P(p["t0"]),
              P(p["t1"]), P(p["rgb"]), P(p["inv_s"]), P(p["packed"]), R, n, P(trans), P(op),
              P(cn_raw), P(p["g_w"]), P(p["g_op"]), P(p["g_dep"]), P(p["g_rgb"]), P(p["g_cn"]),
              P(gs), P(gn), P(gr), P(gi))
    if noalloc:
        nat.check(lib.am_composite_bwd_noalloc(*common, nat.stream()), "am_composite_bwd_noalloc")
    else:
        ws = nat.workspace(lib.am_composite_bwd_workspace_size(R, n), dev)
        nat.check(lib.am_composite_bwd(*common, P(ws), ws.numel(), nat.stream()),
                  "am_composite_bwd")
    torch.cuda.synchronize()
    return gs, gn, gr, gi


def test_composite_bwd_noalloc_matches_workspace_form(dev):
    print("no-alloc compositing backward: 6000 rays, then 900 rays", flush=True)
    for R, spr, seed in ((6000, 96, 1), (900, 40, 2)):
        p = _problem(R, spr, seed, dev)
        ref = _run(p, False, dev)
        got = _run(p, True, dev)
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        assert torch.isfinite(got[3]).all() and got[3].item() != 7.0


def test_composite_bwd_noalloc_rejects_oversized_batch(dev):
    from angelo_amd import _native as nat
    lib = nat.load()
    cd = nat.CompositeDesc()
    dummy = torch.zeros(16, device=dev)
    P = nat.ptr(dummy)
    rc = lib.am_composite_bwd_noalloc(nat.sptr(cd), *([P] * 9), (1 << 20) + 1, 16,
                                      *([P] * 12), nat.stream())
    assert rc == -1
    assert b"at most 1048576 rays" in lib.am_last_error()
