"""Hash-grid oracle -- test infrastructure only.

Restates the tiny-cuda-nn HashGrid that ProgressiveBandHashGrid wraps
(models/network_utils.py:40-66; built by get_encoding :83-93).  tcnn is not
vendored in the reference; its published algorithm is restated:

  level table   scale_l = 2^(l*log2(s))*base - 1 (fp32, exp2 correctly
                rounded), res_l = ceil(scale_l)+1, rows_l = min(next_mult8(
                res_l^3), 2^log2_T), hashed iff res_l^3 > rows_l.
  encode        pos = fmaf(scale_l, x, 0.5); cell = floor(pos); frac = pos-cell;
                corner c: cell + (c&1, c>>1&1, c>>2&1); dense row x+y*r+z*r^2,
                hashed row x ^ y*2654435761 ^ z*805459861 (uint32), mod rows_l;
                weight = prod over x,y,z of frac or 1-frac; output column l*F+f
                (level-major, proven by the mask layout at :53-54,58).
  mask          enc * mask, mask[:current_level*F] = 1 (:53-54,58,66).

Pins (tests/test_oracle_hashgrid.py): the level tables quoted for the
benchmark configurations; dense index and prime hash recomputed with Python
integers; partition of unity of the weights; a table that is linear in the
vertex coordinates is reproduced exactly by trilinear interpolation; the
torch encode agrees with the exact C encode.  Not pinned: NVIDIA's own fp32
exp2f/log2f (the table is explicit and can be overridden).
"""
from dataclasses import dataclass

import numpy as np
import torch

from . import clib


@dataclass
class LevelTable:
    n_levels: int
    n_features: int
    scale: np.ndarray   # float32 [L]
    res: np.ndarray     # uint32 [L]
    offset: np.ndarray  # uint32 [L]
    size: np.ndarray    # uint32 [L]
    hashed: np.ndarray  # uint32 [L]
    n_entries: int


def level_table(n_levels, base_resolution, per_level_scale, log2_hashmap_size, n_features=2):
    lib = clib.load()
    sc = np.zeros(32, np.float32)
    r, off, sz, h = (np.zeros(32, np.uint32) for _ in range(4))
    tot = lib.or_level_table(n_levels, base_resolution, float(np.float32(per_level_scale)),
                             log2_hashmap_size, clib.p(sc), clib.p(r), clib.p(off), clib.p(sz),
                             clib.p(h))
    if tot < 0:
        raise ValueError("bad level table arguments")
    L = n_levels
    return LevelTable(L, n_features, sc[:L].copy(), r[:L].copy(), off[:L].copy(), sz[:L].copy(),
                      h[:L].copy(), int(tot))


def _table_args(lt):
    return (np.ascontiguousarray(lt.scale), np.ascontiguousarray(lt.res),
            np.ascontiguousarray(lt.offset), np.ascontiguousarray(lt.size),
            np.ascontiguousarray(lt.hashed))


def corners(x, lt):
    """Exact per (point, level): cell [n,L,3], frac [n,L,3], level-local rows
    [n,L,8] and weights [n,L,8] (C, fp32 with fmaf)."""
    x = clib.f32(np.asarray(x).reshape(-1, 3))
    n, L = x.shape[0], lt.n_levels
    cell = np.zeros((n, L, 3), np.uint32)
    frac = np.zeros((n, L, 3), np.float32)
    idx = np.zeros((n, L, 8), np.uint32)
    w = np.zeros((n, L, 8), np.float32)
    sc, r, off, sz, h = _table_args(lt)
    clib.load().or_hashgrid_corners(clib.p(x), n, L, clib.p(sc), clib.p(r), clib.p(sz), clib.p(h),
                                    clib.p(cell), clib.p(frac), clib.p(idx), clib.p(w))
    return cell, frac, idx, w


def encode_exact(x, params, lt, active_levels=None):
    """tcnn-order fp32 encode (fma blend in corner order) -> [n, L*F] float32."""
    x = clib.f32(np.asarray(x).reshape(-1, 3))
    params = clib.f32(np.asarray(params).reshape(-1))
    n, L, F = x.shape[0], lt.n_levels, lt.n_features
    out = np.zeros((n, L * F), np.float32)
    sc, r, off, sz, h = _table_args(lt)
    clib.load().or_hashgrid_fwd(clib.p(x), n, clib.p(params), L, F, clib.p(sc), clib.p(r),
                                clib.p(off), clib.p(sz), clib.p(h), clib.p(out))
    if active_levels is not None:
        out[:, active_levels * F:] = 0.0
    return out


def _rows_torch(cell, lt):
    """Level-local rows of the 8 corners from integer cells [n,L,3] (int64),
    uint32 wrap emulated with masks."""
    m = 0xFFFFFFFF
    out = torch.zeros(cell.shape[0], lt.n_levels, 8, dtype=torch.int64, device=cell.device)
    for l in range(lt.n_levels):
        res, size = int(lt.res[l]), int(lt.size[l])
        for c in range(8):
            cx = (cell[:, l, 0] + (c & 1)) & m
            cy = (cell[:, l, 1] + ((c >> 1) & 1)) & m
            cz = (cell[:, l, 2] + ((c >> 2) & 1)) & m
            if lt.hashed[l]:
                idx = cx ^ ((cy * 2654435761) & m) ^ ((cz * 805459861) & m)
            else:
                idx = (cx + cy * res + cz * res * res) & m
            out[:, l, c] = idx % size
    return out


def encode_torch(x, params, lt, active_levels=None, mode="f32"):
    """Differentiable encode (autograd w.r.t. params [n_entries, F] and x).

    mode 'f32': rows and fractional positions from the exact C restatement
    (fp32 fmaf), the fraction carrying d frac / d x = scale_l.  mode 'f64':
    positions, cells and fractions computed in float64 (error budgets).
    Levels at or above active_levels are multiplied by zero (the mask)."""
    xs = x.reshape(-1, 3)
    n, L, F = xs.shape[0], lt.n_levels, lt.n_features
    dev = xs.device
    scale = torch.from_numpy(lt.scale.astype(np.float64)).to(dev, xs.dtype)   # [L]
    if mode == "f32":
        _, frac, idx, _ = corners(xs.detach().cpu().numpy(), lt)
        frac_t = torch.from_numpy(frac).to(dev, xs.dtype)                # [n,L,3]
        frac_t = frac_t + scale[None, :, None] * (xs - xs.detach())[:, None, :]
        idx = torch.from_numpy(idx.astype(np.int64)).to(dev)
    else:
        pos = scale[None, :, None] * xs[:, None, :] + 0.5
        cell = torch.floor(pos.detach())
        frac_t = pos - cell
        idx = _rows_torch(cell.to(torch.int64), lt)
    rows = idx + torch.from_numpy(lt.offset.astype(np.int64)).to(dev)[None, :, None]
    one = torch.ones((), dtype=xs.dtype, device=dev)
    outs = []
    for l in range(L):
        acc = torch.zeros(n, F, dtype=params.dtype, device=dev)
        for c in range(8):
            w = one
            for d in range(3):
                fd = frac_t[:, l, d]
                w = w * (fd if (c >> d) & 1 else 1.0 - fd)
            acc = acc + w[:, None].to(params.dtype) * params[rows[:, l, c]]
        outs.append(acc)
    out = torch.cat(outs, dim=-1)
    if active_levels is not None and active_levels < L:
        mask = torch.zeros(L * F, dtype=out.dtype, device=dev)
        mask[: active_levels * F] = 1.0
        out = out * mask
    return out.reshape(*x.shape[:-1], L * F)


def dense_row(c, res):
    """Python-integer dense row (the closed form)."""
    return int(c[0]) + int(c[1]) * int(res) + int(c[2]) * int(res) * int(res)


def hash_row(c):
    """Python-integer prime hash of a corner, uint32 wrap."""
    m = 0xFFFFFFFF
    return (int(c[0]) ^ ((int(c[1]) * 2654435761) & m) ^ ((int(c[2]) * 805459861) & m)) & m
