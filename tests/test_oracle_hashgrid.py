"""Pins of the hash-grid oracle: level tables, closed-form indices, weights."""
import numpy as np
import pytest
import torch

from oracle import hashgrid as ohg

S1 = 1.3195079107728942


def test_level_table_benchmark_configs():
    lt = ohg.level_table(8, 16, 1.5, 15)
    assert lt.res.tolist() == [16, 24, 36, 54, 81, 122, 183, 274]
    assert lt.hashed.tolist() == [0, 0, 1, 1, 1, 1, 1, 1]
    assert lt.size.tolist()[:2] == [4096, 13824] and lt.n_entries == 214528
    lt = ohg.level_table(16, 32, S1, 21)
    assert lt.res.tolist() == [32, 43, 56, 74, 98, 129, 169, 223, 295, 389, 513, 676, 892, 1177,
                               1553, 2049]
    assert lt.hashed.tolist() == [0] * 5 + [1] * 11
    assert lt.size.tolist()[:5] == [32768, 79512, 175616, 405224, 941192]
    assert lt.n_entries == 24702984


def test_level_table_closed_form():
    for L, base, s, log2T in [(8, 16, 1.5, 15), (16, 32, S1, 19), (12, 8, 2.0, 12)]:
        lt = ohg.level_table(L, base, s, log2T)
        assert lt.res[0] == base and lt.scale[0] == base - 1
        off = 0
        for l in range(L):
            r = int(lt.res[l])
            assert r == int(np.ceil(lt.scale[l])) + 1
            dense = min(((r ** 3 + 7) // 8) * 8, 2 ** log2T)
            assert lt.size[l] == dense
            assert lt.hashed[l] == (1 if r ** 3 > dense else 0)
            assert lt.offset[l] == off
            off += dense
        assert lt.n_entries == off
        # scales grow by the per-level factor in real arithmetic (fp32 rounding aside)
        ratio = (lt.scale[1:] + 1) / (lt.scale[:-1] + 1)
        np.testing.assert_allclose(ratio, s, rtol=1e-5)


def test_indices_match_integer_formulas():
    lt = ohg.level_table(8, 16, 1.5, 15)
    rng = np.random.default_rng(0)
    x = rng.random((300, 3)).astype(np.float32)
    x[:4] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1, 0, 1]]
    cell, frac, idx, w = ohg.corners(x, lt)
    for i in range(x.shape[0]):
        for l in range(lt.n_levels):
            s = np.float32(lt.scale[l])
            for d in range(3):
                p = float(s) * float(x[i, d]) + 0.5          # exact in float64 here
                assert cell[i, l, d] == np.uint32(np.int64(np.floor(np.float32(p))) & 0xFFFFFFFF)
            for c in range(8):
                corner = [int(cell[i, l, d]) + ((c >> d) & 1) for d in range(3)]
                if lt.hashed[l]:
                    ref = ohg.hash_row(corner) % int(lt.size[l])
                else:
                    ref = (ohg.dense_row(corner, lt.res[l]) & 0xFFFFFFFF) % int(lt.size[l])
                assert idx[i, l, c] == ref


def test_weights_partition_of_unity():
    lt = ohg.level_table(16, 32, S1, 16)
    x = np.random.default_rng(1).random((500, 3)).astype(np.float32)
    _, frac, _, w = ohg.corners(x, lt)
    np.testing.assert_allclose(w.sum(-1), 1.0, atol=2e-6)
    assert (frac >= 0).all() and (frac < 1).all()


def test_linear_table_is_reproduced():
    """A table linear in the vertex coordinates is reproduced exactly by
    trilinear interpolation: enc = a . (scale*x + 0.5) + b on dense levels."""
    lt = ohg.level_table(2, 16, 1.5, 15)      # both levels dense
    a = np.array([0.3, -0.2, 0.7])
    b0 = 0.1
    params = np.zeros((lt.n_entries, 2), np.float32)
    for l in range(2):
        r = int(lt.res[l])
        ii = np.arange(r ** 3)
        coords = np.stack([ii % r, (ii // r) % r, ii // (r * r)], -1)
        params[lt.offset[l] + ii, 0] = coords @ a + b0
        params[lt.offset[l] + ii, 1] = -(coords @ a)
    rng = np.random.default_rng(2)
    x = (rng.random((200, 3)) * 0.9).astype(np.float32)   # no wrap-around at the top
    enc = ohg.encode_exact(x, params, lt)
    for l in range(2):
        pos = lt.scale[l].astype(np.float64) * x.astype(np.float64) + 0.5
        np.testing.assert_allclose(enc[:, 2 * l], pos @ a + b0, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(enc[:, 2 * l + 1], -(pos @ a), rtol=1e-5, atol=1e-4)


def test_torch_encode_matches_exact_and_masks():
    lt = ohg.level_table(8, 16, 1.5, 15)
    rng = np.random.default_rng(3)
    x = rng.random((100, 3)).astype(np.float32)
    params = (rng.random((lt.n_entries, 2)) * 2 - 1).astype(np.float32)
    e = ohg.encode_exact(x, params, lt)
    t = ohg.encode_torch(torch.from_numpy(x), torch.from_numpy(params), lt).numpy()
    np.testing.assert_allclose(t, e, atol=2e-7)
    t64 = ohg.encode_torch(torch.from_numpy(x).double(), torch.from_numpy(params).double(), lt,
                           mode="f64").numpy()
    # fp32 positions round to ulp(scale*x) (~3e-5 at scale 272): float64
    # positions move the fraction by up to half of that
    np.testing.assert_allclose(t64, e, atol=4e-5)
    m = ohg.encode_torch(torch.from_numpy(x), torch.from_numpy(params), lt, active_levels=3).numpy()
    assert (m[:, 6:] == 0).all() and np.allclose(m[:, :6], e[:, :6], atol=2e-7)


def test_torch_encode_gradcheck():
    lt = ohg.level_table(3, 4, 1.5, 6)
    g = torch.Generator().manual_seed(0)
    x = (torch.rand(5, 3, generator=g, dtype=torch.float64) * 0.8 + 0.1).requires_grad_(True)
    p = torch.rand(lt.n_entries, 2, generator=g, dtype=torch.float64).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda xx, pp: ohg.encode_torch(xx, pp, lt, mode="f64"),
                                    (x, p), eps=1e-7, atol=1e-5)
