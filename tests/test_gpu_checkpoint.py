"""Checkpoint round trip (the reference trains under PyTorch Lightning, whose
checkpoint holds the module state_dict, the optimizer state and the AMP
scaler state).  A run is interrupted after K steps; its state is saved
through torch.save (model state_dict, the occupancy grid's buffers among
it; AdamW state; the loss scaler; the device RNG and the ray sampler's
generator, which the occupancy update and ray sampling draw from) and
restored into a freshly built workload, which then continues: every
parameter, every AdamW moment and the occupancy grid equal the
uninterrupted run's bit for bit.  The band mask of the progressive hash grid
is not state: update_step rebuilds it from the global step, as in the
reference.  The marched workloads cross an occupancy update (every 16th
global step) after the restore."""
import io

import pytest
import torch

pytestmark = pytest.mark.gpu


def _save(wl):
    buf = io.BytesIO()
    torch.save({"model": wl.model.state_dict(), "optimizer": wl.optimizer.state_dict(),
                "scaler": wl.scaler.state_dict() if wl.scaler else None,
                "cuda_rng": torch.cuda.get_rng_state(),
                "sample_gen": (wl.sample_gen.get_state() if hasattr(wl, "sample_gen")
                               else None)}, buf)
    buf.seek(0)
    return buf


def _load(wl, buf):
    st = torch.load(buf, weights_only=True)
    wl.model.load_state_dict(st["model"])
    wl.optimizer.load_state_dict(st["optimizer"])
    if wl.scaler is not None:
        wl.scaler.load_state_dict(st["scaler"])
    torch.cuda.set_rng_state(st["cuda_rng"])
    if st["sample_gen"] is not None:
        wl.sample_gen.set_state(st["sample_gen"])


def _state(wl):
    out = {k: v.detach().clone() for k, v in wl.model.state_dict().items()}
    for i, p in enumerate(wl.model.parameters()):
        st = wl.optimizer.state.get(p, {})
        for k in ("exp_avg", "exp_avg_sq"):
            if k in st:
                out[f"opt{i}.{k}"] = st[k].clone()
    return out


@pytest.mark.parametrize("name,rays,k,n", [("sphere-fp32", 1024, 3, 3),
                                           ("neuralangelo-fp32", 1024, 10, 8),
                                           ("neuralangelo-mixed", 1024, 10, 8)])
def test_checkpoint_round_trip_continues_bit_identically(dev, name, rays, k, n):
    from angelo_amd.workloads import Workload
    print(f"checkpoint round trip: {name}, {rays} rays, {k} + {n} steps", flush=True)
    torch.manual_seed(0)
    a = Workload(name, dev, rays=rays)
    for i in range(k):
        a.step(i)
    torch.cuda.synchronize()
    ckpt = _save(a)
    for i in range(k, k + n):
        a.step(i)
    torch.cuda.synchronize()
    ref = _state(a)
    if a.scaler is not None:
        ref_scale = a.scaler.state_dict()
    del a
    torch.manual_seed(1)   # the restored run must not depend on the fresh build's draws
    b = Workload(name, dev, rays=rays)
    _load(b, ckpt)
    for i in range(k, k + n):
        b.step(i)
    torch.cuda.synchronize()
    got = _state(b)
    assert got.keys() == ref.keys()
    assert any(key.startswith("opt") for key in got)
    for key in ref:
        assert torch.equal(got[key], ref[key]), key
    if b.scaler is not None:
        assert b.scaler.state_dict() == ref_scale
