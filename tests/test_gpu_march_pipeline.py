"""The pipelined training loop: the next step's rays are drawn and marched
on a second stream while the current step's backward and optimizer step
run; before a step that updates the occupancy grid the loop marches in
order.  The global generator's draws (march jitter, grid updates) keep their
order, so after a run that crosses a grid update (every 16th global step)
every parameter, AdamW moment and grid buffer equals the sequential loop's
bit for bit, and so do the sample counts of every step."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run(name, dev, rays, steps, pipeline):
    from angelo_amd.workloads import Workload
    wl = Workload(name, dev, rays=rays, pipeline=pipeline)
    counts = [wl.step(i) for i in range(steps)]
    wl.drain()
    torch.cuda.synchronize()
    state = {k: v.detach().clone() for k, v in wl.model.state_dict().items()}
    for i, p in enumerate(wl.model.parameters()):
        for k, v in wl.optimizer.state.get(p, {}).items():
            if torch.is_tensor(v) and v.numel() > 1:
                state[f"opt{i}.{k}"] = v.clone()
    return counts, state, wl.pipeline


@pytest.mark.parametrize("name,rays,steps", [("neuralangelo-fp32", 1024, 18),
                                             ("neuralangelo-mixed", 1024, 18),
                                             ("neuralangelo-bg", 1024, 18),
                                             ("neuralangelo-config", 1024, 18)])
def test_pipelined_loop_equals_sequential(dev, name, rays, steps):
    print(f"pipelined loop: {name}, {rays} rays, {steps} steps", flush=True)
    c_seq, s_seq, p_seq = _run(name, dev, rays, steps, False)
    c_pip, s_pip, p_pip = _run(name, dev, rays, steps, True)
    assert not p_seq and p_pip
    assert c_pip == c_seq
    assert s_pip.keys() == s_seq.keys()
    assert any(k.startswith("opt") for k in s_pip)
    for k in s_seq:
        assert torch.equal(s_pip[k], s_seq[k]), k
