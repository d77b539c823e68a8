"""GPU: the learned preconditioner's training harness (the counterpart of
nsrno/train_preconditioner.py:45-109) runs on the project's kernels.

* One Adam step (helm_adam_step) on the gradient that the BF16 MFMA backward
  pass (helm_net_backward) gives for the minimal-residual loss lowers that
  loss on the same batch.
* ``train_preconditioner.main`` runs end to end on a small grid for both of
  its losses (on-policy warm start, unrolled FGMRES subspace): finite losses,
  parameters written and loadable.
"""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_adam_step_lowers_the_minimal_residual_loss():
    from nsrno_amd.learned import LearnedNet, attach, canvas, init_params
    from nsrno_amd.train_preconditioner import Family, loss_and_grad, set_medium
    fam = Family(interior=128, frequency_mhz=3.0)
    plan = fam.plan(restart=8)
    set_medium(plan, fam.medium(5), dict(levels=3, shift=0.5, pre=2, post=2, coarse_sweeps=10))
    net = LearnedNet(canvas(fam.nz), canvas(fam.nx), fam.R, training=True, params=init_params(seed=1, out_scale=0.5))
    attach(plan, net)
    b = plan.inject(fam.ports, fam.amp).to(torch.complex64)
    r = (b / b.abs().square().mean(dim=(1, 2), keepdim=True).sqrt()).contiguous()
    l0, dout = loss_and_grad(plan, net, r, "angle")
    grad = net.backward(dout)
    assert torch.isfinite(grad).all() and float(grad.abs().max()) > 0
    net.adam(lr=2e-4)
    l1, _ = loss_and_grad(plan, net, r, "angle")
    print(f"minimal-residual loss before {l0:.5f}, after one Adam step {l1:.5f}")
    assert l1 < l0
    net.close()


@pytest.mark.parametrize("loss", ["warmstart", "subspace"])
def test_training_harness_end_to_end(tmp_path, loss, capsys):
    from nsrno_amd import train_preconditioner as tp
    from nsrno_amd.learned import LearnedNet
    out = tmp_path / "params.npz"
    args = ["--out", str(out), "--media", "2", "--steps", "4", "--interior", "96", "--loss", loss,
            "--batch", "4", "--unroll", "3", "--cycles", "1", "--log-every", "2"]
    tp.main(args)
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    train = [x for x in lines if x.get("phase") == "train"]
    assert train and all(np.isfinite(x["loss"]) for x in train)
    assert lines[-1]["phase"] == "done" and np.isfinite(lines[-1]["final_loss"])
    params = LearnedNet.load_params(out)
    assert params.size == 12 * (9 * 32 * 32 + 32) and np.isfinite(params).all()
