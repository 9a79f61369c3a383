"""The torch rollout (mif_amd.autograd.Rollout) with the inputs autograd
hands it in practice: an expanded (non-contiguous) output gradient, a
non-contiguous initial field, a body force on the GPU; and a graph whose
Rollout has run another forward since, which must raise instead of
replaying the wrong checkpoints."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

mif = pytest.importorskip("mif_amd")
torch = pytest.importorskip("torch")

from mif_amd.autograd import Rollout  # noqa: E402
from mif_amd.problems import channel_grid  # noqa: E402


def _setup(nsteps=4):
    gs = channel_grid(16, 12, 16, Re=100.0, dt=1e-3, steps=nsteps)
    ctx = mif.Context(gs, device=0)
    roll = Rollout(ctx, nsteps, checkpoint_every=2)
    rng = np.random.default_rng(3)
    u0 = [torch.tensor(0.1 * rng.standard_normal(s), device="cuda") for s in roll.shapes]
    return ctx, roll, u0


def _grad(roll, u0, f, loss):
    us = [t.clone().requires_grad_(True) for t in u0]
    ff = f.clone().requires_grad_(True)
    out = roll(us, ff)
    loss(out).backward()
    return [t.grad.detach().clone() for t in us], ff.grad.detach().clone()


def test_expanded_gradient_and_noncontiguous_input():
    print("[case] rollout inputs: expanded gradient, non-contiguous u0, CUDA force", flush=True)
    ctx, roll, u0 = _setup()
    f = torch.tensor([0.3, -0.2, 0.1], dtype=torch.float64)
    # out[0].sum() hands the rollout an expanded, non-contiguous gradient
    g_sum, gf_sum = _grad(roll, u0, f, lambda out: out[0].sum())
    g_ref, gf_ref = _grad(roll, u0, f, lambda out: (out[0] * torch.ones_like(out[0])).sum())
    for a, b in zip(g_sum, g_ref):
        assert torch.equal(a, b)
    assert torch.equal(gf_sum, gf_ref)
    # a non-contiguous view as the initial field (a transposed copy's transpose)
    u_nc = [t.transpose(0, 2).contiguous().transpose(0, 2) for t in u0]
    assert not u_nc[0].is_contiguous()
    g_nc, _ = _grad(roll, u_nc, f, lambda out: (out[0] * out[0]).sum())
    g_c, _ = _grad(roll, u0, f, lambda out: (out[0] * out[0]).sum())
    for a, b in zip(g_nc, g_c):
        assert torch.equal(a, b)
    # the force on the GPU: its gradient comes back on the GPU
    fc = f.to("cuda")
    _, gf_cuda = _grad(roll, u0, fc, lambda out: out[0].sum())
    assert gf_cuda.device.type == "cuda"
    assert torch.equal(gf_cuda.cpu(), gf_sum)
    ctx.close()


def test_second_forward_invalidates_the_first_graph():
    print("[case] rollout: two forwards, one backward", flush=True)
    ctx, roll, u0 = _setup()
    f = torch.tensor([0.3, -0.2, 0.1], dtype=torch.float64)
    us = [t.clone().requires_grad_(True) for t in u0]
    out_a = roll(us, f)
    out_b = roll([t * 1.5 for t in u0], torch.tensor([0.0, 0.0, 0.0], dtype=torch.float64))
    with pytest.raises(RuntimeError, match="another forward"):
        out_a[0].sum().backward()
    # the latest graph still differentiates, with its own body force
    us_b = [t.clone().requires_grad_(True) for t in u0]
    fb = torch.tensor([0.1, 0.0, -0.1], dtype=torch.float64)
    g1, _ = _grad(roll, u0, fb, lambda out: out[1].sum())
    roll(us_b, f)  # a forward with another force in between ...
    g2, _ = _grad(roll, u0, fb, lambda out: out[1].sum())  # ... does not leak into this one
    for a, b in zip(g1, g2):
        assert torch.equal(a, b)
    del out_b
    ctx.close()
