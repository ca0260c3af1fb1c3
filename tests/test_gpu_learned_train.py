"""Training kernels of the learned preconditioner against the PyTorch
restatement on the CPU: the trainer's forward pass, the loss and every layer's
weight gradient (BF16 roundings straight-through on both sides), and that Adam
steps lower the loss."""
import numpy as np
import pytest
import torch

from ariel_amd import FlipSim, _abi, learned, scenes
from ariel_amd.learned import net as lnet
from ariel_amd.learned.trainer import Trainer

pytestmark = pytest.mark.gpu


def _batch(n, seeds):
    """Pairs on the matrix of a stepped dam break: random FLUID residuals r and
    e = r / (N^2 diag), an approximate solution (the loss needs r.e > 0 only)."""
    sim = FlipSim(scenes.dam_break(n))
    sim.Init()
    sim.Step(2)
    flags, L = sim.get_flags(), sim.get_field(_abi.FIELD_L)
    sim.close()
    diag, fluid = lnet.laplacian(torch.from_numpy(flags.reshape(n, n, n)),
                                 torch.from_numpy(L.reshape(n, n, n)))
    rs, es = [], []
    for s in seeds:
        g = torch.Generator().manual_seed(s)
        r = torch.randn(n, n, n, generator=g) * fluid
        e = torch.where(fluid > 0, r / (diag * n * n).clamp(min=1e-6), torch.zeros_like(r))
        rs.append(r)
        es.append(e)
    sh = (len(seeds), 1, n, n, n)
    rep = lambda t: t.expand(len(seeds), n, n, n).reshape(sh).contiguous()
    return torch.stack(rs).reshape(sh), torch.stack(es).reshape(sh), rep(diag), rep(fluid)


def _cuda(*ts):
    return [t.to("cuda:0").contiguous() for t in ts]


@pytest.mark.parametrize("n", [16, 32])
def test_gradients_match_torch(n):
    print(f"trainer gradients n={n}", flush=True)
    r, e, diag, fluid = _batch(n, [1, 2])
    net = lnet.PressureNet(seed=5)
    flat = learned.layers_to_flat([p.detach().numpy() for p in net.w])
    tr = Trainer(flat)
    rg, eg, dg, fg = _cuda(r, e, diag, fluid)
    # forward of the trainer (materialised skip sums) against the BF16 emulation
    z = torch.empty_like(rg)
    tr.forward(rg, dg, fg, z)
    dw = lnet.diag_weight(diag, fluid)
    with torch.no_grad():
        zr = net(r, fluid, dw, emulate_bf16=True)
    rel = float((z.cpu() - zr).norm() / zr.norm())
    print(f"  forward rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-2
    loss_g, grad_g = tr.step(rg, eg, dg, fg, update=False, want_grad=True)
    zt = net(r, fluid, dw, emulate_bf16="ste")
    loss = lnet.a_angle_loss(zt, r, e, diag, fluid)
    loss.backward()
    grad_t = learned.layers_to_flat([p.grad.numpy() for p in net.w])
    print(f"  loss {loss_g:.6f} vs {loss.item():.6f}", flush=True)
    assert abs(loss_g - loss.item()) <= 2e-3 * max(1.0, abs(loss.item()))
    off, total = learned.layout()
    bounds = list(off) + [total]
    for i, (name, *_r) in enumerate(learned.LAYERS):
        a, b = grad_g[bounds[i]:bounds[i + 1]].astype(np.float64), grad_t[bounds[i]:bounds[i + 1]].astype(np.float64)
        nb = np.linalg.norm(b)
        cos = float(a @ b / max(np.linalg.norm(a) * nb, 1e-300))
        rel = float(np.linalg.norm(a - b) / max(nb, 1e-300))
        print(f"  {name:4s} |g| {nb:.3e} cos {cos:.6f} rel {rel:.2e}", flush=True)
        assert cos > 0.999 and rel < 5e-2, name


def test_adam_steps_lower_the_loss():
    n = 32
    r, e, diag, fluid = _cuda(*_batch(n, [3, 4]))
    tr = Trainer(lnet.PressureNet(seed=6).blob())
    first = tr.step(r, e, diag, fluid, update=False)
    for _ in range(30):
        tr.step(r, e, diag, fluid, lr=2e-3)
    last = tr.step(r, e, diag, fluid, update=False)
    print(f"trainer Adam: loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.8 * first
