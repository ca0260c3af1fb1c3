"""PyTorch restatement of the learned preconditioner (training and checking).

Same network as the library's MFMA kernels: 4 levels, channels 16/32/48/64,
3x3x3 convolutions (stride 2 going down), 1x1x1 projections plus
nearest-neighbour upsampling and additive skips going up, ReLU, no biases:

  e0 = relu(conv(x)), x = (w r/s, w)
  e1..e3 = relu(conv_s2(.))
  d2 = relu(conv(e2 + up(p2(e3)))), d1, d0 likewise
  z  = s * w * conv(d0)

w = sqrt(6 / diag) on FLUID cells and 0 elsewhere (symmetric diagonal scaling;
it also marks the FLUID cells); s = sqrt(sum_FLUID r^2 / (8 * active
segments)), where an active segment is a run of 8 z-consecutive cells holding
at least one FLUID cell. With
``emulate_bf16`` every activation is rounded to BF16 where the kernels store
BF16, so the result matches the library to accumulation-order rounding.

Grid tensors are (batch, channel, i, j, k), k fastest, as the library's cells.
"""
import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import LAYERS, pack, unpack

FLUID, AIR, SOLID = 1, 0, 2


def _bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


def _bf16_ste(x):
    """BF16 rounding in the forward pass, identity in the backward pass (the
    library's training kernels propagate gradients through the roundings)."""
    return x + (_bf16(x) - x).detach()


class PressureNet(nn.Module):
    def __init__(self, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        ws = []
        for name, taps, _s, cin, cout, relu in LAYERS:
            k = 3 if taps == 27 else 1
            fan_in = cin * k ** 3
            std = (2.0 / fan_in) ** 0.5 if relu else (1.0 / fan_in) ** 0.5
            ws.append(nn.Parameter(torch.randn(cout, cin, k, k, k, generator=g) * std))
        self.w = nn.ParameterList(ws)

    @staticmethod
    def scale(r, fluid):
        """The kernels' input scale s per batch item (float32 of an FP64 sum)."""
        b = r.shape[0]
        f = fluid.reshape(b, -1, 8)
        nseg = (f > 0).any(dim=2).sum(dim=1).clamp(min=1).to(torch.float64)
        r2 = ((r.to(torch.float64) * fluid) ** 2).reshape(b, -1).sum(dim=1)
        return torch.sqrt(r2 / (8.0 * nseg)).to(torch.float32)

    def forward(self, r, fluid, dw, emulate_bf16=False):
        """r, fluid, dw = diag_weight(...): (B, 1, N, N, N) float32; returns z.
        emulate_bf16: False (FP32), True (BF16 roundings as the kernels store) or
        "ste" (the same roundings, straight-through in the backward pass)."""
        q = (_bf16_ste if emulate_bf16 == "ste" else _bf16) if emulate_bf16 else (lambda t: t)
        w = [q(p) for p in self.w]
        s = self.scale(r, fluid).view(-1, 1, 1, 1, 1)
        inv = torch.where(s > 0, 1.0 / s, torch.zeros_like(s))
        x = q(torch.cat([r * inv * dw, dw], dim=1))
        e0 = q(F.relu(F.conv3d(x, w[0], padding=1)))
        e1 = q(F.relu(F.conv3d(e0, w[1], stride=2, padding=1)))
        e2 = q(F.relu(F.conv3d(e1, w[2], stride=2, padding=1)))
        e3 = q(F.relu(F.conv3d(e2, w[3], stride=2, padding=1)))

        def up(skip, coarse, wp, wd):
            p = q(F.conv3d(coarse, wp))
            t = q(skip + F.interpolate(p, scale_factor=2, mode="nearest"))
            return q(F.relu(F.conv3d(t, wd, padding=1)))

        d2 = up(e2, e3, w[4], w[5])
        d1 = up(e1, d2, w[6], w[7])
        d0 = up(e0, d1, w[8], w[9])
        return F.conv3d(d0, w[10], padding=1) * s * dw

    def blob(self):
        return pack([p.detach().cpu().numpy() for p in self.w])

    @classmethod
    def from_blob(cls, blob):
        net = cls()
        with torch.no_grad():
            for p, w in zip(net.w, unpack(blob)):
                p.copy_(torch.from_numpy(w))
        return net


def laplacian(flags, L):
    """Coefficients of the pressure matrix for one grid: (diag, fluid mask).

    Follows the library's operator (the reference's ADiag/ARef,
    src/fluidCore/solver/solver.inl:59-101): off-diagonals -1 between FLUID
    neighbours; diag = 6 - #(SOLID or out-of-grid neighbours)
    - sum over AIR neighbours q of L[q] / min(1e-6, L[c]).
    flags: uint8 (N,N,N) tensor, L: float32 (N,N,N)."""
    fl = flags.to(torch.int64)
    padf = F.pad(fl.view(1, 1, *fl.shape).float(), (1, 1, 1, 1, 1, 1), value=float(SOLID))[0, 0]
    padL = F.pad(L.view(1, 1, *L.shape), (1, 1, 1, 1, 1, 1), value=0.0)[0, 0]
    n0, n1, n2 = fl.shape
    denom = torch.clamp(L, max=1.0e-6)
    diag = torch.full_like(L, 6.0)
    for d in range(3):
        for sgn in (-1, 1):
            sl = [slice(1, n0 + 1), slice(1, n1 + 1), slice(1, n2 + 1)]
            sl[d] = slice(1 + sgn, [n0, n1, n2][d] + 1 + sgn)
            nf = padf[tuple(sl)]
            nL = padL[tuple(sl)]
            diag = diag - (nf == SOLID).float()
            diag = diag - torch.where(nf == AIR, nL / denom, torch.zeros_like(L))
    fluid = (fl == FLUID).float()
    return diag * fluid, fluid


def diag_weight(diag, fluid):
    """sqrt(6 / diag) on FLUID cells with diag > 0, else 0 (float32, IEEE)."""
    ok = (fluid > 0) & (diag > 0)
    return torch.where(ok, torch.sqrt(6.0 / torch.where(ok, diag, torch.ones_like(diag))),
                       torch.zeros_like(diag))


def apply_laplacian(x, diag, fluid):
    """y = A x on FLUID rows (zero elsewhere), with the library's grid scaling
    N^2 (unit box, spacing 1/N); x, diag, fluid: (B, 1, N, N, N)."""
    xf = x * fluid
    p = F.pad(xf, (1, 1, 1, 1, 1, 1))
    nb = (p[:, :, :-2, 1:-1, 1:-1] + p[:, :, 2:, 1:-1, 1:-1] + p[:, :, 1:-1, :-2, 1:-1]
          + p[:, :, 1:-1, 2:, 1:-1] + p[:, :, 1:-1, 1:-1, :-2] + p[:, :, 1:-1, 1:-1, 2:])
    n = x.shape[-1]
    return (diag * xf - nb) * fluid * float(n * n)


def a_angle_loss(z, r, x, diag, fluid):
    """sin^2 of the A-angle between z and the exact solution x of A x = r,
    signed: 1 - (r.z)|r.z| / ((r.x)(z.Az)), invariant to positive scalings of z
    and above 1 when z points away from x (flexible PCG needs z.r > 0)."""
    b = z.shape[0]
    rz = (r * z).reshape(b, -1).sum(1)
    rx = (r * x).reshape(b, -1).sum(1).clamp(min=1e-30)
    zaz = (z * apply_laplacian(z, diag, fluid)).reshape(b, -1).sum(1).clamp(min=1e-30)
    return (1.0 - rz * rz.abs() / (rx * zaz)).mean()


def numpy_scale(r, flags):
    """float32 s of the kernels for one grid (numpy), for tests."""
    f = (np.asarray(flags).reshape(-1, 8) == FLUID)
    nseg = max(int(f.any(axis=1).sum()), 1)
    rr = np.asarray(r, np.float64).reshape(-1)[f.reshape(-1)]
    return np.float32(np.sqrt(np.dot(rr, rr) / (8.0 * nseg)))
