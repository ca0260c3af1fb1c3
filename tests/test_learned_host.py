"""Host-side pins of the learned preconditioner's tooling (CPU only): the weight
blob round trip, the PyTorch pressure operator against the oracle's operator,
the network's input scale and its positive homogeneity."""
import numpy as np
import pytest
import torch

from ariel_amd import _abi, learned, scenes
from ariel_amd.learned import net as lnet
from oracle import OracleSim


def test_blob_round_trip_and_layout():
    off, total = learned.layout()
    assert total == 248064 and off[-1] == total - 27 * 16 * 16
    m = lnet.PressureNet(seed=3)
    blob = m.blob()
    assert len(blob) == 24 + 2 * total and blob[:8] == b"ARIELNN1"
    back = learned.unpack(blob)
    for p, w in zip(m.w, back):
        ref = p.detach().to(torch.bfloat16).to(torch.float32).numpy()
        assert np.array_equal(ref, w)
    # packing is a pure relayout: packing the BF16 values again gives the same bytes
    assert learned.pack(back) == blob


def test_pack_places_taps_where_the_kernels_read_them():
    """A single nonzero weight lands at the documented kernel index."""
    ws = [np.zeros((cout, cin, 3 if taps == 27 else 1, 3 if taps == 27 else 1,
                    3 if taps == 27 else 1), np.float32)
          for _n, taps, _s, cin, cout, _r in learned.LAYERS]
    ws[0][5, 1, 2, 0, 1] = 1.0      # in: co=5, c=1, tap = 9*2 + 3*0 + 1 = 19 -> k = 39
    ws[1][7, 3, 0, 1, 2] = 2.0      # e1: tap 5, co 7, ci 3 -> [5][7][3] of [27][32][16]
    ws[4][9, 60, 0, 0, 0] = 3.0     # p2: [co][ci] = [9][60] of [48][64]
    ws[10][0, 4, 1, 1, 1] = 4.0     # out: tap 13, col 0, ci 4
    blob = learned.pack(ws)
    off, total = learned.layout()
    vals = learned.bf16_bits_to_f32(np.frombuffer(blob, "<u2", total, 24))
    nz = np.flatnonzero(vals)
    want = [off[0] + 5 * 64 + 39, off[1] + (5 * 32 + 7) * 16 + 3, off[4] + 9 * 64 + 60,
            off[10] + (13 * 16 + 0) * 16 + 4]
    assert list(nz) == want and list(vals[nz]) == [1.0, 2.0, 3.0, 4.0]


def test_torch_operator_matches_oracle():
    d = scenes.dam_break(16)
    o = OracleSim(d, threads=4)
    o.init()
    for _ in range(2):
        o.step()
    n = o.n
    flags, L = o.get_flags(), o.get_field(_abi.FIELD_L)
    diag, fluid = lnet.laplacian(torch.from_numpy(flags.reshape(n, n, n)),
                                 torch.from_numpy(L.reshape(n, n, n)))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(n ** 3).astype(np.float32)
    y = lnet.apply_laplacian(torch.from_numpy(x).view(1, 1, n, n, n), diag.view(1, 1, n, n, n),
                             fluid.view(1, 1, n, n, n)).reshape(-1).numpy()
    ref = o.apply_A(x)
    assert np.abs(y - ref).max() <= 1e-5 * np.abs(ref).max()
    # the air terms are present (ghost-fluid diagonal differs from the plain count)
    assert (diag[fluid > 0] % 1.0 != 0).any()


def test_scale_and_homogeneity():
    torch.manual_seed(0)
    n = 16
    flags = np.zeros((n, n, n), np.uint8)
    flags[2:9, 3:12, :] = lnet.FLUID
    r = np.random.default_rng(5).standard_normal((n, n, n)).astype(np.float32)
    fl = torch.from_numpy((flags == lnet.FLUID).astype(np.float32)).view(1, 1, n, n, n)
    rt = torch.from_numpy(r).view(1, 1, n, n, n)
    s = lnet.PressureNet.scale(rt, fl)
    assert float(s) == float(lnet.numpy_scale(r, flags))
    m = lnet.PressureNet(seed=1)
    diag = torch.full_like(fl, 6.0) - 2.0 * torch.rand(fl.shape, generator=torch.Generator().manual_seed(1))
    dw = lnet.diag_weight(diag * fl, fl)
    assert torch.all((dw > 0) == (fl > 0))
    with torch.no_grad():
        z1 = m(rt, fl, dw)
        z2 = m(rt * 4.0, fl, dw)
    assert torch.allclose(z2, 4.0 * z1, rtol=1e-5, atol=1e-6 * float(z1.abs().max()))
    assert torch.all(z1[fl == 0] == 0)


def test_a_angle_loss_is_zero_at_the_solution():
    n = 8
    flags = torch.zeros(n, n, n, dtype=torch.uint8)
    flags[1:7, 1:5, 1:7] = lnet.FLUID
    flags[1:7, 5:7, 1:7] = lnet.AIR
    L = torch.where(flags == lnet.FLUID, torch.tensor(-0.5), torch.tensor(0.5))
    diag, fluid = lnet.laplacian(flags, L)
    # assemble the matrix from the operator and solve exactly
    idx = torch.nonzero(fluid.reshape(-1)).reshape(-1)
    cols = []
    for c in idx:
        e = torch.zeros(n ** 3)
        e[c] = 1.0
        cols.append(lnet.apply_laplacian(e.view(1, 1, n, n, n), diag.view(1, 1, n, n, n),
                                         fluid.view(1, 1, n, n, n)).reshape(-1)[idx])
    A = torch.stack(cols, 1).double()
    assert torch.allclose(A, A.T)
    r = torch.zeros(n ** 3, dtype=torch.float64)
    r[idx] = torch.randn(idx.numel(), dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    x = torch.zeros_like(r)
    x[idx] = torch.linalg.solve(A, r[idx])
    sh = (1, 1, n, n, n)
    args = (r.float().view(sh), x.float().view(sh), diag.view(sh), fluid.view(sh))
    assert float(lnet.a_angle_loss(x.float().view(sh) * 3.0, *args)) == pytest.approx(0.0, abs=1e-5)
    # signed: the reversed solution is the worst direction, not a second minimum
    assert float(lnet.a_angle_loss(x.float().view(sh) * -3.0, *args)) == pytest.approx(2.0, abs=1e-5)
    assert float(lnet.a_angle_loss(r.float().view(sh), *args)) > 1e-3
