"""The allocation-free level set (ariel_export_levelset_noalloc) through the C
ABI against ariel_export_levelset on the same handle: bit-identical volumes on
two calls in a row, the second on a smaller input (a smaller grid with fewer
particles), and the documented limits refused with ARIEL_ERR_CAPACITY."""
import ctypes

import numpy as np
import pytest
import torch

from ariel_amd import FlipSim, _abi, read_levelset, scenes

pytestmark = pytest.mark.gpu


def _noalloc(sim, cap=None):
    m = sim.n + 1
    out = torch.full((m ** 3,), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = sim._lib.ariel_export_levelset_noalloc(sim._h, ctypes.c_void_p(out.data_ptr()),
                                                m ** 3 if cap is None else cap)
    if rc != 0:
        return rc, None
    sim.sync()
    return rc, out.cpu().numpy().reshape(m, m, m)


def _reference(sim, path):
    sim.export_levelset(path)
    return read_levelset(path)[0]


def test_noalloc_levelset_matches_export(tmp_path):
    big = FlipSim(scenes.pool_and_cube(48))
    big.Init()
    big.Step(2)
    small = FlipSim(scenes.dam_break(24))
    small.Init()
    small.Step(1)
    assert big.num_particles > small.num_particles > 0
    for name, sim in (("48^3 pool + cube", big), ("24^3 dam break", small)):
        print("no-alloc level set:", name, sim.num_particles, "particles", flush=True)
        rc, phi = _noalloc(sim)
        assert rc == 0, sim._lib.ariel_last_error(sim._h).decode()
        ref = _reference(sim, tmp_path / "ref.levelset")
        assert np.array_equal(phi, ref), f"{name}: {int((phi != ref).sum())} nodes differ"
        assert (phi < 0).any() and (phi == 3.0).any()
    # buffer too small
    rc, _ = _noalloc(small, cap=10)
    assert rc == _abi.ERR_CAPACITY
    big.close()
    small.close()


def test_noalloc_levelset_limits():
    d = scenes.dam_break(72)
    sim = FlipSim(d)
    sim.Init()
    rc, _ = _noalloc(sim)
    assert rc == _abi.ERR_CAPACITY
    assert b"above the limit" in sim._lib.ariel_last_error(sim._h)
    sim.close()
