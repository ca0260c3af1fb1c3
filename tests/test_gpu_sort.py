"""The counting sort's per-cell order (ParticleGrid::Sort, src/grid/particlegrid.cpp:
140-170: each cell lists its particles in particle-vector order) on inputs the
simulated scenes do not produce: particles handed over in a shuffled order (every
cell's segment needs reordering) and a crowded run of 64 cells holding more
particles than a wave stages in LDS (the per-cell sort's global-memory path)."""
import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from oracle import OracleSim
from tests.helpers import assert_bitexact

pytestmark = pytest.mark.gpu


def _sorted_pair(pos, vel, rho, n=32):
    d = scenes.dam_break(n)
    g = FlipSim(d)
    g.Init()
    o = OracleSim(d)
    o.init()
    g.set_particles(pos, vel, rho)
    o.set_particles(pos, vel, rho)
    g.run_stage(_abi.STAGE_SORT)
    o.stage(_abi.STAGE_SORT)
    ig, og = g.get_sort()
    io, oo = o.get_sort()
    g.close()
    o.close()
    return ig, og, io, oo


def _base(n=32):
    g = FlipSim(scenes.dam_break(n))
    g.Init()
    pos, vel, rho = g.GetParticles()
    g.close()
    return pos, vel, rho


def test_shuffled_particles_sort_like_the_oracle():
    pos, vel, rho = _base()
    perm = np.random.default_rng(1).permutation(len(pos))
    ig, og, io, oo = _sorted_pair(pos[perm], vel[perm], rho[perm])
    assert_bitexact(np.diff(og), np.diff(oo), "per-cell counts")
    assert_bitexact(ig, io, "sort ids")


def test_crowded_cells_sort_like_the_oracle():
    n = 32
    pos, vel, rho = _base(n)
    rng = np.random.default_rng(2)
    # cells (5, 6, 0..31) and (5, 7, 0..31): 64 consecutive cell indices starting
    # at a multiple of 64, 40 particles each (2560 > the 1024 a wave stages)
    cells = [(5, j, k) for j in (6, 7) for k in range(n)]
    extra = np.concatenate([(np.array(c, np.float32) + rng.uniform(0.05, 0.95, (40, 3)).astype(np.float32))
                            / np.float32(n) for c in cells])
    allp = np.concatenate([pos, extra])
    allv = np.concatenate([vel, np.zeros_like(extra)])
    allr = np.concatenate([rho, np.zeros(len(extra), np.float32)])
    perm = rng.permutation(len(allp))
    ig, og, io, oo = _sorted_pair(allp[perm], allv[perm], allr[perm], n)
    counts = np.diff(og)
    c0 = (5 * n + 6) * n
    assert c0 % 64 == 0 and counts[c0:c0 + 64].sum() > 1024
    assert_bitexact(counts, np.diff(oo), "per-cell counts")
    assert_bitexact(ig, io, "sort ids")
