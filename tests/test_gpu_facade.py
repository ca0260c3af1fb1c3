"""The C++ facade drives the GPU path as the reference's callers drive
fluidCore::FlipSim: a scene in the reference's JSON format, Init() once, Step()
in a loop, GetParticles() after each step; the result equals the Python path's."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ariel_amd import FlipSim, _abi, scenes
from tests.test_scene_json import SOLVER, _box, _scene

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "ariel_flipsim.hpp"
#include <cstdio>
#include <fstream>
#include <sstream>
int main(int argc, char** argv) {
  std::ifstream f(argv[1]);
  std::stringstream ss;
  ss << f.rdbuf();
  ariel_scene_desc d;
  if (ariel_desc_from_json(ss.str().c_str(), &d) != ARIEL_OK) {
    std::printf("json error %s\n", ariel_last_error(nullptr));
    return 1;
  }
  d.wall = 1; d.jitter = 1; d.seed = 0; d.tol = 1e-6f; d.max_iters = 500;
  try {
    ariel::FlipSim sim(d);
    sim.Init();
    for (int s = 0; s < 3; ++s) sim.Step();
    std::vector<ariel::Particle> ps = sim.GetParticles();
    double sx = 0, sv = 0;
    for (const auto& p : ps) { sx += p.p[0] + p.p[1] + p.p[2]; sv += p.u[1]; }
    std::printf("%zu %d %.9e %.9e\n", ps.size(), sim.m_frame, sx, sv);
  } catch (const ariel::Error& e) {
    std::printf("error %d %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
'''


def test_facade_with_json_scene_matches_python_path():
    n = 32
    text = _scene(n, [_box((0, 0, 0), (0.4 * n, 0.6 * n, n))])
    libdir = os.path.dirname(_abi.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        cpp, exe, js = (os.path.join(d, x) for x in ("f.cpp", "f", "scene.json"))
        open(cpp, "w").write(SRC)
        open(js, "w").write(text)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), cpp, "-o", exe,
                        "-L", libdir, "-lariel_amd", f"-Wl,-rpath,{libdir}",
                        "-Wl,-rpath,/opt/rocm/lib"], check=True)
        print("facade program built; running 3 steps", flush=True)
        out = subprocess.run([exe, js], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    count, frame, sx, sv = out.stdout.split()
    sim = FlipSim(scenes.from_json(text, **SOLVER))
    sim.Init()
    sim.Step(3)
    pos, vel, _ = sim.GetParticles()
    sim.close()
    assert int(count) == len(pos) and int(frame) == 3
    # same descriptor, same kernels: identical state, so identical float sums
    assert float(sx) == pytest.approx(float(np.sum(pos, dtype=np.float64)), rel=1e-9)
    assert float(sv) == pytest.approx(float(np.sum(vel[:, 1], dtype=np.float64)), rel=1e-9)
