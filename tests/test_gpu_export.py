"""Frame export on the GPU (Scene::ExportParticles, src/scene/scene.cpp:34-101):
the GPU-rasterised particle level sets (spheres, and the reference's velocity
trails) equal the oracle's bit for bit, the OpenVDB file holds the same volume,
the exported mesh is the host mesher's on that level set, the particle files
(.geo, .bgeo, gzipped) hold exactly GetParticles() in grid units, and the C++
facade writes all three files per frame."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

from ariel_amd import ArielError, FlipSim, _abi, mesh_sdf, read_levelset, read_partio, read_vdb, scenes
from oracle.levelset import BAND, particle_levelset, trail_levelset

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _sim(desc, steps):
    sim = FlipSim(desc)
    sim.Init()
    if steps:
        sim.Step(steps)
    return sim


@pytest.mark.parametrize("name,desc,steps", [
    ("dam break 32, step 3", lambda: scenes.dam_break(32), 3),
    ("dam break + sphere 48, step 2", lambda: scenes.dam_break_sphere(48), 2),
    ("pool + cube 64, initial", lambda: scenes.pool_and_cube(64, seed=1), 0),
])
def test_levelset_matches_oracle(tmp_path, name, desc, steps):
    print("case:", name, flush=True)
    d = desc()
    sim = _sim(d, steps)
    path = tmp_path / "fluid.levelset"
    sim.export_levelset(path)
    phi, bg = read_levelset(path)
    pos, _vel, _rho = sim.GetParticles()
    n = sim.n
    ref = particle_levelset(pos, n, n)
    assert bg == BAND and phi.shape == (n + 1,) * 3
    assert np.array_equal(phi, ref), f"{np.count_nonzero(phi != ref)} nodes differ"
    assert (phi < 0).any() and (phi == BAND).any()
    sim.close()


@pytest.mark.parametrize("name,desc,steps", [
    ("dam break 24, step 3", lambda: scenes.dam_break(24), 3),
    ("dam break + sphere 32, step 2", lambda: scenes.dam_break_sphere(32), 2),
])
def test_trail_levelset_matches_oracle(tmp_path, name, desc, steps):
    """The reference's rasterizeTrails (levelset.cpp:109) on the GPU against the
    oracle's restatement, bit for bit; the .vdb file holds the same volume."""
    print("case:", name, flush=True)
    sim = _sim(desc(), steps)
    sim.set_levelset_mode(_abi.LEVELSET_TRAILS)
    sim.export_levelset(tmp_path / "t.levelset")
    phi, bg = read_levelset(tmp_path / "t.levelset")
    pos, vel, _rho = sim.GetParticles()
    n = sim.n
    ref = trail_levelset(pos, vel, n, n)
    assert np.array_equal(phi, ref), f"{np.count_nonzero(phi != ref)} nodes differ"
    sph = particle_levelset(pos, n, n)
    assert np.all(phi <= sph + 2.5e-7)
    if np.linalg.norm(vel, axis=1).max() > 0.5:   # a trail longer than the first step of 0.25
        assert (phi < sph - 1e-4).any()
    sim.export_levelset(tmp_path / "t.vdb")
    back, vbg, active, meta = read_vdb(tmp_path / "t.vdb", n + 1)
    assert vbg == bg and np.array_equal(back, phi) and np.array_equal(active, np.abs(phi) < BAND)
    sim.close()


def test_vdb_export_is_the_raw_volume(tmp_path):
    sim = _sim(scenes.pool_and_cube(40, seed=1), 1)
    sim.export_levelset(tmp_path / "s.levelset")
    sim.export_levelset(tmp_path / "s.vdb")
    phi, bg = read_levelset(tmp_path / "s.levelset")
    back, vbg, active, meta = read_vdb(tmp_path / "s.vdb", sim.n + 1)
    assert np.array_equal(back, phi) and vbg == bg
    assert meta["file_voxel_count"] == int((np.abs(phi) < BAND).sum()) > 0
    sim.close()


def test_mesh_is_the_host_meshers_and_particles_file_is_getparticles(tmp_path):
    sim = _sim(scenes.dam_break_sphere(48), 2)
    sim.export_levelset(tmp_path / "f.levelset")
    sim.export_mesh(tmp_path / "f.obj")
    phi, _ = read_levelset(tmp_path / "f.levelset")
    nv, nq = mesh_sdf(phi, tmp_path / "host.obj")
    assert nv > 1000 and nq > 1000
    assert filecmp.cmp(tmp_path / "f.obj", tmp_path / "host.obj", shallow=False)

    sim.export_particles(tmp_path / "f.geo")
    pos, vel, _ = sim.GetParticles()
    lines = open(tmp_path / "f.geo").read().splitlines()
    assert lines[0] == "PGEOMETRY V5"
    assert lines[1] == f"NPoints {len(pos)} NPrims 0"
    assert lines[5:7] == ["v 3 vector 0 0 0", "id 1 int 0"]
    assert lines[-2:] == ["beginExtra", "endExtra"]
    body = lines[7:-2]
    assert len(body) == len(pos)
    vals = np.array([ln.replace("(", " ").replace(")", " ").split() for ln in body], np.float64)
    assert np.array_equal(vals[:, 0:3].astype(np.float32), pos * np.float32(sim.n))
    assert np.all(vals[:, 3] == 1.0)
    assert np.array_equal(vals[:, 4:7].astype(np.float32), vel)
    assert np.array_equal(vals[:, 7], np.arange(len(pos)))

    for ext in (".bgeo", ".geo.gz", ".bgeo.gz"):
        sim.export_particles(tmp_path / ("f" + ext))
        p, v, i = read_partio(tmp_path / ("f" + ext))
        assert np.array_equal(p, pos * np.float32(sim.n)) and np.array_equal(v, vel)
        assert np.array_equal(i, np.arange(len(pos)))
    with pytest.raises(ArielError) as e:
        sim.export_particles(tmp_path / "f.pdb")
    assert e.value.code == -6
    sim.close()


FACADE = r'''
#include "ariel_flipsim.hpp"
#include <cstdio>
int main(int argc, char** argv) {
  ariel_scene_desc d;
  if (ariel_desc_from_json("{\"settings\":[{\"density\":0.5,\"step_size\":0.004166667,"
                           "\"dim\":{\"x\":32,\"y\":32,\"z\":32}}],"
                           "\"globalforces\":[{\"x\":0,\"y\":-9.8,\"z\":0}],"
                           "\"meshfiles\":[{\"id\":\"b\",\"mesh_gen\":\"box\","
                           "\"point0\":{\"x\":1,\"y\":1,\"z\":1},\"point1\":{\"x\":12,\"y\":18,\"z\":31}}],"
                           "\"transforms\":[{\"id\":\"t\"}],"
                           "\"geoms\":[{\"id\":\"g\",\"type\":\"mesh\",\"geom_frames\":[\"b\"],"
                           "\"transform_frames\":[\"t\"],\"pre_persist\":false,\"post_persist\":false,"
                           "\"frame_interval\":1,\"frame_offset\":0}],"
                           "\"sim\":[{\"geom\":\"g\",\"type\":\"liquid\"}]}", &d) != ARIEL_OK) {
    std::printf("json error %s\n", ariel_last_error(nullptr));
    return 1;
  }
  try {
    ariel::FlipSim sim(d);
    sim.Init();
    sim.Step(false, false, false, argv[1]);
    sim.Step(true, true, true, argv[1]);
    std::printf("frame %d\n", sim.m_frame);
  } catch (const ariel::Error& e) {
    std::printf("error %d %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
'''


def test_facade_step_writes_the_three_files_of_the_frame(tmp_path):
    src = tmp_path / "f.cpp"
    src.write_text(FACADE)
    exe = tmp_path / "f"
    libdir = os.path.dirname(_abi.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lariel_amd", f"-Wl,-rpath,{libdir}", "-Wl,-rpath,/opt/rocm/lib"],
                   check=True)
    print("facade program built; running 2 steps with export", flush=True)
    base = str(tmp_path / "out")
    r = subprocess.run([str(exe), base], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "frame 2" in r.stdout
    names = sorted(os.listdir(tmp_path))
    assert "out.0002.vdb" in names and "out.0002.obj" in names and "out.0002.geo" in names
    assert not any(n.startswith("out.0001") for n in names)
    phi, _bg, _active, meta = read_vdb(tmp_path / "out.0002.vdb", 33)
    assert meta["class"] == "level set"
    mesh_sdf(phi, tmp_path / "again.obj")
    assert filecmp.cmp(tmp_path / "out.0002.obj", tmp_path / "again.obj", shallow=False)
