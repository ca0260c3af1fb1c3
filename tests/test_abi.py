"""C ABI checks that run without a GPU: every entry point declared in
include/ariel.h is exported by libariel_amd.so and bound by the Python mirror,
struct layouts agree between C, ctypes and the oracle, descriptor validation
fails loudly, and the C++ FlipSim facade compiles and links."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from ariel_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ariel.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ariel_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _abi.load()
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} not exported"
        assert n in _abi.PROTOTYPES, f"{n} not bound in ariel_amd/_abi.py"
    nm = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r" T (ariel_\w+)", nm))
    assert set(names) == exported


def test_struct_layouts_match_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "ariel.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(ariel_scene_desc), sizeof(ariel_step_stats),
         sizeof(ariel_liquid), offsetof(ariel_scene_desc, spring), sizeof(ariel_kernel_stats));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(_abi.SceneDesc)
    assert vals[1] == ctypes.sizeof(_abi.StepStats)
    assert vals[2] == ctypes.sizeof(_abi.Liquid)
    assert vals[3] == _abi.SceneDesc.spring.offset
    assert vals[4] == ctypes.sizeof(_abi.KernelStats)
    from oracle.oracle import OracleStats
    assert ctypes.sizeof(OracleStats) == ctypes.sizeof(_abi.StepStats)


def test_desc_defaults_are_the_reference_defaults():
    d = scenes.defaults()
    assert list(d.dims) == [32, 32, 32]          # sceneloader.cpp:413
    assert d.density == pytest.approx(0.5)       # sceneloader.cpp:412
    assert d.dt == pytest.approx(0.005)          # sceneloader.cpp:418
    assert d.picflip == pytest.approx(0.95)      # flip.cpp:27
    assert d.sharpen_re == pytest.approx(1.4)    # particlegridoperations.inl:149
    assert d.spring == pytest.approx(50.0)       # particleresampler.inl:42
    assert d.mic_sigma == pytest.approx(0.25)    # solver.inl:107
    raw = scenes.raw_defaults()
    for f, _ in _abi.SceneDesc._fields_:
        if f in ("forces", "liquids", "reserved", "dims"):
            continue
        assert getattr(raw, f) == getattr(d, f), f


@pytest.mark.parametrize("field,value,code", [
    ("dims", [32, 32, 16], -1),
    ("dims", [30, 30, 30], -6),
    ("density", 0.3, -1),
    ("precond", 1, -6),
    ("tile", 4, -6),
    ("max_iters", 0, -1),
])
def test_invalid_descriptors_fail_loudly(field, value, code):
    lib = _abi.load()
    d = scenes.dam_break(32)
    if field == "dims":
        d.dims[:] = value
    else:
        setattr(d, field, value)
    h = ctypes.c_void_p()
    rc = lib.ariel_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert len(lib.ariel_last_error(None)) > 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_hip_error():
    from ariel_amd import ArielError, FlipSim
    with pytest.raises(ArielError) as e:
        FlipSim(scenes.dam_break(32))
    assert e.value.code == -2


def test_cpp_facade_compiles_and_links():
    src = r'''
#include "ariel_flipsim.hpp"
#include <cstdio>
int main() {
  ariel_scene_desc d;
  ariel_desc_defaults(&d);
  try {
    ariel::FlipSim sim(d);
    sim.Init();
    sim.Step();
    std::printf("%zu %d\n", sim.GetParticles().size(), sim.m_frame);
  } catch (const ariel::Error& e) {
    std::printf("error %d\n", e.code());
  }
  try {
    ariel::FlipSim js(std::string("{\"sim\": []}"), std::string("."));
  } catch (const ariel::Error& e) {
    std::printf("json error %d\n", e.code());
  }
  return 0;
}
'''
    libdir = os.path.dirname(_abi.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "f.cpp")
        exe = os.path.join(d, "f")
        open(cpp, "w").write(src)
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), cpp, "-o", exe,
                        "-L", libdir, "-lariel_amd", f"-Wl,-rpath,{libdir}",
                        "-Wl,-rpath,/opt/rocm/lib"], check=True)
        if not os.path.exists("/dev/kfd"):
            out = subprocess.run([exe], capture_output=True, text=True, timeout=60).stdout
            assert out.startswith("error -2")


def test_batch_rejects_mixed_solver_settings():
    """Tanks of a batch share one launch sequence: mixed preconditioners are
    rejected before any device work."""
    import ctypes
    from ariel_amd import _abi, scenes
    lib = _abi.load()
    a = scenes.tank(16, (0.3, 0.5, 0.5))
    b = scenes.tank(24, (0.3, 0.5, 0.5))
    b.precond = _abi.PRECOND_JACOBI
    arr = (_abi.SceneDesc * 2)(a, b)
    h = ctypes.c_void_p()
    rc = lib.ariel_batch_create(arr, 2, ctypes.byref(h))
    assert rc == _abi.ERR_INVALID
    assert b"share" in lib.ariel_batch_last_error(None)
    assert lib.ariel_batch_create(arr, 0, ctypes.byref(h)) == _abi.ERR_INVALID
    arr[0].precond = _abi.PRECOND_LEARNED
    assert lib.ariel_batch_create(arr, 1, ctypes.byref(h)) == _abi.ERR_UNSUPPORTED
    assert b"single tanks" in lib.ariel_batch_last_error(None)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_trainer_without_gpu_fails_loudly():
    import ctypes
    from ariel_amd import _abi
    lib = _abi.load()
    h = ctypes.c_void_p()
    assert lib.ariel_trainer_create(ctypes.byref(h)) == _abi.ERR_HIP
    assert not h.value
    assert b"trainer_create" in lib.ariel_last_error(None)
