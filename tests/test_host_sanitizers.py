"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds the library's translation units with the sanitizers on their host side
(device code is compiled as usual and never runs here), the CPU oracle and the
driver tests/host_sanitize_main.cpp with the same sanitizers, links them into
one executable and runs it: the JSON scene reader on valid, truncated and
mutated scenes, the mesher, the descriptor checks and the creation and error
paths that end without a GPU, and the oracle through steps in both solver
modes, stage runs, the operator, every preconditioner and the FP64 solve. Any
sanitizer report aborts the run (-fno-sanitize-recover), leaks included.
"""
import concurrent.futures
import glob
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN = "address,undefined"


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    return cmd, r


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="needs hipcc and clang++")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    out = str(tmp_path)
    hip_flags = ["--offload-arch=gfx950", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off",
                 "-Wno-unused-value",
                 "-Xarch_host", "-g",
                 "-Xarch_host", f"-fsanitize={SAN}",
                 "-Xarch_host", "-fno-sanitize-recover=all",
                 "-Xarch_host", "-fno-omit-frame-pointer"]
    host_flags = ["-O1", "-g", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off",
                  f"-fsanitize={SAN}", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
    jobs, objs = [], []
    for src in sorted(glob.glob(os.path.join(ROOT, "ariel_amd", "csrc", "*.hip"))):
        obj = os.path.join(out, os.path.basename(src) + ".o")
        jobs.append([HIPCC, *hip_flags, "-c", src, "-o", obj])
        objs.append(obj)
    for src in (os.path.join(ROOT, "oracle", "flip_oracle.cpp"), os.path.join(HERE, "host_sanitize_main.cpp")):
        obj = os.path.join(out, os.path.basename(src) + ".o")
        jobs.append([CLANG, *host_flags, "-c", src, "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for cmd, r in ex.map(lambda c: _run(c, out), jobs):
            assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stderr[-4000:]}"
    exe = os.path.join(out, "host_sanitize")
    cmd, r = _run([CLANG, f"-fsanitize={SAN}", "-pthread", "-o", exe, *objs, "-L/opt/rocm/lib",
                   "-lamdhip64", "-lz", "-Wl,-rpath,/opt/rocm/lib"], out)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, out], cwd=out, capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout[-2000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert r.stdout.strip().endswith("ok"), report
