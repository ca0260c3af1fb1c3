"""Build libariel_amd.so (hand-written HIP for gfx950) in-tree.

    python -m ariel_amd.build

Every translation unit under ariel_amd/csrc/*.hip is compiled with hipcc for
gfx950 only and linked into ariel_amd/lib/libariel_amd.so. Floating point is kept
IEEE (no FMA contraction, correctly rounded division and sqrt) because the
kernels must agree bit for bit with a serial evaluation of the same expressions.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libariel_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(SRC, "*.hip")))
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, obj, r in ex.map(_compile, srcs):
            msgs = [ln for ln in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in ln and ln.strip()]
            if r.returncode != 0:
                sys.stderr.write("\n".join(msgs) + "\n")
                raise RuntimeError(f"hipcc failed on {src}")
            if verbose and msgs:
                sys.stderr.write("\n".join(msgs) + "\n")
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
