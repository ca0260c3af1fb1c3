"""Writes hash4_vectors.json: fixed input/output vectors of the counter-based hash
that replaces the reference's rand() (particleresampler.inl:69-71) and jitters the
seeding lattice.

The hash is C. Wellons' published 32-bit integer finaliser "lowbias32" (hash
prospector, 2018), chained over four words:
    hash4(seed, a, b, c) = lowbias32(lowbias32(lowbias32(lowbias32(seed) ^ a) ^ b) ^ c)
This script restates it in plain Python integer arithmetic, independently of the
C++ oracle and the HIP kernels, and records its outputs on a fixed set of inputs
(edge words, the keys the seeding and the resampler use, and random words).

    python tests/golden/make_hash_vectors.py
"""
import json
import os
import random

M32 = 0xFFFFFFFF


def lowbias32(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def hash4(seed, a, b, c):
    return lowbias32(lowbias32(lowbias32(lowbias32(seed) ^ a) ^ b) ^ c)


def inputs():
    edge = [0, 1, 2, 3, 100, 0x7FFFFFFF, 0x80000000, M32]
    out = [(s, a, b, c) for s in (0, 1, 7) for a in (0, 1, 2) for b in edge[:4] for c in edge[::3]]
    # seeding keys (seed, 3l + axis, a*N*m + b, c) and resampler keys (seed, 3 frame + axis, id_p, id_q)
    out += [(1, 3 * l + ax, 17 * 512 + 33, 250) for l in range(2) for ax in range(3)]
    out += [(2, 3 * f + ax, 123456, 123457) for f in (1, 5, 99) for ax in range(3)]
    rng = random.Random(20261204)
    out += [tuple(rng.getrandbits(32) for _ in range(4)) for _ in range(64)]
    return out


def main():
    vec = [{"in": list(k), "out": hash4(*k)} for k in inputs()]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash4_vectors.json")
    with open(path, "w") as f:
        f.write('{"hash": "lowbias32 chained: hash4(seed,a,b,c)",\n "vectors": [\n')
        f.write(",\n".join(json.dumps(v) for v in vec))
        f.write("\n]}\n")
    print(f"wrote {len(vec)} vectors to {path}")


if __name__ == "__main__":
    main()
