"""Planted mistakes: each copy of the oracle source below carries one plausible
error (a wrong window, radius, constant, face offset, ghost ratio, sign,
averaging rule, weight, accept rule, hash key or jitter scale) and must fail the brute-force and closed-form pins of
test_oracle_pins.py (and the reference solver's direct-solve pin). The copies are compiled on the CPU into a temporary
folder; the committed oracle is left untouched."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "flip_oracle.cpp")

MISTAKES = {
    "density window of two cells": ("for_cell_neighbors(s, position, 1,", "for_cell_neighbors(s, position, 2,"),
    "density radius 3 instead of 4 lattice spacings": (
        "const float hsm = 4.0f * s.d.density / maxd;", "const float hsm = 3.0f * s.d.density / maxd;"),
    "cell SDF constant": ("return 0.2f * n0 - accm;", "return 0.25f * n0 - accm;"),
    "splat x window symmetric": ("for_wall_neighbors(s, i, j, k, 1, 2, 2,", "for_wall_neighbors(s, i, j, k, 2, 2, 2,"),
    "splat y face at the cell centre": (
        "V3 py = {(float)i + 0.5f, (float)j, (float)k + 0.5f};",
        "V3 py = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};"),
    "ghost pressure ratio inverted": (
        "else pb = Lc(bi, bj, bk) / gmin(1.0e-6f, Lc(i, j, k)) * Pc(i, j, k);",
        "else pb = Lc(i, j, k) / gmin(1.0e-6f, Lc(bi, bj, bk)) * Pc(i, j, k);"),
    "pressure gradient sign": ("return (pf - pb) / h;", "return (pb - pf) / h;"),
    "extrapolation into every unmarked face": (
        "if (!(!mark[n][f] && wallmark[n][f])) continue;", "if (!(!mark[n][f])) continue;"),
    "extrapolation mean over one extra face": (
        "if (wsum) s.uf[n][f] = sum / (float)wsum;", "if (wsum) s.uf[n][f] = sum / (float)(wsum + 1);"),
    "resample spring pulls instead of pushing": (
        "spring.x += w * (pp.x - np_.x) / dist * re;", "spring.x += w * (np_.x - pp.x) / dist * re;"),
    "resample radius of a whole cell": (
        "const float re = s.d.density / maxd;", "const float re = 2.0f * s.d.density / maxd;"),
    "resample velocity with Smooth weights": (
        "float w = 1.0f * sharpen(dist2, re);", "float w = 1.0f * smooth(dist2, re);"),
    "resample accept rule inverted": ("if (!inside_wall(s, s.t[n])) {", "if (inside_wall(s, s.t[n])) {"),
    "PIC/FLIP ratio on the PIC part": (
        "s.u[i].y = (1.0f - ratio) * pic.y + ratio * t.y;", "s.u[i].y = ratio * pic.y + (1.0f - ratio) * t.y;"),
    "advection clamp of two cells": ("const float r = 1.0f / s.maxd;", "const float r = 2.0f / s.maxd;"),
    "boundary pass between like cells": (
        "if (i < N && i > 0 && check_wall(s, i, j, k) * check_wall(s, i - 1, j, k) < 0)",
        "if (i < N && i > 0 && check_wall(s, i, j, k) * check_wall(s, i - 1, j, k) > 0)"),
    "divergence y faces swapped": (
        "s.uf[1][s.fidx(1, i, j + 1, k)] - s.uf[1][s.fidx(1, i, j, k)] +\n                            s.uf[2][s.fidx(2, i, j, k + 1)] - s.uf[2][s.fidx(2, i, j, k)]) / h;\n                s.D",
        "s.uf[1][s.fidx(1, i, j, k)] - s.uf[1][s.fidx(1, i, j + 1, k)] +\n                            s.uf[2][s.fidx(2, i, j, k + 1)] - s.uf[2][s.fidx(2, i, j, k)]) / h;\n                s.D"),
    "coincident-pair jitter scale 0.05": ("spring.x += 0.01f * re / dt", "spring.x += 0.05f * re / dt"),
    "coincident-pair hash keys swapped": (
        "hash4(seed, 3u * fr + 0u, (uint32_t)n0, (uint32_t)q)", "hash4(seed, 3u * fr + 0u, (uint32_t)q, (uint32_t)n0)"),
    "hash key words mixed in the wrong order": (
        "return mix32(mix32(mix32(mix32(seed) ^ a) ^ b) ^ c);", "return mix32(mix32(mix32(mix32(seed) ^ b) ^ a) ^ c);"),
    "hash multiplier off by two": ("x *= 0x7feb352dU;", "x *= 0x7feb352bU;"),
    "seeding jitter axes swapped": (
        "jx = hash_unit(hash4(s.d.seed, 3u * l + 0u, key, c));", "jx = hash_unit(hash4(s.d.seed, 3u * l + 1u, key, c));"),
    "reference solver stop ten times looser": (
        "float eps = 1.0e-2f * (float)(N * N * N);", "float eps = 1.0e-1f * (float)(N * N * N);"),
    "reference solver beta inverted": ("float beta = a2 / a;", "float beta = a / a2;"),
}


def _run_pins(tmp_path, text):
    src = tmp_path / "flip_oracle.cpp"
    src.write_text(text)
    lib = tmp_path / "liboracle.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-shared",
                    "-I", os.path.join(ROOT, "oracle"), str(src), "-o", str(lib)], check=True)
    env = dict(os.environ, ARIEL_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"),
                        os.path.join(ROOT, "tests", "test_oracle_solver.py") +
                        "::test_faithful_reference_solver_matches_direct_solve"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


def test_unplanted_copy_passes_the_pins(tmp_path):
    rc, out = _run_pins(tmp_path, open(SRC).read())
    assert rc == 0, out


@pytest.mark.parametrize("name", list(MISTAKES))
def test_planted_mistake_fails_the_pins(tmp_path, name):
    old, new = MISTAKES[name]
    text = open(SRC).read()
    assert text.count(old) == 1, "the planted site must exist exactly once"
    rc, out = _run_pins(tmp_path, text.replace(old, new))
    assert rc == 1, f"{name}: the pins passed on a broken oracle\n{out}"
