"""The reference scene's full geometry (ariel_create_from_json / ariel_scene_emit,
host only): liquid geoms emit the lattice points they contain by ray parity
(src/scene/scene.cpp:257-300, 362-378, 436-455) through mesh files, the box and
sphere generators, rotated and interpolated transforms, animated sequences and
frame ranges (src/geom/mesh.cpp:61-109). Each case is checked against what the
geometry fixes by an independent method: the oracle's analytic box seeding,
half-space and radius bounds, a generalized winding number."""
import json

import numpy as np
import pytest

from ariel_amd import ArielError, scenes
from oracle import OracleSim

N = 32


def _doc(meshfiles, transforms, geoms, sim, n=N, animated=None):
    doc = {"settings": [{"density": 0.5, "step_size": 1.0 / 240.0, "dim": {"x": n, "y": n, "z": n}}],
           "globalforces": [{"x": 0, "y": -9.8, "z": 0}],
           "meshfiles": meshfiles, "transforms": transforms, "geoms": geoms, "sim": sim}
    if animated:
        doc["animatedmeshes"] = animated
    return json.dumps(doc)


def _geom(gid, meshes, xforms, offset=0, interval=1, pre=False, post=False, kind="mesh", anim=None):
    g = {"id": gid, "type": kind, "transform_frames": xforms, "pre_persist": pre, "post_persist": post,
         "frame_interval": interval, "frame_offset": offset}
    if kind == "mesh":
        g["geom_frames"] = meshes
    else:
        g["anim_sequence"] = anim
    return g


def _box(mid, p0, p1):
    return {"id": mid, "mesh_gen": "box", "point0": dict(zip("xyz", p0)), "point1": dict(zip("xyz", p1))}


def _one(meshfile, transform=None, **geom_kw):
    return _doc([meshfile], [dict(id="t", **(transform or {}))],
                [_geom("g", [meshfile["id"]], ["t"], **geom_kw)], [{"geom": "g", "type": "liquid"}])


def _solver(jitter=1):
    return scenes.dam_break(N, jitter=jitter)


def _analytic_box(lo, hi, jitter=1):
    """The oracle's seeding of the equivalent built-in box liquid (grid units)."""
    text = _one(_box("m", lo, hi))
    d = scenes.from_json(text, jitter=jitter, seed=0)
    o = OracleSim(d)
    o.init()
    pos, _, _ = o.get_particles()
    o.close()
    return pos


def test_box_mesh_emits_the_built_in_box_seeding():
    lo, hi = (1, 1, 1), (12.8, 19.2, 31)
    pos, vel = scenes.emit(_one(_box("m", lo, hi)), 0, solver=_solver())
    assert np.array_equal(pos, _analytic_box(lo, hi))
    assert np.all(vel == 0)


def test_rotated_box_is_the_box_with_its_extents_swapped():
    # a 6 x 10 x 8 box centred at (16.5, 14.5, 12), turned 90 degrees about z:
    # 10 x 6 x 8 around the same centre (faces on integer planes, lattice points
    # a quarter cell away from them)
    c = np.array([16.5, 14.5, 12.0])
    half = np.array([3.0, 5.0, 4.0])
    m = _box("m", tuple(-half), tuple(half))
    text = _one(m, {"translation": dict(zip("xyz", c)), "rotation": {"x": 0, "y": 0, "z": 90}})
    pos, _ = scenes.emit(text, 0, solver=_solver(jitter=0))
    ref = _analytic_box(tuple(c - half[[1, 0, 2]]), tuple(c + half[[1, 0, 2]]), jitter=0)
    assert len(pos) == len(ref) > 0
    assert np.array_equal(pos[np.lexsort(pos.T)], ref[np.lexsort(ref.T)])


def test_sphere_generator_is_an_inscribed_polyhedron():
    c, r = np.array([15.3, 16.2, 14.7]), 7.5
    text = _one({"id": "m", "mesh_gen": "sphere", "center": dict(zip("xyz", c)), "radius": r})
    pos, _ = scenes.emit(text, 0, solver=_solver())
    g = pos.astype(np.float64) * N
    dist = np.linalg.norm(g - c, axis=1)
    assert dist.max() < r                                  # vertices on the sphere: inside it
    # every lattice point deeper than the 20 x 20 tessellation's sagitta is kept
    pall, _ = scenes.emit(_one(_box("m", (1, 1, 1), (31, 31, 31))), 0, solver=_solver())
    ga = pall.astype(np.float64) * N
    deep = np.linalg.norm(ga - c, axis=1) < r * np.cos(np.pi / 20) ** 2
    kept = {tuple(p) for p in np.round(g, 5)}
    assert all(tuple(p) in kept for p in np.round(ga[deep], 5))
    assert len(pos) > 0.9 * 4.0 / 3.0 * np.pi * r ** 3 * 8


def _winding(points, verts, tris):
    """Generalized winding number (Jacobson et al. 2013): sum of the signed solid
    angles of the triangles, over 4 pi (Van Oosterom and Strackee's formula)."""
    w = np.zeros(len(points))
    for t in tris:
        a, b, c = (verts[t[k]] - points for k in range(3))
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        det = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la + \
            np.einsum("ij,ij->i", c, a) * lb
        w += 2.0 * np.arctan2(det, den)
    return w / (4.0 * np.pi)


def _l_prism(path):
    """An L-shaped prism (non-convex), outward quads, written as an OBJ file."""
    ring = [(4, 4), (20, 4), (20, 10), (11, 10), (11, 26), (4, 26)]
    z0, z1 = 5.0, 27.0
    verts = [(x, y, z0) for x, y in ring] + [(x, y, z1) for x, y in ring]
    n = len(ring)
    faces = []
    for k in range(n):
        a, b = k, (k + 1) % n
        faces.append((a, b, n + b, n + a))
    # caps as triangles (fan from a convex corner is not valid here: split by hand)
    bottom = [(0, 5, 4), (0, 4, 3), (0, 3, 2), (0, 2, 1)]
    top = [(n + a, n + b, n + c) for a, c, b in bottom]
    with open(path, "w") as f:
        for v in verts:
            f.write("v %g %g %g\n" % v)
        for q in faces:
            f.write("f " + " ".join(str(i + 1) for i in q) + "\n")
        for t in bottom + top:
            f.write("f " + " ".join("%d/%d/%d" % (i + 1, i + 1, i + 1) for i in t) + "\n")
    tris = [(q[0], q[1], q[2]) for q in faces] + [(q[0], q[2], q[3]) for q in faces] + bottom + top
    return np.array(verts, np.float64), tris


def test_obj_mesh_against_the_winding_number(tmp_path):
    verts, tris = _l_prism(tmp_path / "l.obj")
    text = _one({"id": "m", "file": "l.obj"})
    pos, _ = scenes.emit(text, 0, base_dir=str(tmp_path), solver=_solver())
    pall, _ = scenes.emit(_one(_box("m", (1, 1, 1), (31, 31, 31))), 0, solver=_solver())
    g = pall.astype(np.float64) * N
    w = _winding(g, verts, tris)
    inside = np.abs(w - 1.0) < 1e-3
    assert np.all((np.abs(w) < 1e-3) | inside)             # every point clearly in or out
    ref = pall[inside]
    assert np.array_equal(np.sort(pos.view("V12"), axis=0), np.sort(ref.view("V12"), axis=0))
    # the same shape translated: the winding set moves with it
    text2 = _one({"id": "m", "file": "l.obj"}, {"translation": {"x": 2, "y": 0, "z": -1}})
    pos2, _ = scenes.emit(text2, 0, base_dir=str(tmp_path), solver=_solver())
    w2 = _winding(g, verts + np.array([2.0, 0.0, -1.0]), tris)
    assert np.array_equal(np.sort(pos2.view("V12"), axis=0), np.sort(pall[np.abs(w2 - 1) < 1e-3].view("V12"), axis=0))


def test_frame_ranges_and_persistence():
    m = _box("m", (2, 2, 2), (10, 10, 10))
    def count(frame, **kw):
        return len(scenes.emit(_one(m, **kw), frame, solver=_solver())[0])
    full = count(0)
    assert full > 0
    # one frame at offset 0: frame 0 only (a one-shot liquid)
    assert [count(f) for f in range(3)] == [full, 0, 0]
    # offset 2: from frame 2, once; pre_persist also before; post_persist also after
    assert [count(f, offset=2) for f in range(5)] == [0, 0, full, 0, 0]
    assert [count(f, offset=2, pre=True) for f in range(4)] == [full, full, full, 0]
    assert [count(f, offset=2, post=True) for f in range(5)] == [0, 0, full, full, full]


def test_interpolated_transform_and_animated_sequence():
    m = _box("m", (2, 2, 2), (8, 8, 8))
    # two transform frames 4 cells apart along x, frame interval 2: frame 1 is
    # halfway (the box moved 2 cells)
    doc = _doc([m], [{"id": "a"}, {"id": "b", "translation": {"x": 4, "y": 0, "z": 0}}],
               [_geom("g", ["m", "m"], ["a", "b"], interval=2)], [{"geom": "g", "type": "liquid"}])
    p1, _ = scenes.emit(doc, 1, solver=_solver(jitter=0))
    ref = _analytic_box((4, 2, 2), (10, 8, 8), jitter=0)
    assert np.array_equal(p1, ref)
    # an animated sequence of two meshes: frame k takes mesh k
    doc2 = _doc([_box("m0", (2, 2, 2), (8, 8, 8)), _box("m1", (12, 2, 2), (18, 8, 8))],
                [{"id": "t"}], [_geom("g", None, ["t", "t"], kind="animated_mesh", anim="s")],
                [{"geom": "g", "type": "liquid"}], animated=[{"id": "s", "frames": ["m0", "m1"]}])
    q0, _ = scenes.emit(doc2, 0, solver=_solver(jitter=0))
    q1, _ = scenes.emit(doc2, 1, solver=_solver(jitter=0))
    assert np.array_equal(q0, _analytic_box((2, 2, 2), (8, 8, 8), jitter=0))
    assert np.array_equal(q1, _analytic_box((12, 2, 2), (18, 8, 8), jitter=0))


def test_liquid_velocity_and_refusals(tmp_path):
    m = _box("m", (2, 2, 2), (8, 8, 8))
    doc = _doc([m], [{"id": "t"}], [_geom("g", ["m"], ["t"])],
               [{"geom": "g", "type": "liquid", "velocity": {"x": 0.5, "y": -1, "z": 2}}])
    _, vel = scenes.emit(doc, 0, solver=_solver())
    assert np.all(vel == np.array([0.5, -1, 2], np.float32))
    solid = _doc([m], [{"id": "t"}], [_geom("g", ["m"], ["t"])], [{"geom": "g", "type": "solid"}])
    with pytest.raises(ArielError) as e:
        scenes.emit(solid, 0, solver=_solver())
    assert e.value.code == -6
    missing = _doc([{"id": "m", "file": "nope.obj"}], [{"id": "t"}], [_geom("g", ["m"], ["t"])],
                   [{"geom": "g", "type": "liquid"}])
    with pytest.raises(ArielError) as e:
        scenes.emit(missing, 0, base_dir=str(tmp_path), solver=_solver())
    assert e.value.code == -7


def _torus(path, R=9.0, r=3.5, nu=32, nv=16):
    """A torus of nu x nv quads, written as an OBJ file; the reference tests a
    quad as the triangles (0, 1, 2) and (0, 3, 2), listed here with consistent
    orientation for the winding number."""
    a = 2 * np.pi * np.arange(nu)[:, None] / nu
    b = 2 * np.pi * np.arange(nv)[None, :] / nv
    verts = np.stack([(R + r * np.cos(b)) * np.cos(a), np.broadcast_to(r * np.sin(b), (nu, nv)),
                      (R + r * np.cos(b)) * np.sin(a)], -1).reshape(-1, 3)
    quads = []
    for i in range(nu):
        for j in range(nv):
            quads.append((i * nv + j, ((i + 1) % nu) * nv + j, ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv))
    with open(path, "w") as f:
        for v in verts:
            f.write("v %.6f %.6f %.6f\n" % tuple(v))
        for q in quads:
            f.write("f " + " ".join(str(i + 1) for i in q) + "\n")
    verts = np.loadtxt(path, usecols=(1, 2, 3), max_rows=len(verts)).astype(np.float32).astype(np.float64)
    tris = [(q[0], q[1], q[2]) for q in quads] + [(q[0], q[2], q[3]) for q in quads]
    return verts, tris


def test_rotated_scaled_torus_against_the_winding_number(tmp_path):
    # a 1024-triangle torus turned about all three axes and scaled unevenly:
    # the binned ray test must keep exactly the lattice points of winding number 1
    verts, tris = _torus(tmp_path / "t.obj")
    t, r, s = np.array([16.3, 15.1, 16.7]), np.array([23.0, -41.0, 67.0]), np.array([1.2, 0.8, 1.05])
    text = _one({"id": "m", "file": "t.obj"}, {"translation": dict(zip("xyz", t)),
                                               "rotation": dict(zip("xyz", r)), "scale": dict(zip("xyz", s))})
    pos, _ = scenes.emit(text, 0, base_dir=str(tmp_path), solver=_solver())
    def rot(axis, deg):
        c, sn = np.cos(np.radians(deg)), np.sin(np.radians(deg))
        m = np.eye(3)
        i, j = [(1, 2), (2, 0), (0, 1)][axis]
        m[i, i], m[i, j], m[j, i], m[j, j] = c, -sn, sn, c
        return m
    R = rot(2, r[2]) @ rot(1, r[1]) @ rot(0, r[0])          # T Rz Ry Rx S
    world = verts * s @ R.T + t
    pall, _ = scenes.emit(_one(_box("m", (1, 1, 1), (31, 31, 31))), 0, solver=_solver())
    g = pall.astype(np.float64) * N
    near = np.all((g > world.min(0) - 1) & (g < world.max(0) + 1), axis=1)   # outside: winding 0
    w = np.zeros(len(g))
    w[near] = np.abs(_winding(g[near], world, tris))      # the quads face inwards: -1 inside
    inside = np.abs(w - 1.0) < 1e-3
    assert np.all((np.abs(w) < 1e-3) | inside)
    assert inside.sum() > 1000
    assert np.array_equal(np.sort(pos.view("V12"), axis=0), np.sort(pall[inside].view("V12"), axis=0))


@pytest.mark.parametrize("r", [(23.0, 0.0, 0.0), (0.0, -41.0, 0.0), (23.0, -41.0, 67.0), (90.0, 45.0, 0.0)])
def test_box_turned_about_every_axis(r):
    # the ray from a lattice point is tilted in the box's frame: only hits in
    # front of the point count, so the emitted set is exactly the lattice points
    # inside the turned box (R = Rz Ry Rx, degrees)
    h, c = np.array([5.0, 3.0, 7.0]), np.array([16.0, 16.0, 16.0])
    text = _one(_box("m", tuple(-h), tuple(h)), {"translation": dict(zip("xyz", c)), "rotation": dict(zip("xyz", r))})
    pos, _ = scenes.emit(text, 0, solver=_solver())
    def rot(axis, deg):
        cs, sn = np.cos(np.radians(deg)), np.sin(np.radians(deg))
        m = np.eye(3)
        i, j = [(1, 2), (2, 0), (0, 1)][axis]
        m[i, i], m[i, j], m[j, i], m[j, j] = cs, -sn, sn, cs
        return m
    R = rot(2, r[2]) @ rot(1, r[1]) @ rot(0, r[0])
    pall, _ = scenes.emit(_one(_box("m", (1, 1, 1), (31, 31, 31))), 0, solver=_solver())
    local = (pall.astype(np.float64) * N - c) @ R
    margin = np.min(h - np.abs(local), axis=1)
    clear = np.abs(margin) > 1e-4                          # leave out points on a face
    ref = pall[clear & (margin > 0)]
    assert len(ref) > 1000 and (~clear).sum() < 10
    near = {tuple(p) for p in pall[~clear].view(np.uint32).tolist()}
    got = np.array([p for p in pos if tuple(p.view(np.uint32).tolist()) not in near], np.float32).reshape(-1, 3)
    assert np.array_equal(np.sort(got.view("V12"), axis=0), np.sort(ref.view("V12"), axis=0))
