"""The reference's JSON scene format (src/scene/sceneloader.cpp) read into a
descriptor: the schema's dam break and sphere give byte-identical descriptors
to the hand-built scenes, transforms map points as s * p + t, and every
unsupported or malformed input fails loudly (CPU only)."""
import ctypes
import json

import pytest

from ariel_amd import ArielError, _abi, scenes


def _scene(n=32, liquids=(), forces=((0.0, -9.8, 0.0),), step=1.0 / 240.0, extra=None):
    meshfiles, transforms, geoms, sim = [], [], [], []
    for i, lq in enumerate(liquids):
        mid, tid, gid = f"m{i}", f"t{i}", f"g{i}"
        meshfiles.append(dict(id=mid, **lq["mesh"]))
        transforms.append(dict(id=tid, **lq.get("transform", {})))
        geoms.append({"id": gid, "type": "mesh", "geom_frames": [mid], "transform_frames": [tid],
                      "pre_persist": False, "post_persist": lq.get("post_persist", False),
                      "frame_interval": 1, "frame_offset": 0})
        s = {"geom": gid, "type": lq.get("type", "liquid")}
        if "velocity" in lq:
            s["velocity"] = dict(zip("xyz", lq["velocity"]))
        sim.append(s)
    doc = {"settings": [{"density": 0.5, "step_size": step, "dim": {"x": n, "y": n, "z": n}}],
           "camera": [{"fovx": 45}],
           "globalforces": [dict(zip("xyz", f)) for f in forces],
           "transforms": transforms, "meshfiles": meshfiles, "geoms": geoms, "sim": sim}
    doc.update(extra or {})
    return json.dumps(doc)


def _box(p0, p1, **kw):
    return dict(mesh={"mesh_gen": "box", "point0": dict(zip("xyz", p0)), "point1": dict(zip("xyz", p1))}, **kw)


def _sphere(c, r, **kw):
    return dict(mesh={"mesh_gen": "sphere", "center": dict(zip("xyz", c)), "radius": r}, **kw)


def _fields(d):
    out = {}
    for name, _t in d._fields_:
        v = getattr(d, name)
        if name == "liquids":
            v = [bytes(v[i]) for i in range(d.n_liquids)]
        elif name == "forces":
            v = [tuple(v[i]) for i in range(d.n_forces)]
        elif hasattr(v, "__len__"):
            v = tuple(v)
        out[name] = v
    return out


SOLVER = dict(wall=1, jitter=1, seed=0, tol=1e-6, max_iters=500)


def test_dam_break_matches_the_hand_built_scene():
    n = 32
    text = _scene(n, [_box((0, 0, 0), (0.4 * n, 0.6 * n, n))])
    d = scenes.from_json(text, **SOLVER)
    assert _fields(d) == _fields(scenes.dam_break(n))


def test_sphere_with_transform_and_velocity():
    n = 128
    # sphere of radius 0.12 at (0.7, 0.8, 0.5) in unit coordinates, written as a
    # unit sphere at the origin scaled by 0.12 n and translated
    text = _scene(n, [_box((0, 0, 0), (0.4 * n, 0.6 * n, n)),
                      _sphere((0, 0, 0), 1.0, transform={"translation": dict(x=0.7 * n, y=0.8 * n, z=0.5 * n),
                                                         "scale": dict(x=0.12 * n, y=0.12 * n, z=0.12 * n)},
                              velocity=(0.0, -1.0, 0.0))])
    d = scenes.from_json(text, **SOLVER)
    ref = scenes.dam_break_sphere(n)
    got, want = _fields(d), _fields(ref)
    assert {k: v for k, v in got.items() if k != "liquids"} == {k: v for k, v in want.items() if k != "liquids"}
    assert d.liquids[0].kind == ref.liquids[0].kind and bytes(d.liquids[0]) == bytes(ref.liquids[0])
    s, r = d.liquids[1], ref.liquids[1]
    assert s.kind == _abi.LIQUID_SPHERE
    assert tuple(s.velocity) == (0.0, -1.0, 0.0)
    for a in range(3):
        assert s.center[a] == pytest.approx(r.center[a], rel=1e-6)
    assert s.radius == pytest.approx(r.radius, rel=1e-6)


def test_box_transform_is_scale_then_translate():
    n = 64
    text = _scene(n, [_box((-1, 2, 3), (1, 4, 5), transform={"translation": dict(x=10, y=0, z=-2),
                                                             "scale": dict(x=2, y=-1, z=1)})])
    lq = scenes.from_json(text).liquids[0]
    # x: [-2, 2] + 10 = [8, 12]; y: [-2, -4] -> [-4, -2]; z: [3, 5] - 2 = [1, 3]
    assert [v * n for v in lq.lo] == pytest.approx([8, -4, 1])
    assert [v * n for v in lq.hi] == pytest.approx([12, -2, 3])


def test_defaults_without_settings():
    d = scenes.from_json(json.dumps({"sim": []}))
    assert tuple(d.dims) == (32, 32, 32) and d.density == 0.5 and d.dt == pytest.approx(0.005)
    assert d.n_liquids == 0 and d.n_forces == 0


@pytest.mark.parametrize("case,code", [
    ({"transform": {"rotation": {"x": 0, "y": 30, "z": 0}}}, _abi.ERR_UNSUPPORTED),
    ({"type": "solid"}, _abi.ERR_UNSUPPORTED),
    ({"post_persist": True}, _abi.ERR_UNSUPPORTED),
    ({"mesh": {"file": "bunny.obj"}}, _abi.ERR_UNSUPPORTED),
])
def test_unsupported_scene_features_fail_loudly(case, code):
    lq = _box((0, 0, 0), (8, 8, 8))
    lq.update(case)
    with pytest.raises(ArielError) as e:
        scenes.from_json(_scene(32, [lq]))
    assert e.value.code == code


def test_nonuniform_sphere_scale_and_animation_are_unsupported():
    with pytest.raises(ArielError):
        scenes.from_json(_scene(32, [_sphere((8, 8, 8), 2, transform={"scale": dict(x=1, y=2, z=1)})]))
    with pytest.raises(ArielError) as e:
        scenes.from_json(_scene(32, [], extra={"animatedmeshes": [{"id": "a", "frames": ["m0"]}]}))
    assert e.value.code == _abi.ERR_UNSUPPORTED


@pytest.mark.parametrize("text", ['{"settings": [', '{"sim": [1 2]}', '[]', '{"a": tru}', ''])
def test_malformed_json_is_invalid(text):
    lib = _abi.load()
    d = _abi.SceneDesc()
    assert lib.ariel_desc_from_json(text.encode(), ctypes.byref(d)) == _abi.ERR_INVALID
    assert lib.ariel_last_error(None).startswith(b"desc_from_json")


def test_unknown_geom_is_invalid():
    doc = json.loads(_scene(32, [_box((0, 0, 0), (8, 8, 8))]))
    doc["sim"][0]["geom"] = "nope"
    with pytest.raises(ArielError) as e:
        scenes.from_json(json.dumps(doc))
    assert e.value.code == _abi.ERR_INVALID
