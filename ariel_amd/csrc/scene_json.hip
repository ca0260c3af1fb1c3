// scene_json.hip — the reference's JSON scene format (src/scene/sceneloader.cpp,
// schema of SceneLoader::SceneLoader and its Load* methods) read into an
// ariel_scene_desc. Host code only.
//
// Supported subset: "settings" (density, step_size, dim), "globalforces",
// "meshfiles" generated by mesh_gen "box" (point0, point1) or "sphere"
// (center, radius), "transforms" with translation and scale (rotation zero),
// "geoms" of type "mesh" with one geom frame and one transform frame,
// emitting once ("post_persist": false, or absent), and "sim" entries of type
// "liquid" with an optional velocity. Geometry is in grid-cell units (world =
// p * maxd, scene.cpp:439) and becomes unit-box coordinates here. Meshes read
// from .obj files, rotations, animated meshes, persistent emitters and solid
// geometry are rejected with ARIEL_ERR_UNSUPPORTED; "camera" and the output
// paths are ignored. Transforms apply as T R S (utilities.inl:191-199), so
// with no rotation a point maps to s * p + t. The reference seeds liquids by
// ray-parity tests against tessellated meshes (scene.cpp:436-455); here boxes
// and spheres are analytic, so particles within a tessellation facet of a
// sphere's surface may differ.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ariel.h"
#include "ariel_internal.h"

namespace {

struct JValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;   // in document order

    const JValue* get(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const char* s;
    size_t n, i = 0;
    std::string err;

    void ws() {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char* w) {
        const size_t l = std::strlen(w);
        if (i + l <= n && std::strncmp(s + i, w, l) == 0) {
            i += l;
            return true;
        }
        return false;
    }
    bool string(std::string& out) {
        if (i >= n || s[i] != '"') return fail("expected a string");
        ++i;
        while (i < n && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= n) return fail("unterminated escape");
                const char e = s[i++];
                switch (e) {
                    case '"': c = '"'; break;
                    case '\\': c = '\\'; break;
                    case '/': c = '/'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'n': c = '\n'; break;
                    case 'r': c = '\r'; break;
                    case 't': c = '\t'; break;
                    case 'u':
                        if (i + 4 > n) return fail("bad \\u escape");
                        i += 4;     // identifiers in scenes are ASCII; keep a placeholder
                        c = '?';
                        break;
                    default: return fail("bad escape");
                }
            }
            out.push_back(c);
        }
        if (i >= n) return fail("unterminated string");
        ++i;
        return true;
    }
    bool value(JValue& v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (i >= n) return fail("unexpected end");
        const char c = s[i];
        if (c == '{') {
            v.kind = JValue::Object;
            ++i;
            ws();
            if (i < n && s[i] == '}') {
                ++i;
                return true;
            }
            for (;;) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (i >= n || s[i] != ':') return fail("expected ':'");
                ++i;
                JValue child;
                if (!value(child, depth + 1)) return false;
                v.obj.emplace_back(std::move(key), std::move(child));
                ws();
                if (i < n && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < n && s[i] == '}') {
                    ++i;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JValue::Array;
            ++i;
            ws();
            if (i < n && s[i] == ']') {
                ++i;
                return true;
            }
            for (;;) {
                JValue child;
                if (!value(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (i < n && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < n && s[i] == ']') {
                    ++i;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::String;
            return string(v.str);
        }
        if (lit("true")) {
            v.kind = JValue::Bool;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JValue::Bool;
            return true;
        }
        if (lit("null")) return true;
        if (c == '-' || (c >= '0' && c <= '9')) {
            char* end = nullptr;
            v.num = std::strtod(s + i, &end);
            if (end == s + i) return fail("bad number");
            v.kind = JValue::Number;
            i = (size_t)(end - s);
            return true;
        }
        return fail("unexpected character");
    }
};

// jsoncpp's asFloat / asInt / asBool / asString on a possibly missing member
float as_float(const JValue* v) { return v && v->kind == JValue::Number ? (float)v->num : 0.0f; }
int as_int(const JValue* v) { return v && v->kind == JValue::Number ? (int)v->num : 0; }
bool as_bool(const JValue* v) { return v && (v->kind == JValue::Bool ? v->b : (v->kind == JValue::Number && v->num != 0)); }
std::string as_string(const JValue* v) { return v && v->kind == JValue::String ? v->str : std::string(); }
size_t size_of(const JValue* v) {
    if (!v) return 0;
    return v->kind == JValue::Array ? v->arr.size() : (v->kind == JValue::Object ? v->obj.size() : 0);
}
const JValue* at(const JValue* v, size_t k) {
    return v && v->kind == JValue::Array && k < v->arr.size() ? &v->arr[k] : nullptr;
}
bool vec3(const JValue* v, float out[3]) {
    if (!v || v->kind != JValue::Object) return false;
    out[0] = as_float(v->get("x"));
    out[1] = as_float(v->get("y"));
    out[2] = as_float(v->get("z"));
    return true;
}

struct Mesh {
    int kind = 0;   // ARIEL_LIQUID_BOX or _SPHERE
    float p0[3] = {0, 0, 0}, p1[3] = {0, 0, 0}, c[3] = {0, 0, 0}, r = 0;
};
struct Xform {
    float t[3] = {0, 0, 0}, s[3] = {1, 1, 1};
    bool rotated = false;
};
struct Geom {
    std::string mesh, xform;
};

}  // namespace

extern "C" int ariel_desc_from_json(const char* text, ariel_scene_desc* out) {
    using ariel::set_last_error;
    if (!text || !out) {
        set_last_error("desc_from_json: null argument");
        return ARIEL_ERR_INVALID;
    }
    Parser P{text, std::strlen(text)};
    JValue root;
    if (!P.value(root, 0)) {
        set_last_error("desc_from_json: " + P.err);
        return ARIEL_ERR_INVALID;
    }
    P.ws();
    if (P.i != P.n || root.kind != JValue::Object) {
        set_last_error("desc_from_json: the document is not one JSON object");
        return ARIEL_ERR_INVALID;
    }
    auto unsupported = [](const std::string& m) {
        set_last_error("desc_from_json: " + m);
        return ARIEL_ERR_UNSUPPORTED;
    };
    ariel_scene_desc d;
    int rc = ariel_desc_defaults(&d);
    if (rc != ARIEL_OK) return rc;
    // settings[0] (LoadSettings, sceneloader.cpp:411-443)
    float density = 0.5f, step = 0.005f;
    int dims[3] = {32, 32, 32};
    if (const JValue* st = at(root.get("settings"), 0)) {
        if (st->get("density")) density = as_float(st->get("density"));
        if (st->get("step_size")) step = as_float(st->get("step_size"));
        if (const JValue* dim = st->get("dim")) {
            dims[0] = as_int(dim->get("x"));
            dims[1] = as_int(dim->get("y"));
            dims[2] = as_int(dim->get("z"));
        }
    }
    for (int a = 0; a < 3; ++a) d.dims[a] = dims[a];
    d.density = density;
    d.dt = step;
    const float maxd = (float)std::max(dims[0], std::max(dims[1], dims[2]));
    if (!(maxd > 0.0f)) {
        set_last_error("desc_from_json: settings.dim must be positive");
        return ARIEL_ERR_INVALID;
    }
    // globalforces (LoadGlobalForces, :400-409)
    d.n_forces = 0;
    if (const JValue* gf = root.get("globalforces")) {
        for (size_t k = 0; k < size_of(gf); ++k) {
            if (d.n_forces >= ARIEL_MAX_FORCES) return unsupported("too many global forces");
            vec3(at(gf, k), d.forces[d.n_forces]);
            ++d.n_forces;
        }
    }
    // transforms (LoadGeomTransforms, :305-344)
    std::map<std::string, Xform> xforms;
    if (const JValue* tr = root.get("transforms")) {
        for (size_t k = 0; k < size_of(tr); ++k) {
            const JValue* t = at(tr, k);
            if (!t || !t->get("id")) continue;          // the reference skips it
            const std::string id = as_string(t->get("id"));
            if (xforms.count(id)) continue;             // first definition wins
            Xform x;
            if (t->get("translation")) vec3(t->get("translation"), x.t);
            if (t->get("scale")) vec3(t->get("scale"), x.s);
            float rot[3] = {0, 0, 0};
            if (t->get("rotation")) vec3(t->get("rotation"), rot);
            x.rotated = rot[0] != 0.0f || rot[1] != 0.0f || rot[2] != 0.0f;
            xforms[id] = x;
        }
    }
    // meshfiles (LoadMeshFiles, :346-398)
    std::map<std::string, Mesh> meshes;
    std::map<std::string, bool> from_file;
    if (const JValue* mf = root.get("meshfiles")) {
        for (size_t k = 0; k < size_of(mf); ++k) {
            const JValue* m = at(mf, k);
            if (!m || !m->get("id")) continue;
            const std::string id = as_string(m->get("id"));
            if (meshes.count(id) || from_file.count(id)) continue;
            if (m->get("file")) {
                from_file[id] = true;
                continue;
            }
            const std::string gen = as_string(m->get("mesh_gen"));
            Mesh g;
            if (gen == "box") {
                g.kind = ARIEL_LIQUID_BOX;
                vec3(m->get("point0"), g.p0);
                vec3(m->get("point1"), g.p1);
            } else if (gen == "sphere") {
                g.kind = ARIEL_LIQUID_SPHERE;
                vec3(m->get("center"), g.c);
                g.r = as_float(m->get("radius"));
            } else {
                from_file[id] = true;   // an empty mesh in the reference
                continue;
            }
            meshes[id] = g;
        }
    }
    if (root.get("animatedmeshes") && size_of(root.get("animatedmeshes")) > 0)
        return unsupported("animated meshes");
    // geoms (LoadGeom, :173-253)
    std::map<std::string, Geom> geoms;
    if (const JValue* gs = root.get("geoms")) {
        for (size_t k = 0; k < size_of(gs); ++k) {
            const JValue* g = at(gs, k);
            if (!g || !g->get("id")) continue;
            const std::string id = as_string(g->get("id"));
            if (geoms.count(id)) continue;
            const std::string type = as_string(g->get("type"));
            if (type == "animated_mesh") return unsupported("geom \"" + id + "\": animated mesh");
            if (type != "mesh") continue;
            const size_t frames = std::min(size_of(g->get("geom_frames")), size_of(g->get("transform_frames")));
            if (frames != 1) return unsupported("geom \"" + id + "\": " + std::to_string(frames) +
                                                " frames (one is supported)");
            if (as_bool(g->get("post_persist")))
                return unsupported("geom \"" + id + "\": a persistent emitter (post_persist)");
            Geom G;
            G.mesh = as_string(at(g->get("geom_frames"), 0));
            G.xform = as_string(at(g->get("transform_frames"), 0));
            geoms[id] = G;
        }
    }
    // sim (LoadSim, :153-171)
    d.n_liquids = 0;
    if (const JValue* sim = root.get("sim")) {
        for (size_t k = 0; k < size_of(sim); ++k) {
            const JValue* s = at(sim, k);
            if (!s) continue;
            const std::string type = as_string(s->get("type"));
            const std::string gid = as_string(s->get("geom"));
            if (type == "solid") return unsupported("solid geometry (geom \"" + gid + "\")");
            if (type != "liquid") continue;
            auto git = geoms.find(gid);
            if (git == geoms.end()) {
                set_last_error("desc_from_json: sim entry names unknown geom \"" + gid + "\"");
                return ARIEL_ERR_INVALID;
            }
            if (from_file.count(git->second.mesh))
                return unsupported("geom \"" + gid + "\": meshes read from files");
            auto mit = meshes.find(git->second.mesh);
            auto xit = xforms.find(git->second.xform);
            if (mit == meshes.end() || xit == xforms.end()) {
                set_last_error("desc_from_json: geom \"" + gid + "\" names an unknown mesh or transform");
                return ARIEL_ERR_INVALID;
            }
            const Mesh& m = mit->second;
            const Xform& x = xit->second;
            if (x.rotated) return unsupported("geom \"" + gid + "\": rotated transform");
            if (d.n_liquids >= ARIEL_MAX_LIQUIDS) return unsupported("too many liquids");
            ariel_liquid& lq = d.liquids[d.n_liquids];
            std::memset(&lq, 0, sizeof(lq));
            lq.kind = m.kind;
            if (m.kind == ARIEL_LIQUID_BOX) {
                for (int a = 0; a < 3; ++a) {
                    const float q0 = m.p0[a] * x.s[a] + x.t[a], q1 = m.p1[a] * x.s[a] + x.t[a];
                    lq.lo[a] = std::min(q0, q1) / maxd;
                    lq.hi[a] = std::max(q0, q1) / maxd;
                }
            } else {
                if (!(x.s[0] == x.s[1] && x.s[1] == x.s[2]))
                    return unsupported("geom \"" + gid + "\": non-uniform scale of a sphere");
                for (int a = 0; a < 3; ++a) lq.center[a] = (m.c[a] * x.s[a] + x.t[a]) / maxd;
                lq.radius = m.r * std::fabs(x.s[0]) / maxd;
            }
            if (s->get("velocity")) vec3(s->get("velocity"), lq.velocity);
            ++d.n_liquids;
        }
    }
    *out = d;
    return ARIEL_OK;
}

namespace ariel {

// The whole reference scene (SceneLoader, sceneloader.cpp:11-474): settings and
// forces into d, and every geometry node into geom: mesh files (OBJ, read from
// base_dir, or the box and sphere generators tessellated as the reference
// does), transforms with rotation, animated mesh sequences (a mesh per frame;
// the reference interpolates between consecutive meshes by the fraction of the
// frame number, which is 0 on the integer frames a simulation steps through),
// geoms with their frame ranges and persistence, and the liquids of "sim".
// Solid geometry is refused (no solid constraints in this build).
int scene_from_json(const char* text, const char* base_dir, ariel_scene_desc* d, SceneGeom* geom,
                    std::string& err) {
    Parser P{text, std::strlen(text)};
    JValue root;
    if (!P.value(root, 0)) {
        err = P.err;
        return ARIEL_ERR_INVALID;
    }
    P.ws();
    if (P.i != P.n || root.kind != JValue::Object) {
        err = "the document is not one JSON object";
        return ARIEL_ERR_INVALID;
    }
    int rc = ariel_desc_defaults(d);
    if (rc != ARIEL_OK) return rc;
    float density = 0.5f, step = 0.005f;
    int dims[3] = {32, 32, 32};
    if (const JValue* st = at(root.get("settings"), 0)) {
        if (st->get("density")) density = as_float(st->get("density"));
        if (st->get("step_size")) step = as_float(st->get("step_size"));
        if (const JValue* dim = st->get("dim")) {
            dims[0] = as_int(dim->get("x"));
            dims[1] = as_int(dim->get("y"));
            dims[2] = as_int(dim->get("z"));
        }
    }
    for (int a = 0; a < 3; ++a) d->dims[a] = dims[a];
    d->density = density;
    d->dt = step;
    d->n_forces = 0;
    if (const JValue* gf = root.get("globalforces")) {
        for (size_t k = 0; k < size_of(gf); ++k) {
            if (d->n_forces >= ARIEL_MAX_FORCES) {
                err = "too many global forces";
                return ARIEL_ERR_UNSUPPORTED;
            }
            vec3(at(gf, k), d->forces[d->n_forces]);
            ++d->n_forces;
        }
    }
    d->n_liquids = 0;
    SceneGeom& G = *geom;
    G = SceneGeom();
    std::map<std::string, int> xform_id, mesh_id, anim_id, geom_id;
    if (const JValue* tr = root.get("transforms")) {
        for (size_t k = 0; k < size_of(tr); ++k) {
            const JValue* t = at(tr, k);
            if (!t || !t->get("id")) continue;
            const std::string id = as_string(t->get("id"));
            if (xform_id.count(id)) continue;
            GeomXform x;
            if (t->get("translation")) vec3(t->get("translation"), x.t);
            if (t->get("rotation")) vec3(t->get("rotation"), x.r);
            if (t->get("scale")) vec3(t->get("scale"), x.s);
            xform_id[id] = (int)G.xforms.size();
            G.xforms.push_back(x);
        }
    }
    const std::string base = base_dir && *base_dir ? std::string(base_dir) + "/" : std::string();
    if (const JValue* mf = root.get("meshfiles")) {
        for (size_t k = 0; k < size_of(mf); ++k) {
            const JValue* m = at(mf, k);
            if (!m || !m->get("id")) continue;
            const std::string id = as_string(m->get("id"));
            if (mesh_id.count(id)) continue;
            GeomMesh g;
            if (m->get("file")) {
                if (!read_obj_mesh(base + as_string(m->get("file")), g, err)) return ARIEL_ERR_IO;
            } else {
                const std::string gen = as_string(m->get("mesh_gen"));
                if (gen == "box") {
                    float p0[3], p1[3];
                    vec3(m->get("point0"), p0);
                    vec3(m->get("point1"), p1);
                    cube_mesh(g, p0, p1);
                } else if (gen == "sphere") {
                    float c[3];
                    vec3(m->get("center"), c);
                    sphere_mesh(g, c, as_float(m->get("radius")), 20);
                }
            }
            mesh_id[id] = (int)G.meshes.size();
            G.meshes.push_back(std::move(g));
        }
    }
    std::vector<std::vector<int>> anims;
    if (const JValue* am = root.get("animatedmeshes")) {
        for (size_t k = 0; k < size_of(am); ++k) {
            const JValue* a = at(am, k);
            if (!a || !a->get("id")) continue;
            const std::string id = as_string(a->get("id"));
            if (anim_id.count(id)) continue;
            std::vector<int> seq;
            for (size_t f = 0; f < size_of(a->get("frames")); ++f) {
                auto it = mesh_id.find(as_string(at(a->get("frames"), f)));
                if (it == mesh_id.end()) {
                    err = "animated mesh \"" + id + "\" names an unknown mesh file";
                    return ARIEL_ERR_INVALID;
                }
                seq.push_back(it->second);
            }
            anim_id[id] = (int)anims.size();
            anims.push_back(seq);
        }
    }
    if (const JValue* gs = root.get("geoms")) {
        for (size_t k = 0; k < size_of(gs); ++k) {
            const JValue* g = at(gs, k);
            if (!g || !g->get("id")) continue;
            const std::string id = as_string(g->get("id"));
            if (geom_id.count(id)) continue;
            GeomNode n;
            n.pre_persist = as_bool(g->get("pre_persist"));
            n.post_persist = as_bool(g->get("post_persist"));
            n.interval = as_int(g->get("frame_interval"));
            n.offset = as_int(g->get("frame_offset"));
            if (n.interval <= 0) {
                err = "geom \"" + id + "\": frame_interval must be positive";
                return ARIEL_ERR_INVALID;
            }
            const std::string type = as_string(g->get("type"));
            std::vector<int> meshes;
            if (type == "mesh") {
                for (size_t f = 0; f < size_of(g->get("geom_frames")); ++f) {
                    auto it = mesh_id.find(as_string(at(g->get("geom_frames"), f)));
                    if (it == mesh_id.end()) {
                        err = "geom \"" + id + "\" names an unknown mesh file";
                        return ARIEL_ERR_INVALID;
                    }
                    meshes.push_back(it->second);
                }
            } else if (type == "animated_mesh") {
                auto it = anim_id.find(as_string(g->get("anim_sequence")));
                if (it == anim_id.end()) {
                    err = "geom \"" + id + "\" names an unknown animated mesh";
                    return ARIEL_ERR_INVALID;
                }
                meshes = anims[it->second];
            } else {
                continue;
            }
            std::vector<int> xf;
            for (size_t f = 0; f < size_of(g->get("transform_frames")); ++f) {
                auto it = xform_id.find(as_string(at(g->get("transform_frames"), f)));
                if (it == xform_id.end()) {
                    err = "geom \"" + id + "\" names an unknown transform";
                    return ARIEL_ERR_INVALID;
                }
                xf.push_back(it->second);
            }
            const size_t frames = std::min(meshes.size(), xf.size());
            if (frames == 0) {
                err = "geom \"" + id + "\" has no frames";
                return ARIEL_ERR_INVALID;
            }
            n.mesh.assign(meshes.begin(), meshes.begin() + frames);
            n.xform.assign(xf.begin(), xf.begin() + frames);
            geom_id[id] = (int)G.geoms.size();
            G.geoms.push_back(n);
        }
    }
    if (const JValue* sim = root.get("sim")) {
        for (size_t k = 0; k < size_of(sim); ++k) {
            const JValue* s = at(sim, k);
            if (!s) continue;
            const std::string type = as_string(s->get("type"));
            const std::string gid = as_string(s->get("geom"));
            if (type == "solid") {
                err = "solid geometry (geom \"" + gid + "\"): no solid constraints in this build";
                return ARIEL_ERR_UNSUPPORTED;
            }
            if (type != "liquid") continue;
            auto git = geom_id.find(gid);
            if (git == geom_id.end()) {
                err = "sim entry names unknown geom \"" + gid + "\"";
                return ARIEL_ERR_INVALID;
            }
            GeomLiquid lq;
            lq.geom = git->second;
            lq.velocity[0] = lq.velocity[1] = lq.velocity[2] = 0.0f;
            if (s->get("velocity")) vec3(s->get("velocity"), lq.velocity);
            G.liquids.push_back(lq);
        }
    }
    return ARIEL_OK;
}

}  // namespace ariel
