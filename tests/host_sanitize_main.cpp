// Host-code driver for the AddressSanitizer/UndefinedBehaviorSanitizer build of
// tests/test_host_sanitizers.py (test infrastructure, CPU only). It runs every
// library entry point whose work happens on the host -- the JSON scene reader
// (also on every prefix of a valid scene and on mutated bytes), the mesher, the
// descriptor checks, the creation and error paths that end without a GPU -- and
// the CPU oracle through initialisation, steps in both solver modes, stage runs,
// the operator, the preconditioners and the FP64 solve. The sanitizers abort on
// the first report; the driver itself checks return codes and prints "ok".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/ariel.h"
#include "../oracle/flip_oracle.h"

static_assert(sizeof(oracle_desc) == sizeof(ariel_scene_desc), "oracle_desc mirrors ariel_scene_desc");

namespace {

int g_fail = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what);
        ++g_fail;
    }
}

const char* kScene =
    "{\"settings\": [{\"density\": 0.5, \"step_size\": 0.004166667, \"dim\": {\"x\": 16, \"y\": 16, \"z\": 16}}],"
    " \"camera\": [{\"fovx\": 45}],"
    " \"globalforces\": [{\"x\": 0.0, \"y\": -9.8, \"z\": 0.0}],"
    " \"transforms\": [{\"id\": \"t0\", \"translation\": {\"x\": 1, \"y\": 2, \"z\": 3},"
    "                  \"scale\": {\"x\": 1, \"y\": 1, \"z\": 1}},"
    "                 {\"id\": \"t1\"}],"
    " \"meshfiles\": [{\"id\": \"m0\", \"mesh_gen\": \"box\", \"point0\": {\"x\": 0, \"y\": 0, \"z\": 0},"
    "                  \"point1\": {\"x\": 6, \"y\": 9, \"z\": 16}},"
    "                 {\"id\": \"m1\", \"mesh_gen\": \"sphere\", \"center\": {\"x\": 11, \"y\": 12, \"z\": 8},"
    "                  \"radius\": 2.5}],"
    " \"geoms\": [{\"id\": \"g0\", \"type\": \"mesh\", \"geom_frames\": [\"m0\"], \"transform_frames\": [\"t0\"],"
    "             \"pre_persist\": false, \"post_persist\": false, \"frame_interval\": 1, \"frame_offset\": 0},"
    "            {\"id\": \"g1\", \"type\": \"mesh\", \"geom_frames\": [\"m1\"], \"transform_frames\": [\"t1\"]}],"
    " \"sim\": [{\"geom\": \"g0\", \"type\": \"liquid\"},"
    "          {\"geom\": \"g1\", \"type\": \"liquid\", \"velocity\": {\"x\": 0, \"y\": -1, \"z\": 0}}]}";

void json_paths() {
    ariel_scene_desc d;
    check(ariel_desc_from_json(kScene, &d) == ARIEL_OK, "valid JSON scene");
    check(d.n_liquids == 2 && d.n_forces == 1 && d.dims[0] == 16, "JSON scene contents");
    check(ariel_desc_from_json(nullptr, &d) != ARIEL_OK, "null JSON text");
    check(ariel_desc_from_json(kScene, nullptr) != ARIEL_OK, "null JSON output");
    // every proper prefix is malformed or incomplete: an error, never a crash
    const std::string s(kScene);
    int ok_prefixes = 0;
    for (size_t n = 0; n < s.size(); ++n) {
        const std::string p = s.substr(0, n);
        if (ariel_desc_from_json(p.c_str(), &d) == ARIEL_OK) ++ok_prefixes;
    }
    check(ok_prefixes == 0, "no proper prefix of the scene parses");
    // byte mutations: any status, no memory error
    uint32_t x = 12345u;
    const char repl[] = "{}[]\",:0123456789.-eE \\tnul";
    for (int it = 0; it < 4000; ++it) {
        std::string m = s;
        for (int k = 0; k < 1 + it % 4; ++k) {
            x = x * 1664525u + 1013904223u;
            const size_t at = (x >> 8) % m.size();
            x = x * 1664525u + 1013904223u;
            const int op = (x >> 24) % 3;
            const char c = repl[(x >> 4) % (sizeof(repl) - 1)];
            if (op == 0) m[at] = c;
            else if (op == 1) m.insert(m.begin() + at, c);
            else m.erase(m.begin() + at);
        }
        (void)ariel_desc_from_json(m.c_str(), &d);
    }
    // deep nesting
    std::string deep(20000, '[');
    (void)ariel_desc_from_json(deep.c_str(), &d);
}

void mesher_paths(const char* dir) {
    const int M = 17;
    std::vector<float> phi((size_t)M * M * M);
    const std::string path = std::string(dir) + "/m.obj";
    int64_t nv = -1, nq = -1;
    for (int shape = 0; shape < 4; ++shape) {
        uint32_t x = 777u + shape;
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j)
                for (int k = 0; k < M; ++k) {
                    const float fx = i - 8.f, fy = j - 8.f, fz = k - 8.f;
                    float v = 0.f;
                    if (shape == 0) v = std::sqrt(fx * fx + fy * fy + fz * fz) - 5.3f;
                    if (shape == 1) {
                        const float q = std::sqrt(fx * fx + fz * fz) - 5.f;
                        v = std::sqrt(q * q + fy * fy) - 2.2f;
                    }
                    if (shape == 2) {
                        x = x * 1664525u + 1013904223u;
                        v = (float)((x >> 8) & 0xFFFF) / 65536.f - 0.5f;
                    }
                    if (shape == 3) v = (i + j + k) % 2 ? 1.f : -1.f;
                    phi[((size_t)i * M + j) * M + k] = v;
                }
        check(ariel_mesh_sdf(phi.data(), M, path.c_str(), &nv, &nq) == ARIEL_OK, "mesher");
        check(nv > 0 && nq > 0, "mesher output");
    }
    std::vector<float> full(phi.size(), -1.f);
    check(ariel_mesh_sdf(full.data(), M, path.c_str(), &nv, &nq) == ARIEL_OK && nq == 0, "full volume");
    check(ariel_mesh_sdf(nullptr, M, path.c_str(), &nv, &nq) != ARIEL_OK, "null volume");
    check(ariel_mesh_sdf(phi.data(), 1, path.c_str(), &nv, &nq) != ARIEL_OK, "volume of one node");
    check(ariel_mesh_sdf(phi.data(), M, "/nonexistent-dir/x.obj", &nv, &nq) != ARIEL_OK,
          "unwritable path");
}

ariel_scene_desc dam_break(int n);

// The file writers (VDB, Partio) and the scene geometry's emission: valid
// inputs, truncated OBJ files and refusals.
void files_and_scene_paths(const char* dir) {
    const int M = 21;
    std::vector<float> phi((size_t)M * M * M);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j)
            for (int k = 0; k < M; ++k) {
                const float fx = i - 10.f, fy = j - 9.f, fz = k - 11.f;
                const float v = std::sqrt(fx * fx + fy * fy + fz * fz) - 6.1f;
                phi[((size_t)i * M + j) * M + k] = v < -3.f ? -3.f : (v > 3.f ? 3.f : v);
            }
    const std::string vdb = std::string(dir) + "/s.vdb";
    check(ariel_write_vdb(phi.data(), M, 3.0f, vdb.c_str()) == ARIEL_OK, "vdb write");
    check(ariel_write_vdb(phi.data(), M, 3.0f, "/nonexistent-dir/s.vdb") == ARIEL_ERR_IO, "vdb unwritable");
    check(ariel_write_vdb(nullptr, M, 3.0f, vdb.c_str()) == ARIEL_ERR_INVALID, "vdb null");
    std::vector<float> pos(300), vel(300);
    for (int i = 0; i < 300; ++i) {
        pos[i] = 0.1f * i;
        vel[i] = -0.05f * i;
    }
    for (const char* ext : {".geo", ".bgeo", ".geo.gz", ".bgeo.gz"}) {
        const std::string p = std::string(dir) + "/p" + ext;
        check(ariel_write_particles(p.c_str(), 100, pos.data(), vel.data()) == ARIEL_OK, "partio write");
    }
    check(ariel_write_particles((std::string(dir) + "/p.pdb").c_str(), 100, pos.data(), vel.data()) ==
              ARIEL_ERR_UNSUPPORTED, "partio format refused");
    // scenes: generators, a rotation, an OBJ file (and a truncated one), frames
    const std::string obj = std::string(dir) + "/t.obj";
    if (FILE* f = std::fopen(obj.c_str(), "w")) {
        std::fputs("v 4 4 4\nv 20 4 4\nv 4 20 4\nv 4 4 20\nf 1 3 2\nf 1 2 4\nf 1 4 3\nf 2 3 4\n", f);
        std::fclose(f);
    }
    const std::string bad = std::string(dir) + "/bad.obj";
    if (FILE* f = std::fopen(bad.c_str(), "w")) {
        std::fputs("v 1 2\nf 1 9 3\n", f);
        std::fclose(f);
    }
    const char* scene =
        "{\"settings\":[{\"density\":0.5,\"step_size\":0.004,\"dim\":{\"x\":24,\"y\":24,\"z\":24}}],"
        "\"meshfiles\":[{\"id\":\"b\",\"mesh_gen\":\"box\",\"point0\":{\"x\":2,\"y\":2,\"z\":2},"
        "\"point1\":{\"x\":9,\"y\":12,\"z\":20}},{\"id\":\"s\",\"mesh_gen\":\"sphere\","
        "\"center\":{\"x\":15,\"y\":15,\"z\":12},\"radius\":4},{\"id\":\"o\",\"file\":\"t.obj\"}],"
        "\"transforms\":[{\"id\":\"t\"},{\"id\":\"r\",\"rotation\":{\"x\":10,\"y\":20,\"z\":30}}],"
        "\"animatedmeshes\":[{\"id\":\"a\",\"frames\":[\"b\",\"s\"]}],"
        "\"geoms\":[{\"id\":\"g0\",\"type\":\"mesh\",\"geom_frames\":[\"b\"],\"transform_frames\":[\"r\"],"
        "\"frame_interval\":1,\"frame_offset\":0},{\"id\":\"g1\",\"type\":\"animated_mesh\","
        "\"anim_sequence\":\"a\",\"transform_frames\":[\"t\",\"t\"],\"frame_interval\":2,"
        "\"frame_offset\":0,\"post_persist\":true},{\"id\":\"g2\",\"type\":\"mesh\","
        "\"geom_frames\":[\"o\"],\"transform_frames\":[\"t\"],\"frame_interval\":1,\"frame_offset\":0}],"
        "\"sim\":[{\"geom\":\"g0\",\"type\":\"liquid\"},{\"geom\":\"g1\",\"type\":\"liquid\"},"
        "{\"geom\":\"g2\",\"type\":\"liquid\",\"velocity\":{\"x\":1,\"y\":0,\"z\":0}}]}";
    ariel_scene_desc solver = dam_break(24);
    for (int frame = 0; frame < 4; ++frame) {
        int64_t n = -1;
        check(ariel_scene_emit(scene, dir, &solver, frame, nullptr, nullptr, 0, &n) == ARIEL_OK && n > 0,
              "scene emission count");
        std::vector<float> p(3 * (size_t)n), v(3 * (size_t)n);
        check(ariel_scene_emit(scene, dir, &solver, frame, p.data(), v.data(), n, &n) == ARIEL_OK,
              "scene emission");
    }
    std::string broken(scene);
    const size_t at = broken.find("t.obj");
    broken.replace(at, 5, "bad.obj");
    int64_t n = 0;
    check(ariel_scene_emit(broken.c_str(), dir, &solver, 0, nullptr, nullptr, 0, &n) != ARIEL_OK,
          "truncated OBJ refused");
    check(ariel_scene_emit(scene, "/nonexistent-dir", &solver, 0, nullptr, nullptr, 0, &n) == ARIEL_ERR_IO,
          "missing mesh file");
    for (size_t cut = 1; cut < std::strlen(scene); cut += 37) {
        const std::string part(scene, cut);
        (void)ariel_scene_emit(part.c_str(), dir, &solver, 0, nullptr, nullptr, 0, &n);
    }
}

ariel_scene_desc dam_break(int n) {
    ariel_scene_desc d;
    ariel_desc_defaults(&d);
    d.dims[0] = d.dims[1] = d.dims[2] = n;
    d.density = 0.5f;
    d.dt = 1.0f / 240.0f;
    d.wall = 1;
    d.jitter = 1;
    d.n_forces = 1;
    d.forces[0][0] = 0.f;
    d.forces[0][1] = -9.8f;
    d.forces[0][2] = 0.f;
    d.n_liquids = 1;
    d.liquids[0].kind = ARIEL_LIQUID_BOX;
    d.liquids[0].lo[0] = d.liquids[0].lo[1] = d.liquids[0].lo[2] = 0.f;
    d.liquids[0].hi[0] = 0.4f;
    d.liquids[0].hi[1] = 0.6f;
    d.liquids[0].hi[2] = 1.0f;
    return d;
}

void boundary_paths() {
    ariel_scene_desc d;
    check(ariel_desc_defaults(&d) == ARIEL_OK, "defaults");
    check(ariel_desc_defaults(nullptr) != ARIEL_OK, "defaults on null");
    check(std::strlen(ariel_last_error(nullptr)) > 0, "last error of a null handle");
    ariel_sim* h = nullptr;
    check(ariel_create(nullptr, &h) != ARIEL_OK, "create from null");
    ariel_scene_desc bad = dam_break(16);
    bad.dims[1] = 8;
    check(ariel_create(&bad, &h) != ARIEL_OK && h == nullptr, "non-cubic grid");
    bad = dam_break(16);
    bad.n_liquids = 99;
    check(ariel_create(&bad, &h) != ARIEL_OK, "too many liquids");
    bad = dam_break(16);
    bad.precond = 42;
    check(ariel_create(&bad, &h) != ARIEL_OK, "unknown preconditioner");
    bad = dam_break(16);
    bad.dt = -1.f;
    check(ariel_create(&bad, &h) != ARIEL_OK, "negative step");
    // a valid descriptor: creation reaches the device and fails without one
    const ariel_scene_desc ok = dam_break(16);
    const int rc = ariel_create(&ok, &h);
    if (rc == ARIEL_OK) ariel_destroy(h);
    check(ariel_destroy(nullptr) == ARIEL_OK, "destroy null");
    ariel_batch* b = nullptr;
    ariel_scene_desc two[2] = {dam_break(16), dam_break(24)};
    two[1].tol = 1e-3f;
    check(ariel_batch_create(two, 2, &b) != ARIEL_OK, "batch with mixed solver settings");
    check(ariel_batch_create(two, 0, &b) != ARIEL_OK, "empty batch");
    ariel_trainer* t = nullptr;
    const int trc = ariel_trainer_create(&t);
    if (trc == ARIEL_OK) ariel_trainer_destroy(t);
}

void oracle_paths() {
    for (int mode = 0; mode < 2; ++mode) {
        ariel_scene_desc a = dam_break(12);
        a.refine_every = 5;
        oracle_desc od;
        std::memcpy(&od, &a, sizeof(od));
        oracle_sim* s = oracle_create(&od, 2);
        check(s != nullptr, "oracle create");
        check(oracle_init(s) == 0, "oracle init");
        for (int st = 0; st < 3; ++st) check(oracle_step(s, mode) == 0, "oracle step");
        oracle_stats stats;
        oracle_get_stats(s, &stats);
        check(stats.particles > 0 && std::isfinite(stats.residual_norm), "oracle stats");
        const int64_t np = oracle_num_particles(s);
        std::vector<float> pos(3 * np), vel(3 * np), rho(np);
        oracle_get_particles(s, pos.data(), vel.data(), rho.data());
        // coincident pairs through the resampler
        for (int64_t i = 1; i < np && i < 200; i += 2)
            for (int c = 0; c < 3; ++c) pos[3 * i + c] = pos[3 * (i - 1) + c];
        oracle_set_particles(s, pos.data(), vel.data(), rho.data(), np);
        for (int stage = 0; stage < 10; ++stage) check(oracle_stage(s, stage, mode) == 0, "oracle stage");
        const int64_t nc = oracle_field_size(s, 6);
        std::vector<float> x(nc), y(nc), z(nc);
        std::vector<double> y64(nc), p64(nc);
        for (int64_t i = 0; i < nc; ++i) x[i] = std::sin(0.1f * (float)i);
        oracle_apply_A(s, x.data(), y.data());
        oracle_apply_A_f64(s, x.data(), y64.data());
        oracle_apply_precond(s, x.data(), z.data());
        oracle_get_precond(s, z.data());
        double rel = 0.0;
        (void)oracle_solve_fp64(s, 1e-8, 200, p64.data(), &rel);
        std::vector<int32_t> ids(oracle_num_particles(s)), offs(nc + 1);
        oracle_get_sort(s, ids.data(), offs.data());
        std::vector<uint8_t> flags(nc), acc(oracle_num_particles(s));
        oracle_get_flags(s, flags.data());
        oracle_get_accept(s, acc.data());
        for (int f = 0; f < 9; ++f) {
            std::vector<float> fv(oracle_field_size(s, f));
            oracle_get_field(s, f, fv.data());
            oracle_set_field(s, f, fv.data());
        }
        const float p[3] = {0.3f, 0.4f, 0.5f};
        float u[3];
        oracle_interpolate_velocity(s, 0, p, u);
        oracle_destroy(s);
    }
    // the other preconditioners of the extended mode
    for (int pc = 1; pc <= 4; ++pc) {
        ariel_scene_desc a = dam_break(16);
        a.precond = pc;
        oracle_desc od;
        std::memcpy(&od, &a, sizeof(od));
        oracle_sim* s = oracle_create(&od, 1);
        check(oracle_init(s) == 0 && oracle_step(s, 0) == 0, "oracle preconditioner step");
        oracle_destroy(s);
    }
}

}  // namespace

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : "/tmp";
    json_paths();
    mesher_paths(dir);
    files_and_scene_paths(dir);
    boundary_paths();
    oracle_paths();
    if (g_fail) {
        std::fprintf(stderr, "%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
