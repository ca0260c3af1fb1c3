// ariel_flipsim.hpp — header-only C++ facade over the C ABI with the shape of
// Ariel's fluidCore::FlipSim (src/sim/flip.hpp:20-34), so that the reference's
// callers (main.cpp:52-54 constructs it; viewer.cpp:66-86 calls Init() once and
// Step() in a loop, reading GetParticles() after each step) keep their calls.
//
// Differences that a caller sees:
//  * the scene is an ariel_scene_desc (liquid primitives, forces, seed) or the
//    reference's JSON scene with its mesh files (the second constructor:
//    meshes, transforms, animated meshes, emitters; no solid geometry) instead
//    of sceneCore::Scene;
//  * GetParticles() returns a snapshot by value (device memory stays on the GPU)
//    in particle-vector order, not a pointer to live particles;
//  * failures throw ariel::Error carrying the ABI status and message.
#ifndef ARIEL_FLIPSIM_HPP
#define ARIEL_FLIPSIM_HPP

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "ariel.h"

namespace ariel {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// One particle as the reference's Particle exposes it (macgrid.inl:36-50):
// position, velocity, density and type (always ARIEL_FLUID here).
struct Particle {
    float p[3];
    float u[3];
    float density;
    int type;
};

struct Dimensions {
    int x, y, z;
};

class FlipSim {
public:
    // FlipSim(maxres, density, stepsize, scene, verbose) (flip.cpp:15-30):
    // dims, density and dt live in the descriptor.
    explicit FlipSim(const ariel_scene_desc& desc) : m_frame(0), desc_(desc) {
        check(ariel_create(&desc, &h_), nullptr);
    }
    // FlipSim(..., Scene*) with the Scene that SceneLoader reads from a JSON scene
    // (sceneloader.cpp:11-474): the scene's text, the folder its mesh files are
    // relative to, and the solver fields (defaults when null).
    FlipSim(const std::string& scene_json, const std::string& base_dir,
            const ariel_scene_desc* solver = nullptr)
        : m_frame(0) {
        check(ariel_create_from_json(scene_json.c_str(), base_dir.c_str(), solver, &h_), nullptr);
        check(ariel_get_desc(h_, &desc_), h_);
    }
    ~FlipSim() { ariel_destroy(h_); }
    FlipSim(const FlipSim&) = delete;
    FlipSim& operator=(const FlipSim&) = delete;

    void Init() {
        check(ariel_init(h_), h_);
        sync_frame();
    }

    // Step(saveVDB, saveOBJ, savePARTIO) (flip.cpp:92-128) with the export of
    // Scene::ExportParticles (scene.cpp:34-101) after the step: the particles'
    // velocity-trail level set (LevelSet(particles, maxd), levelset.cpp:101-111)
    // as an OpenVDB file, its mesh and the particles go to
    // <export_path>.<frame>.vdb, .obj and <partio_extension> (".geo" unless set),
    // the frame padded to 4 digits and placed before the extension as the
    // reference names its files (scene.cpp:48-58, 85-90).
    void Step(bool saveVDB = false, bool saveOBJ = false, bool savePARTIO = false,
              const char* export_path = nullptr) {
        check(ariel_step(h_, 1), h_);
        sync_frame();
        if (!export_path || !(saveVDB || saveOBJ || savePARTIO)) return;
        char frame[16];
        std::snprintf(frame, sizeof(frame), ".%04d", m_frame);
        const std::string base = std::string(export_path) + frame;
        check(ariel_set_levelset_mode(h_, ARIEL_LEVELSET_TRAILS), h_);
        if (saveVDB) check(ariel_export_levelset(h_, (base + ".vdb").c_str()), h_);
        if (saveOBJ) check(ariel_export_mesh(h_, (base + ".obj").c_str()), h_);
        if (savePARTIO) check(ariel_export_particles(h_, (base + partio_extension).c_str()), h_);
    }

    // The Partio format of Step's particle files, by extension (".geo", ".bgeo",
    // either with ".gz"), as the reference's scene path selects it.
    std::string partio_extension = ".geo";

    std::vector<Particle> GetParticles() {
        int64_t n = 0;
        check(ariel_particle_count(h_, &n), h_);
        std::vector<float> pos(3 * n), vel(3 * n), rho(n);
        std::vector<uint8_t> type(n);
        check(ariel_get_particles(h_, pos.data(), vel.data(), rho.data(), type.data(), n, &n), h_);
        std::vector<Particle> out((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) {
                out[i].p[c] = pos[3 * i + c];
                out[i].u[c] = vel[3 * i + c];
            }
            out[i].density = rho[i];
            out[i].type = type[i];
        }
        return out;
    }

    // FlipSim::GetScene (flip.hpp:31): the scene the simulation was made from,
    // here its descriptor.
    const ariel_scene_desc& GetScene() const { return desc_; }

    Dimensions GetDimensions() {
        int32_t d[3];
        check(ariel_get_dimensions(h_, d), h_);
        return Dimensions{d[0], d[1], d[2]};
    }

    // MAC-grid accessor (no counterpart in the reference).
    std::vector<float> GetMacGrid(int which) {
        const Dimensions d = GetDimensions();
        const int64_t n = which <= ARIEL_FIELD_DUZ ? (int64_t)(d.x + 1) * d.y * d.z
                                                   : (int64_t)d.x * d.y * d.z;
        std::vector<float> out((size_t)n);
        check(ariel_get_macgrid(h_, which, out.data(), n), h_);
        return out;
    }

    ariel_step_stats GetStats() {
        ariel_step_stats s;
        check(ariel_get_stats(h_, &s), h_);
        return s;
    }

    ariel_sim* handle() { return h_; }

    int m_frame;   // FlipSim::m_frame (flip.hpp:33)

private:
    void sync_frame() {
        int32_t f = 0;
        check(ariel_get_frame(h_, &f), h_);
        m_frame = f;
    }
    static void check(int rc, const ariel_sim* h) {
        if (rc != ARIEL_OK) throw Error(rc, ariel_last_error(h));
    }
    ariel_scene_desc desc_;
    ariel_sim* h_ = nullptr;
};

}  // namespace ariel

#endif  // ARIEL_FLIPSIM_HPP
