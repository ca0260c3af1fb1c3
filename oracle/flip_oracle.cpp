// flip_oracle.cpp — CPU restatement of Ariel's FLIP step. TEST INFRASTRUCTURE.
//
// This file is the correctness oracle and the timed CPU baseline ("port"); it is
// loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
// as the checker. It follows the reference files line by line (citations are
// relative to the reference repository root):
//   src/sim/flip.cpp, src/sim/solver.inl, src/sim/particlegridoperations.inl,
//   src/sim/particleresampler.inl, src/grid/particlegrid.cpp,
//   src/grid/macgrid.inl:13-92, src/math/kernels.inl:25-38,
//   src/scene/scene.cpp:257-360,436-477.
//
// Determinism: every loop that the reference runs under tbb::parallel_for writes
// disjoint outputs, so running it on the thread pool below or serially gives the
// same bits. The reference's racy preconditioner sweeps (solver.inl:108-134,
// 238-284) and its Product/Sort (solver.inl:180-193, particlegrid.cpp:140-172) run
// serially here, in the reference's i,j,k order.
//
// Extensions the configurations require (each a stated deviation):
//   E1 static SOLID walls `wall` cells thick, OR-ed into MarkCellTypes and used as
//      the solid test of the resampler's accept step;
//   E2 seeded uniform jitter inside each lattice octant (hash below);
//   E3 relative stop ||r|| <= tol ||b||, max_iters;  E4 MIC(0) with tau/sigma;
//   E5 FP64 dot products and norms, true-residual recompute every k iterations;
//   E6 counter-based hash replacing rand() in the resampler;
//   E10 deterministic particle order (liquids in order, cells i,j,k, octants);
//   E11 optional trilinear P2G weights over the reference window.
// Functions with no external pin are marked "parity unpinned" where defined.
//
// Floating point: compiled with -ffp-contract=off and IEEE division/sqrt; the
// helper gmax/gmin reproduce glm::max/glm::min ((x<y)?y:x, (y<x)?y:x).

#include "flip_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace {

const int AIR = 0, FLUID = 1, SOLID = 2;   // src/grid/macgrid.inl:13

inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline float gmin(float x, float y) { return (y < x) ? y : x; }

struct V3 { float x, y, z; };

// ---------------------------------------------------------------- kernels.inl
// Sqrlength (kernels.inl:25-30)
inline float sqrlength(V3 a, V3 b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}
// Smooth (kernels.inl:32-34)
inline float smooth(float r2, float h) { return gmax(1.0f - r2 / (h * h), 0.0f); }
// Sharpen (kernels.inl:36-38)
inline float sharpen(float r2, float h) { return gmax(h * h / gmax(r2, 1.0e-5f) - 1.0f, 0.0f); }

// ---------------------------------------------------------------- hash (E2/E6)
// 32-bit integer finaliser ("lowbias32", C. Wellons' hash-prospector), chained
// over four words. Pinned by tests/golden/hash4_vectors.json (written by an
// independent Python restatement) and by the published inverse of lowbias32
// (tests/test_oracle_pins.py).
inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
inline uint32_t hash4(uint32_t seed, uint32_t a, uint32_t b, uint32_t c) {
    return mix32(mix32(mix32(mix32(seed) ^ a) ^ b) ^ c);
}
inline float hash_unit(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- thread pool
// Emulates tbb::parallel_for over a blocked range split into one contiguous chunk
// per thread.
class Pool {
public:
    explicit Pool(int n) : n_(n < 1 ? 1 : n) {
        for (int t = 1; t < n_; ++t) ws_.emplace_back([this, t] { loop(t); });
    }
    ~Pool() {
        { std::lock_guard<std::mutex> lk(m_); stop_ = true; }
        cv_.notify_all();
        for (auto& w : ws_) w.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)>& f) {
        if (n_ == 1) { f(0); return; }
        { std::lock_guard<std::mutex> lk(m_); job_ = &f; pending_ = n_ - 1; ++gen_; }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }
private:
    void loop(int t) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* j;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_; j = job_;
            }
            (*j)(t);
            { std::lock_guard<std::mutex> lk(m_); if (--pending_ == 0) done_.notify_one(); }
        }
    }
    int n_;
    std::vector<std::thread> ws_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

template <class F>
void pfor(Pool& p, int64_t b, int64_t e, F f) {
    int n = p.size();
    if (n == 1 || e - b < 2) { for (int64_t i = b; i < e; ++i) f(i); return; }
    std::function<void(int)> job = [&](int t) {
        int64_t len = e - b, lo = b + len * t / n, hi = b + len * (t + 1) / n;
        for (int64_t i = lo; i < hi; ++i) f(i);
    };
    p.run(job);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ============================================================== simulation state
struct oracle_sim {
    oracle_desc d;
    int N = 0;            // grid edge
    float maxd = 0.f;     // glm::max(glm::max(dims.x, dims.z), dims.y) (flip.cpp:47)
    Pool* pool = nullptr;

    // particles in particle-vector (id) order; all FLUID, mass 1 (scene.cpp:436-455)
    std::vector<V3> p, u, t, t2;
    std::vector<float> rho;
    std::vector<uint8_t> accept;

    // ParticleGrid (particlegrid.cpp): per-cell lists as offsets + ids
    std::vector<int32_t> offsets, ids;

    // MacGrid m_mgrid / m_mgrid_previous (macgrid.inl:21-34, 72-84)
    std::vector<float> uf[3], prev[3], P, D, L;
    std::vector<int32_t> A;
    std::vector<float> pc;        // preconditioner values of the last build
    std::vector<int32_t> nup;     // upper fluid-neighbour counts (MIC(0))

    float max_density = 0.f;
    int frame = 0;
    oracle_stats st{};

    int64_t nc() const { return (int64_t)N * N * N; }
    int64_t nf() const { return (int64_t)(N + 1) * N * N; }
    int64_t cidx(int i, int j, int k) const { return ((int64_t)i * N + j) * N + k; }
    // face index of component c at (i,j,k)
    int64_t fidx(int c, int i, int j, int k) const {
        if (c == 0) return ((int64_t)i * N + j) * N + k;
        if (c == 1) return ((int64_t)i * (N + 1) + j) * N + k;
        return ((int64_t)i * N + j) * (N + 1) + k;
    }
    bool in_range(int i, int j, int k) const {
        return i >= 0 && i < N && j >= 0 && j < N && k >= 0 && k < N;
    }
    int a_at(int i, int j, int k) const { return A[cidx(i, j, k)]; }
};

namespace {

// ------------------------------------------------------------- ParticleGrid::Sort
// particlegrid.cpp:140-172. Cell = (int)max(0, min((int)maxd-1, int(maxd)*p)) per
// axis; each cell's list holds particles in particle-vector order. A counting sort
// that is stable in id order produces exactly those lists.
void sort_particles(oracle_sim& s, const std::vector<V3>& pos, std::vector<int32_t>& offsets,
                    std::vector<int32_t>& ids) {
    const int N = s.N;
    const float maxd = s.maxd;
    const int64_t nc = s.nc();
    offsets.assign(nc + 1, 0);
    ids.resize(pos.size());
    std::vector<int64_t> cell(pos.size());
    for (size_t i = 0; i < pos.size(); ++i) {
        float fx = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * pos[i].x));
        float fy = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * pos[i].y));
        float fz = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * pos[i].z));
        cell[i] = ((int64_t)(int)fx * N + (int)fy) * N + (int)fz;
        offsets[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) offsets[c + 1] += offsets[c];
    std::vector<int32_t> cursor(offsets.begin(), offsets.end() - 1);
    for (size_t i = 0; i < pos.size(); ++i) ids[cursor[cell[i]]++] = (int32_t)i;
}

// GetCellNeighbors (particlegrid.cpp:28-50) with a float index: loop variables are
// ints initialised by truncation of (index - n) and compared as floats with
// (index + n); out-of-range cells are skipped. Visits cells sx -> sy -> sz, then
// each cell's list in order.
template <class F>
void for_cell_neighbors(const oracle_sim& s, V3 idx, int n, F f) {
    const float dimx = (float)s.N;
    for (int sx = (int)(idx.x - (float)n); (float)sx <= idx.x + (float)n; sx++) {
        for (int sy = (int)(idx.y - (float)n); (float)sy <= idx.y + (float)n; sy++) {
            for (int sz = (int)(idx.z - (float)n); (float)sz <= idx.z + (float)n; sz++) {
                if (sx < 0 || (float)sx > dimx - 1 || sy < 0 || (float)sy > dimx - 1 || sz < 0 ||
                    (float)sz > dimx - 1)
                    continue;
                int64_t c = s.cidx(sx, sy, sz);
                for (int32_t a = s.offsets[c]; a < s.offsets[c + 1]; ++a) f(s.ids[a]);
            }
        }
    }
}

// GetWallNeighbors (particlegrid.cpp:52-73): window [index-n, index+n-1].
template <class F>
void for_wall_neighbors(const oracle_sim& s, int i, int j, int k, int nx, int ny, int nz, F f) {
    for (int sx = i - nx; sx <= i + nx - 1; sx++)
        for (int sy = j - ny; sy <= j + ny - 1; sy++)
            for (int sz = k - nz; sz <= k + nz - 1; sz++) {
                if (!s.in_range(sx, sy, sz)) continue;
                int64_t c = s.cidx(sx, sy, sz);
                for (int32_t a = s.offsets[c]; a < s.offsets[c + 1]; ++a) f(s.ids[a]);
            }
}

// ----------------------------------------------------------------- ComputeDensity
// flip.cpp:697-733. All particles are FLUID (no solid geoms); mass 1.
void compute_density(oracle_sim& s, const std::vector<V3>& pos, std::vector<float>& rho,
                     float max_density) {
    const float maxd = s.maxd;
    const float hsm = 4.0f * s.d.density / maxd;
    rho.resize(pos.size());
    pfor(*s.pool, 0, (int64_t)pos.size(), [&](int64_t i) {
        V3 position = pos[i];
        position.x = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * position.x));
        position.y = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * position.y));
        position.z = (float)(int)gmax(0.0f, gmin((int)maxd - 1.0f, (float)(int)maxd * position.z));
        float weightsum = 0.0f;
        for_cell_neighbors(s, position, 1, [&](int32_t q) {
            float sqd = sqrlength(pos[q], pos[i]);
            float weight = 1.0f * smooth(sqd, hsm);
            weightsum = weightsum + weight;
        });
        rho[i] = weightsum / max_density;
    });
}

// ------------------------------------------------------------- ApplyExternalForces
// flip.cpp:682-695
void apply_forces(oracle_sim& s) {
    pfor(*s.pool, 0, (int64_t)s.u.size(), [&](int64_t i) {
        for (int f = 0; f < s.d.n_forces; ++f) {
            s.u[i].x = s.u[i].x + s.d.forces[f][0] * s.d.dt;
            s.u[i].y = s.u[i].y + s.d.forces[f][1] * s.d.dt;
            s.u[i].z = s.u[i].z + s.d.forces[f][2] * s.d.dt;
        }
    });
}

// ---------------------------------------------------------- SplatParticlesToMACGrid
// particlegridoperations.inl:146-249 (Sharpen, RE=1.4, FLUID only, mass-weighted,
// 0 when the weight sum is 0). E11: trilinear hat weights over the same window.
inline float hat(float d) { return gmax(1.0f - std::fabs(d), 0.0f); }

void splat_to_grid(oracle_sim& s) {
    const int N = s.N;
    const float RE = s.d.sharpen_re;
    const float maxd = (float)N;
    const bool tri = s.d.splat == 1;
    auto weight = [&](int32_t q, V3 face) {
        V3 pos;
        pos.x = gmax(0.0f, gmin(maxd, maxd * s.p[q].x));
        pos.y = gmax(0.0f, gmin(maxd, maxd * s.p[q].y));
        pos.z = gmax(0.0f, gmin(maxd, maxd * s.p[q].z));
        if (tri) return (hat(pos.x - face.x) * hat(pos.y - face.y)) * hat(pos.z - face.z);
        return 1.0f * sharpen(sqrlength(pos, face), RE);
    };
    pfor(*s.pool, 0, N + 1, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k) {
                if (j < N && k < N) {
                    V3 px = {(float)i, (float)j + 0.5f, (float)k + 0.5f};
                    float sumw = 0.0f, sumx = 0.0f;
                    for_wall_neighbors(s, i, j, k, 1, 2, 2, [&](int32_t q) {
                        float w = weight(q, px);
                        sumx += w * s.u[q].x;
                        sumw += w;
                    });
                    s.uf[0][s.fidx(0, i, j, k)] = (sumw > 0) ? sumx / sumw : 0.0f;
                }
                if (i < N && k < N) {
                    V3 py = {(float)i + 0.5f, (float)j, (float)k + 0.5f};
                    float sumw = 0.0f, sumy = 0.0f;
                    for_wall_neighbors(s, i, j, k, 2, 1, 2, [&](int32_t q) {
                        float w = weight(q, py);
                        sumy += w * s.u[q].y;
                        sumw += w;
                    });
                    s.uf[1][s.fidx(1, i, j, k)] = (sumw > 0) ? sumy / sumw : 0.0f;
                }
                if (i < N && j < N) {
                    V3 pz = {(float)i + 0.5f, (float)j + 0.5f, (float)k};
                    float sumw = 0.0f, sumz = 0.0f;
                    for_wall_neighbors(s, i, j, k, 2, 2, 1, [&](int32_t q) {
                        float w = weight(q, pz);
                        sumz += w * s.u[q].z;
                        sumw += w;
                    });
                    s.uf[2][s.fidx(2, i, j, k)] = (sumw > 0) ? sumz / sumw : 0.0f;
                }
            }
    });
}

// ---------------------------------------------------------------- CellSDF / marks
// CellSDF (particlegrid.cpp:75-90): sum of FLUID densities in list order; all
// particles are FLUID here, so the early "return 1.0f" never fires.
float cell_sdf(const oracle_sim& s, int64_t c) {
    float accm = 0.0f;
    for (int32_t a = s.offsets[c]; a < s.offsets[c + 1]; ++a) accm += s.rho[s.ids[a]];
    float dd = s.d.density;
    float n0 = 1.0f / (dd * dd * dd);
    return 0.2f * n0 - accm;
}

bool is_wall(const oracle_sim& s, int i, int j, int k) {
    int w = s.d.wall, N = s.N;
    if (w <= 0) return false;
    return i < w || j < w || k < w || i >= N - w || j >= N - w || k >= N - w;
}

// MarkCellTypes (particlegrid.cpp:108-138) + E1 wall mask.
void mark_cell_types(oracle_sim& s) {
    const int N = s.N;
    pfor(*s.pool, 0, N, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s.cidx(i, j, k);
                int a = AIR;
                if (is_wall(s, i, j, k)) a = SOLID;
                if (a != SOLID) a = (cell_sdf(s, c) < 0.0) ? FLUID : AIR;
                s.A[c] = a;
            }
    });
}

// BuildSDF (particlegrid.cpp:92-106): L = CellSDF for every cell.
void build_sdf(oracle_sim& s) {
    pfor(*s.pool, 0, s.nc(), [&](int64_t c) { s.L[c] = cell_sdf(s, c); });
}

// ------------------------------------------------------------ grid copy / boundary
// StorePreviousGrid (flip.cpp:324-363)
void store_previous(oracle_sim& s) {
    for (int c = 0; c < 3; ++c) s.prev[c] = s.uf[c];
}
// SubtractPreviousGrid (flip.cpp:365-410)
void subtract_previous(oracle_sim& s) {
    for (int c = 0; c < 3; ++c)
        pfor(*s.pool, 0, s.nf(), [&](int64_t f) { s.prev[c][f] = s.uf[c][f] - s.prev[c][f]; });
}

// CheckWall (particlegridoperations.inl:37-43)
inline float check_wall(const oracle_sim& s, int i, int j, int k) {
    return s.a_at(i, j, k) == SOLID ? 1.0f : -1.0f;
}
// EnforceBoundaryVelocity (particlegridoperations.inl:45-104)
void enforce_boundary(oracle_sim& s) {
    const int N = s.N;
    pfor(*s.pool, 0, N + 1, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k) {
                if (j < N && k < N) {
                    if (i == 0 || i == N) s.uf[0][s.fidx(0, i, j, k)] = 0.0f;
                    if (i < N && i > 0 && check_wall(s, i, j, k) * check_wall(s, i - 1, j, k) < 0)
                        s.uf[0][s.fidx(0, i, j, k)] = 0.0f;
                }
                if (i < N && k < N) {
                    if (j == 0 || j == N) s.uf[1][s.fidx(1, i, j, k)] = 0.0f;
                    if (j < N && j > 0 && check_wall(s, i, j, k) * check_wall(s, i, j - 1, k) < 0)
                        s.uf[1][s.fidx(1, i, j, k)] = 0.0f;
                }
                if (i < N && j < N) {
                    if (k == 0 || k == N) s.uf[2][s.fidx(2, i, j, k)] = 0.0f;
                    if (k < N && k > 0 && check_wall(s, i, j, k) * check_wall(s, i, j, k - 1) < 0)
                        s.uf[2][s.fidx(2, i, j, k)] = 0.0f;
                }
            }
    });
}

// Project divergence (flip.cpp:417-437) followed by FlipGrid (solver.inl:42-56).
void divergence(oracle_sim& s) {
    const int N = s.N;
    const float h = 1.0f / s.maxd;
    pfor(*s.pool, 0, N, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                float dv = (s.uf[0][s.fidx(0, i + 1, j, k)] - s.uf[0][s.fidx(0, i, j, k)] +
                            s.uf[1][s.fidx(1, i, j + 1, k)] - s.uf[1][s.fidx(1, i, j, k)] +
                            s.uf[2][s.fidx(2, i, j, k + 1)] - s.uf[2][s.fidx(2, i, j, k)]) / h;
                s.D[s.cidx(i, j, k)] = -dv;
            }
    });
}

// ------------------------------------------------------------------- solver.inl
// XRef (solver.inl:138-154), subcell = 1.
inline float xref(const oracle_sim& s, const float* X, int fi, int fj, int fk, int pi, int pj,
                  int pk) {
    const int N = s.N;
    int i = std::min(std::max(0, pi), N - 1);
    int j = std::min(std::max(0, pj), N - 1);
    int k = std::min(std::max(0, pk), N - 1);
    int a = s.a_at(i, j, k);
    if (a == FLUID) return X[s.cidx(i, j, k)];
    if (a == SOLID) return X[s.cidx(fi, fj, fk)];
    return s.L[s.cidx(i, j, k)] / gmin(1.0e-6f, s.L[s.cidx(fi, fj, fk)]) * X[s.cidx(fi, fj, fk)];
}

// ComputeAx (solver.inl:196-230)
void compute_ax(const oracle_sim& s, const float* X, float* Y, Pool& pool) {
    const int N = s.N;
    float n = (float)N;
    float h = 1.0f / (n * n);
    pfor(pool, 0, N, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s.cidx(i, j, k);
                if (s.A[c] == FLUID) {
                    float result = (6.0f * X[c] - xref(s, X, i, j, k, i + 1, j, k) -
                                    xref(s, X, i, j, k, i - 1, j, k) -
                                    xref(s, X, i, j, k, i, j + 1, k) -
                                    xref(s, X, i, j, k, i, j - 1, k) -
                                    xref(s, X, i, j, k, i, j, k + 1) -
                                    xref(s, X, i, j, k, i, j, k - 1)) / h;
                    Y[c] = result;
                } else {
                    Y[c] = 0.0f;
                }
            }
    });
}

// E5: the operator of ComputeAx evaluated in double precision from FP32 inputs,
// y = N^2 (6 x_c - sum_q X~_q) with the ghost ratios L_q/min(1e-6, L_c) in double.
// Used for the true-residual recompute and the reported true residual.
void compute_ax_f64(const oracle_sim& s, const float* X, double* Y, Pool& pool) {
    const int N = s.N;
    const double scale = (double)N * (double)N;
    pfor(pool, 0, N, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s.cidx(i, j, k);
                if (s.A[c] != FLUID) { Y[c] = 0.0; continue; }
                const double xc = X[c];
                double acc = 6.0 * xc;
                const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k},
                                     {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (int m = 0; m < 6; ++m) {
                    int a = std::min(std::max(0, q[m][0]), N - 1);
                    int b = std::min(std::max(0, q[m][1]), N - 1);
                    int cc = std::min(std::max(0, q[m][2]), N - 1);
                    int t = s.a_at(a, b, cc);
                    if (t == FLUID) acc -= (double)X[s.cidx(a, b, cc)];
                    else if (t == SOLID) acc -= xc;
                    else acc -= (double)s.L[s.cidx(a, b, cc)] / (double)gmin(1.0e-6f, s.L[c]) * xc;
                }
                Y[c] = scale * acc;
            }
    });
}

// r = b - A p with the FP64 operator; returns ||r||^2 (FP64) and stores r in FP32.
double true_residual(const oracle_sim& s, const float* P, float* R, Pool& pool) {
    std::vector<double> ax(s.nc(), 0.0);
    compute_ax_f64(s, P, ax.data(), pool);
    double acc = 0.0;
    for (int64_t c = 0; c < s.nc(); ++c) {
        if (s.A[c] != FLUID) { R[c] = 0.0f; continue; }
        double r = (double)s.D[c] - ax[c];
        acc += r * r;
        R[c] = (float)r;
    }
    return acc;
}

// ADiag (solver.inl:81-101), subcell = 1; out-of-range neighbours count as SOLID.
float adiag(const oracle_sim& s, int i, int j, int k) {
    float diag = 6.0f;
    if (s.a_at(i, j, k) != FLUID) return diag;
    const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                         {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
    for (int m = 0; m < 6; ++m) {
        int qi = q[m][0], qj = q[m][1], qk = q[m][2];
        if (!s.in_range(qi, qj, qk) || s.a_at(qi, qj, qk) == SOLID) {
            diag -= 1.0f;
        } else if (s.a_at(qi, qj, qk) == AIR) {
            diag -= s.L[s.cidx(qi, qj, qk)] / gmin(1.0e-6f, s.L[s.cidx(i, j, k)]);
        }
    }
    return diag;
}

// ARef / PRef (solver.inl:59-78)
float aref(const oracle_sim& s, int i, int j, int k, int qi, int qj, int qk) {
    if (!s.in_range(i, j, k) || s.a_at(i, j, k) != FLUID) return 0.0f;
    if (!s.in_range(qi, qj, qk) || s.a_at(qi, qj, qk) != FLUID) return 0.0f;
    return -1.0f;
}
float pref(const oracle_sim& s, const std::vector<float>& p, int i, int j, int k) {
    if (!s.in_range(i, j, k) || p[s.cidx(i, j, k)] != (float)FLUID) return 0.0f;
    return p[s.cidx(i, j, k)];
}

// BuildPreconditioner verbatim (solver.inl:104-135), run in serial i,j,k order.
void build_reference_precond(oracle_sim& s) {
    const int N = s.N;
    s.pc.assign(s.nc(), 0.0f);
    const float a = 0.25f;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (s.a_at(i, j, k) != FLUID) continue;
                float left = aref(s, i - 1, j, k, i, j, k) * pref(s, s.pc, i - 1, j, k);
                float bottom = aref(s, i, j - 1, k, i, j, k) * pref(s, s.pc, i, j - 1, k);
                float back = aref(s, i, j, k - 1, i, j, k) * pref(s, s.pc, i, j, k - 1);
                float diag = adiag(s, i, j, k);
                float e = diag - (left * left) - (bottom * bottom) - (back * back);
                if (diag > 0) {
                    if (e < a * diag) e = diag;
                    s.pc[s.cidx(i, j, k)] = 1.0f / std::sqrt(e);
                }
            }
}

// ApplyPreconditioner verbatim (solver.inl:232-286), serial.
void apply_reference_precond(oracle_sim& s, const float* R, float* Z) {
    const int N = s.N;
    std::vector<float> Q(s.nc(), 0.0f);
    const std::vector<float>& P = s.pc;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (s.a_at(i, j, k) != FLUID) continue;
                float left = aref(s, i - 1, j, k, i, j, k) * pref(s, P, i - 1, j, k) *
                             pref(s, Q, i - 1, j, k);
                float bottom = aref(s, i, j - 1, k, i, j, k) * pref(s, P, i, j - 1, k) *
                               pref(s, Q, i, j - 1, k);
                float back = aref(s, i, j, k - 1, i, j, k) * pref(s, P, i, j, k - 1) *
                             pref(s, Q, i, j, k - 1);
                float tt = R[s.cidx(i, j, k)] - left - bottom - back;
                Q[s.cidx(i, j, k)] = tt * P[s.cidx(i, j, k)];
            }
    // Backward sweep: j, k descending; i descending (solver.inl:262-284). Z keeps
    // its previous values on non-fluid cells, which are 0 (ComputeAx wrote them).
    std::vector<float> Zs(Z, Z + s.nc());
    for (int j = N - 1; j >= 0; --j)
        for (int k = N - 1; k >= 0; --k)
            for (int i = N - 1; i >= 0; --i) {
                if (s.a_at(i, j, k) != FLUID) continue;
                float right = aref(s, i, j, k, i + 1, j, k) * pref(s, P, i, j, k) *
                              pref(s, Zs, i + 1, j, k);
                float top = aref(s, i, j, k, i, j + 1, k) * pref(s, P, i, j, k) *
                            pref(s, Zs, i, j + 1, k);
                float front = aref(s, i, j, k, i, j, k + 1) * pref(s, P, i, j, k) *
                              pref(s, Zs, i, j, k + 1);
                float tt = Q[s.cidx(i, j, k)] - right - top - front;
                Zs[s.cidx(i, j, k)] = tt * P[s.cidx(i, j, k)];
            }
    std::copy(Zs.begin(), Zs.end(), Z);
}

// ---------------------------------------------------- MIC(0) under an ordering (E4)
// Pivot (Bridson, "Fluid Simulation for Computer Graphics", 2nd ed., sec. 5.x,
// generalised to any ordering): for fluid cell i with fluid lower neighbours j,
//   e = diag_i - sum_j pc_j^2 - tau * sum_j (nup_j - 1) pc_j^2,
//   e = diag_i if e < sigma*diag_i;  pc_i = 1/sqrt(e);
// nup_j = number of fluid upper neighbours of j. Off-diagonals are -1 (ARef).
// Orderings: natural (lexicographic), tiled red-black (tiles of 8 x 8 x Tz cells,
// Tz = desc.tile, coloured by (ti+tj+tk)&1, red tiles first, natural order inside
// a tile), block (tiles independent: cross-tile couplings dropped). Neighbour order for sums:
// lower pass -x,-y,-z,+x,+y,+z; upper pass +x,+y,+z,-x,-y,-z.
const int DIRS[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
const int UDIRS[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

struct Ordering {
    int kind;  // 1 natural, 0 tiled red-black, 4 block
    int T;     // tile edge along x and y (8)
    int Tz;    // tile edge along z
    int color(int i, int j, int k) const { return ((i / T) + (j / T) + (k / Tz)) & 1; }
    bool same_tile(int i, int j, int k, int a, int b, int c) const {
        return i / T == a / T && j / T == b / T && k / Tz == c / Tz;
    }
    // 1 if (a,b,c) precedes (i,j,k), 0 if it follows, -1 if uncoupled.
    int lower(int i, int j, int k, int a, int b, int c) const {
        bool minus = (a < i) || (b < j) || (c < k);
        if (kind == 1) return minus ? 1 : 0;
        if (same_tile(i, j, k, a, b, c)) return minus ? 1 : 0;
        if (kind == 4) return -1;
        return color(a, b, c) == 0 ? 1 : 0;  // red precedes black
    }
};

Ordering ordering_of(const oracle_sim& s) {
    Ordering o;
    // the multigrid preconditioner (6) has no restatement here: its solves are
    // checked against this oracle's tiled MIC(0)-PCG to the same stop rule, so
    // the two agree to the solve tolerance, not iteration by iteration
    o.kind = s.d.precond == 6 ? 0 : s.d.precond;
    o.T = 8;
    o.Tz = s.d.tile > 0 ? s.d.tile : 8;
    return o;
}

// Cells in factorisation order.
std::vector<int64_t> factor_order(const oracle_sim& s, const Ordering& o) {
    std::vector<int64_t> ord;
    ord.reserve(s.nc());
    const int N = s.N;
    if (o.kind == 0) {
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j)
                    for (int k = 0; k < N; ++k)
                        if (o.color(i, j, k) == pass) ord.push_back(s.cidx(i, j, k));
    } else {
        for (int64_t c = 0; c < s.nc(); ++c) ord.push_back(c);
    }
    return ord;
}

void build_mic0(oracle_sim& s) {
    const int N = s.N;
    const Ordering o = ordering_of(s);
    const float tau = s.d.mic_tau, sigma = s.d.mic_sigma;
    s.pc.assign(s.nc(), 0.0f);
    s.nup.assign(s.nc(), 0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (s.a_at(i, j, k) != FLUID) continue;
                int cnt = 0;
                for (int d = 0; d < 6; ++d) {
                    int a = i + DIRS[d][0], b = j + DIRS[d][1], c = k + DIRS[d][2];
                    if (!s.in_range(a, b, c) || s.a_at(a, b, c) != FLUID) continue;
                    if (o.lower(i, j, k, a, b, c) == 0) cnt++;
                }
                s.nup[s.cidx(i, j, k)] = cnt;
            }
    for (int64_t cell : factor_order(s, o)) {
        int i = (int)(cell / ((int64_t)N * N)), j = (int)((cell / N) % N), k = (int)(cell % N);
        if (s.A[cell] != FLUID) continue;
        float diag = adiag(s, i, j, k);
        float e = diag;
        float ssum = 0.0f;
        for (int d = 0; d < 6; ++d) {
            int a = i + DIRS[d][0], b = j + DIRS[d][1], c = k + DIRS[d][2];
            if (!s.in_range(a, b, c) || s.a_at(a, b, c) != FLUID) continue;
            if (o.lower(i, j, k, a, b, c) != 1) continue;
            float pj = s.pc[s.cidx(a, b, c)];
            e = e - pj * pj;
        }
        for (int d = 0; d < 6; ++d) {
            int a = i + DIRS[d][0], b = j + DIRS[d][1], c = k + DIRS[d][2];
            if (!s.in_range(a, b, c) || s.a_at(a, b, c) != FLUID) continue;
            if (o.lower(i, j, k, a, b, c) != 1) continue;
            float pj = s.pc[s.cidx(a, b, c)];
            ssum = ssum + (float)(s.nup[s.cidx(a, b, c)] - 1) * (pj * pj);
        }
        e = e - tau * ssum;
        if (diag > 0) {
            if (e < sigma * diag) e = diag;
            s.pc[cell] = 1.0f / std::sqrt(e);
        }
    }
}

void apply_mic0(const oracle_sim& s, const float* r, float* z) {
    const int N = s.N;
    const Ordering o = ordering_of(s);
    std::vector<int64_t> ord = factor_order(s, o);
    std::vector<float> q(s.nc(), 0.0f);
    for (int64_t cell : ord) {
        if (s.A[cell] != FLUID) continue;
        int i = (int)(cell / ((int64_t)N * N)), j = (int)((cell / N) % N), k = (int)(cell % N);
        float t = r[cell];
        for (int d = 0; d < 6; ++d) {
            int a = i + DIRS[d][0], b = j + DIRS[d][1], c = k + DIRS[d][2];
            if (!s.in_range(a, b, c) || s.a_at(a, b, c) != FLUID) continue;
            if (o.lower(i, j, k, a, b, c) != 1) continue;
            int64_t jc = s.cidx(a, b, c);
            t = t + s.pc[jc] * q[jc];
        }
        q[cell] = t * s.pc[cell];
    }
    std::vector<float> zz(s.nc(), 0.0f);
    for (auto it = ord.rbegin(); it != ord.rend(); ++it) {
        int64_t cell = *it;
        if (s.A[cell] != FLUID) continue;
        int i = (int)(cell / ((int64_t)N * N)), j = (int)((cell / N) % N), k = (int)(cell % N);
        float t = q[cell];
        for (int d = 0; d < 6; ++d) {
            int a = i + UDIRS[d][0], b = j + UDIRS[d][1], c = k + UDIRS[d][2];
            if (!s.in_range(a, b, c) || s.a_at(a, b, c) != FLUID) continue;
            if (o.lower(i, j, k, a, b, c) != 0) continue;
            t = t + s.pc[cell] * zz[s.cidx(a, b, c)];
        }
        zz[cell] = t * s.pc[cell];
    }
    std::copy(zz.begin(), zz.end(), z);
}

// Jacobi: pc holds 1/diag.
void build_jacobi(oracle_sim& s) {
    const int N = s.N;
    s.pc.assign(s.nc(), 0.0f);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (s.a_at(i, j, k) != FLUID) continue;
                float diag = adiag(s, i, j, k);
                if (diag > 0) s.pc[s.cidx(i, j, k)] = 1.0f / diag;
            }
}

void build_precond(oracle_sim& s) {
    int k = s.d.precond;
    if (k == 3) build_reference_precond(s);
    else if (k == 2) build_jacobi(s);
    else build_mic0(s);
}

void apply_precond(oracle_sim& s, const float* r, float* z) {
    int k = s.d.precond;
    if (k == 3) {
        apply_reference_precond(s, r, z);
    } else if (k == 2) {
        for (int64_t c = 0; c < s.nc(); ++c) z[c] = (s.A[c] == FLUID) ? r[c] * s.pc[c] : 0.0f;
    } else {
        apply_mic0(s, r, z);
    }
}

// Product (solver.inl:180-193): FP32, serial i,j,k order, fluid cells only.
float product_f32(const oracle_sim& s, const float* X, const float* Y) {
    float result = 0.0f;
    for (int64_t c = 0; c < s.nc(); ++c)
        if (s.A[c] == FLUID) result += X[c] * Y[c];
    return result;
}
// E5: FP64 dot over fluid cells.
double product_f64(const oracle_sim& s, const float* X, const float* Y) {
    double result = 0.0;
    for (int64_t c = 0; c < s.nc(); ++c)
        if (s.A[c] == FLUID) result += (double)X[c] * (double)Y[c];
    return result;
}

// Op (solver.inl:157-177): target = X + alpha*Y on fluid cells, 0 elsewhere.
void op(const oracle_sim& s, const float* X, const float* Y, float* T, float alpha, Pool& pool) {
    pfor(pool, 0, s.nc(), [&](int64_t c) {
        T[c] = (s.A[c] == FLUID) ? X[c] + alpha * Y[c] : 0.0f;
    });
}

// SolveConjugateGradient verbatim (solver.inl:289-354): FP32 dots, absolute stop.
void solve_faithful(oracle_sim& s) {
    const int64_t n = s.nc();
    std::vector<float> R(n, 0.0f), Z(n, 0.0f), S(n, 0.0f);
    build_reference_precond(s);
    compute_ax(s, s.P.data(), Z.data(), *s.pool);
    op(s, s.D.data(), Z.data(), R.data(), -1.0f, *s.pool);
    float error0 = product_f32(s, R.data(), R.data());
    apply_reference_precond(s, R.data(), Z.data());
    S = Z;
    const int N = s.N;
    float eps = 1.0e-2f * (float)(N * N * N);
    float a = product_f32(s, Z.data(), R.data());
    int k = 0;
    float error1 = error0;
    for (k = 0; k < N * N * N; k++) {
        compute_ax(s, S.data(), Z.data(), *s.pool);
        float alpha = a / product_f32(s, Z.data(), S.data());
        op(s, s.P.data(), S.data(), s.P.data(), alpha, *s.pool);
        op(s, R.data(), Z.data(), R.data(), -alpha, *s.pool);
        error1 = product_f32(s, R.data(), R.data());
        error0 = gmax(error0, error1);
        if (error1 <= eps) break;
        apply_reference_precond(s, R.data(), Z.data());
        float a2 = product_f32(s, Z.data(), R.data());
        float beta = a2 / a;
        op(s, Z.data(), S.data(), S.data(), beta, *s.pool);
        a = a2;
    }
    s.st.iterations = std::min(k + 1, N * N * N);
    s.st.solve_status = error1 <= eps ? 0 : 1;
    s.st.residual_norm = std::sqrt((double)error1);
}

// The operator on an FP64 vector (same terms and order as compute_ax_f64).
void compute_ax_f64d(const oracle_sim& s, const double* X, double* Y, Pool& pool) {
    const int N = s.N;
    const double scale = (double)N * (double)N;
    pfor(pool, 0, N, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s.cidx(i, j, k);
                if (s.A[c] != FLUID) { Y[c] = 0.0; continue; }
                const double xc = X[c];
                double acc = 6.0 * xc;
                const int q[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k},
                                     {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (int m = 0; m < 6; ++m) {
                    int a = std::min(std::max(0, q[m][0]), N - 1);
                    int b = std::min(std::max(0, q[m][1]), N - 1);
                    int cc = std::min(std::max(0, q[m][2]), N - 1);
                    int t = s.a_at(a, b, cc);
                    if (t == FLUID) acc -= X[s.cidx(a, b, cc)];
                    else if (t == SOLID) acc -= xc;
                    else acc -= (double)s.L[s.cidx(a, b, cc)] / (double)gmin(1.0e-6f, s.L[c]) * xc;
                }
                Y[c] = scale * acc;
            }
    });
}

// r = b - A x for an FP64 solution x; returns ||r||^2 and stores r in FP32.
double true_residual_d(const oracle_sim& s, const double* X, float* R, Pool& pool) {
    std::vector<double> ax(s.nc(), 0.0);
    compute_ax_f64d(s, X, ax.data(), pool);
    double acc = 0.0;
    for (int64_t c = 0; c < s.nc(); ++c) {
        if (s.A[c] != FLUID) { R[c] = 0.0f; continue; }
        double r = (double)s.D[c] - ax[c];
        acc += r * r;
        R[c] = (float)r;
    }
    return acc;
}

// E3-E5 PCG: warm start (non-fluid P cleared, as the reference's first Op does),
// FP64 dots/norms, FP32 vectors. With refinement period k > 0 the solution is
// accumulated in FP64 by group-wise reliable updates (van der Vorst & Ye,
// "Residual replacement strategies for Krylov subspace iterative methods", 2000):
// x64 = warm start, the FP32 PCG iterate P holds the correction since the last
// flush; every k iterations x64 += P, P = 0 and r = b - A x64 evaluated in FP64;
// at exit P = (float)(x64 + P).
void solve_extended(oracle_sim& s) {
    const int64_t n = s.nc();
    std::vector<float> R(n, 0.0f), Z(n, 0.0f), S(n, 0.0f), AX(n, 0.0f);
    std::vector<double> X64;
    const int kref = s.d.refine_every;
    build_precond(s);
    for (int64_t c = 0; c < n; ++c)
        if (s.A[c] != FLUID) s.P[c] = 0.0f;
    double bnorm2 = product_f64(s, s.D.data(), s.D.data());
    double rnorm2;
    if (kref > 0) {
        X64.assign(n, 0.0);
        for (int64_t c = 0; c < n; ++c) X64[c] = (double)s.P[c];
        rnorm2 = true_residual_d(s, X64.data(), R.data(), *s.pool);
        std::fill(s.P.begin(), s.P.end(), 0.0f);
    } else {
        compute_ax(s, s.P.data(), AX.data(), *s.pool);
        op(s, s.D.data(), AX.data(), R.data(), -1.0f, *s.pool);
        rnorm2 = product_f64(s, R.data(), R.data());
    }
    double thr;
    if (s.d.stop_rule == 1) thr = (double)(1.0e-2f * (float)(s.N * s.N * s.N));
    else thr = (double)s.d.tol * (double)s.d.tol * bnorm2;
    s.st.rhs_norm = std::sqrt(bnorm2);
    int it = 0, status = 1;
    if (!(rnorm2 <= thr)) {
        apply_precond(s, R.data(), Z.data());
        S = Z;
        double rho = product_f64(s, Z.data(), R.data());
        for (it = 1; it <= s.d.max_iters; ++it) {
            compute_ax(s, S.data(), Z.data(), *s.pool);
            double sig = product_f64(s, Z.data(), S.data());
            if (!std::isfinite(sig) || !std::isfinite(rho)) { status = 3; break; }
            if (!(sig > 0.0)) { status = 2; break; }
            float af = (float)(rho / sig);
            op(s, s.P.data(), S.data(), s.P.data(), af, *s.pool);
            op(s, R.data(), Z.data(), R.data(), -af, *s.pool);
            rnorm2 = product_f64(s, R.data(), R.data());
            if (kref > 0 && it % kref == 0) {
                for (int64_t c = 0; c < n; ++c) {
                    if (s.A[c] == FLUID) X64[c] += (double)s.P[c];
                    s.P[c] = 0.0f;
                }
                rnorm2 = true_residual_d(s, X64.data(), R.data(), *s.pool);
            }
            if (!std::isfinite(rnorm2)) { status = 3; break; }
            if (rnorm2 <= thr) { status = 0; break; }
            apply_precond(s, R.data(), Z.data());
            double rho2 = product_f64(s, Z.data(), R.data());
            float bf = (float)(rho2 / rho);
            op(s, Z.data(), S.data(), S.data(), bf, *s.pool);
            rho = rho2;
        }
        if (it > s.d.max_iters) it = s.d.max_iters;
    } else {
        status = 0;
    }
    if (kref > 0) {
        for (int64_t c = 0; c < n; ++c)
            if (s.A[c] == FLUID) X64[c] += (double)s.P[c];
        s.st.true_residual_norm64 = std::sqrt(true_residual_d(s, X64.data(), R.data(), *s.pool));
        for (int64_t c = 0; c < n; ++c) s.P[c] = (s.A[c] == FLUID) ? (float)X64[c] : 0.0f;
    }
    s.st.iterations = it;
    s.st.solve_status = status;
    s.st.residual_norm = std::sqrt(rnorm2);
    s.st.true_residual_norm = std::sqrt(true_residual(s, s.P.data(), R.data(), *s.pool));
    if (kref <= 0) s.st.true_residual_norm64 = s.st.true_residual_norm;
}

// SubtractPressureGradient (flip.cpp:568-680), subcell = 1.
void subtract_gradient(oracle_sim& s) {
    const int N = s.N;
    const float h = 1.0f / s.maxd;
    auto Pc = [&](int i, int j, int k) { return s.P[s.cidx(i, j, k)]; };
    auto Lc = [&](int i, int j, int k) { return s.L[s.cidx(i, j, k)]; };
    auto grad = [&](int i, int j, int k, int bi, int bj, int bk) {
        float pf = Pc(i, j, k), pb = Pc(bi, bj, bk);
        if (Lc(i, j, k) * Lc(bi, bj, bk) < 0.0f) {
            if (Lc(i, j, k) < 0.0f) pf = Pc(i, j, k);
            else pf = Lc(i, j, k) / gmin(1.0e-3f, Lc(bi, bj, bk)) * Pc(bi, bj, bk);
            if (Lc(bi, bj, bk) < 0.0f) pb = Pc(bi, bj, bk);
            else pb = Lc(bi, bj, bk) / gmin(1.0e-6f, Lc(i, j, k)) * Pc(i, j, k);
        }
        return (pf - pb) / h;
    };
    pfor(*s.pool, 0, N + 1, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k) {
                if (j < N && k < N && i > 0 && i < N) {
                    float& v = s.uf[0][s.fidx(0, i, j, k)];
                    v -= grad(i, j, k, i - 1, j, k);
                }
                if (i < N && k < N && j > 0 && j < N) {
                    float& v = s.uf[1][s.fidx(1, i, j, k)];
                    v -= grad(i, j, k, i, j - 1, k);
                }
                if (i < N && j < N && k > 0 && k < N) {
                    float& v = s.uf[2][s.fidx(2, i, j, k)];
                    v -= grad(i, j, k, i, j, k - 1);
                }
            }
    });
}

// ExtrapolateVelocity (flip.cpp:452-566)
void extrapolate(oracle_sim& s) {
    const int N = s.N;
    auto fluid = [&](int i, int j, int k) { return s.a_at(i, j, k) == FLUID; };
    auto solid = [&](int i, int j, int k) { return s.a_at(i, j, k) == SOLID; };
    std::vector<uint8_t> mark[3], wallmark[3];
    for (int c = 0; c < 3; ++c) { mark[c].assign(s.nf(), 0); wallmark[c].assign(s.nf(), 0); }
    for (int i = 0; i < N + 1; ++i)
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k) {
                if (j < N && k < N) {
                    int64_t f = s.fidx(0, i, j, k);
                    mark[0][f] = (i > 0 && fluid(i - 1, j, k)) || (i < N && fluid(i, j, k));
                    wallmark[0][f] = (i <= 0 || solid(i - 1, j, k)) && (i >= N || solid(i, j, k));
                }
                if (i < N && k < N) {
                    int64_t f = s.fidx(1, i, j, k);
                    mark[1][f] = (j > 0 && fluid(i, j - 1, k)) || (j < N && fluid(i, j, k));
                    wallmark[1][f] = (j <= 0 || solid(i, j - 1, k)) && (j >= N || solid(i, j, k));
                }
                if (i < N && j < N) {
                    int64_t f = s.fidx(2, i, j, k);
                    mark[2][f] = (k > 0 && fluid(i, j, k - 1)) || (k < N && fluid(i, j, k));
                    wallmark[2][f] = (k <= 0 || solid(i, j, k - 1)) && (k >= N || solid(i, j, k));
                }
            }
    pfor(*s.pool, 0, N + 1, [&](int64_t ii) {
        int i = (int)ii;
        for (int j = 0; j < N + 1; ++j)
            for (int k = 0; k < N + 1; ++k)
                for (int n = 0; n < 3; ++n) {
                    if (n != 0 && i > N - 1) continue;
                    if (n != 1 && j > N - 1) continue;
                    if (n != 2 && k > N - 1) continue;
                    int64_t f = s.fidx(n, i, j, k);
                    if (!(!mark[n][f] && wallmark[n][f])) continue;
                    unsigned wsum = 0;
                    float sum = 0.0f;
                    const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                         {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                    for (int qk = 0; qk < 6; ++qk) {
                        int a = q[qk][0], b = q[qk][1], c = q[qk][2];
                        if (a >= 0 && a < N + (n == 0) && b >= 0 && b < N + (n == 1) && c >= 0 &&
                            c < N + (n == 2)) {
                            int64_t g = s.fidx(n, a, b, c);
                            if (mark[n][g]) {
                                wsum++;
                                sum += s.uf[n][g];
                            }
                        }
                    }
                    if (wsum) s.uf[n][f] = sum / (float)wsum;
                }
    });
}

// Interpolate (particlegridoperations.inl:106-118)
float interpolate(const oracle_sim& s, const std::vector<float>& q, int comp, V3 p, V3 n) {
    float x = gmax(0.0f, gmin(n.x, p.x));
    float y = gmax(0.0f, gmin(n.y, p.y));
    float z = gmax(0.0f, gmin(n.z, p.z));
    int i = (int)gmin(x, n.x - 2);
    int j = (int)gmin(y, n.y - 2);
    int k = (int)gmin(z, n.z - 2);
    auto Q = [&](int a, int b, int c) { return q[s.fidx(comp, a, b, c)]; };
    float term1 = (((float)(i + 1) - x) * Q(i, j, k) + (x - (float)i) * Q(i + 1, j, k)) *
                  ((float)(j + 1) - y);
    float term2 = (((float)(i + 1) - x) * Q(i, j + 1, k) + (x - (float)i) * Q(i + 1, j + 1, k)) *
                  (y - (float)j);
    float term3 = (((float)(i + 1) - x) * Q(i, j, k + 1) + (x - (float)i) * Q(i + 1, j, k + 1)) *
                  ((float)(j + 1) - y);
    float term4 =
        (((float)(i + 1) - x) * Q(i, j + 1, k + 1) + (x - (float)i) * Q(i + 1, j + 1, k + 1)) *
        (y - (float)j);
    return ((float)(k + 1) - z) * (term1 + term2) + (z - (float)k) * (term3 + term4);
}

// InterpolateVelocity (particlegridoperations.inl:120-133)
V3 interpolate_velocity(const oracle_sim& s, const std::vector<float>* g, V3 p) {
    const float x = (float)s.N, y = (float)s.N, z = (float)s.N;
    V3 u;
    u.x = interpolate(s, g[0], 0, V3{x * p.x, y * p.y - 0.5f, z * p.z - 0.5f}, V3{x + 1, y, z});
    u.y = interpolate(s, g[1], 1, V3{x * p.x - 0.5f, y * p.y, z * p.z - 0.5f}, V3{x, y + 1, z});
    u.z = interpolate(s, g[2], 2, V3{x * p.x - 0.5f, y * p.y - 0.5f, z * p.z}, V3{x, y, z + 1});
    return u;
}

// SolvePicFlip (flip.cpp:287-322)
void solve_pic_flip(oracle_sim& s) {
    const float ratio = s.d.picflip;
    pfor(*s.pool, 0, (int64_t)s.p.size(), [&](int64_t i) {
        V3 t = s.u[i];
        V3 du = interpolate_velocity(s, s.prev, s.p[i]);
        t = V3{du.x + t.x, du.y + t.y, du.z + t.z};
        V3 pic = interpolate_velocity(s, s.uf, s.p[i]);
        s.u[i].x = (1.0f - ratio) * pic.x + ratio * t.x;
        s.u[i].y = (1.0f - ratio) * pic.y + ratio * t.y;
        s.u[i].z = (1.0f - ratio) * pic.z + ratio * t.z;
    });
}

// AdvectParticles (flip.cpp:229-285); the solid-particle push is inert (no SOLID
// particles) and the Sort it calls only serves that push.
void advect(oracle_sim& s) {
    const float dt = s.d.dt;
    pfor(*s.pool, 0, (int64_t)s.p.size(), [&](int64_t i) {
        V3 v = interpolate_velocity(s, s.uf, s.p[i]);
        s.p[i].x = s.p[i].x + dt * v.x;
        s.p[i].y = s.p[i].y + dt * v.y;
        s.p[i].z = s.p[i].z + dt * v.z;
    });
    const float r = 1.0f / s.maxd;
    pfor(*s.pool, 0, (int64_t)s.p.size(), [&](int64_t i) {
        s.p[i].x = gmax(r, gmin(1.0f - r, s.p[i].x));
        s.p[i].y = gmax(r, gmin(1.0f - r, s.p[i].y));
        s.p[i].z = gmax(r, gmin(1.0f - r, s.p[i].z));
    });
}

// E1 solid test used by the resampler's accept step (CheckPointInsideSolidGeom,
// scene.cpp:382-400, with the static walls as the only solid).
bool inside_wall(const oracle_sim& s, V3 t) {
    if (s.d.wall <= 0) return false;
    const float lo = (float)s.d.wall, hi = (float)(s.N - s.d.wall);
    float x = t.x * s.maxd, y = t.y * s.maxd, z = t.z * s.maxd;
    return x < lo || y < lo || z < lo || x > hi || y > hi || z > hi;
}

// ResampleParticles + Resample (particleresampler.inl:35-148). The rand() of the
// coincident-neighbour branch is replaced by hash4(seed, 3*frame+axis, id_p, id_q)
// (E6). The reference's draws are unseeded and nondeterministic, so no run of it
// can be matched there; the replacement is pinned by isolated coincident groups
// whose displacement is recomputed from an independent hash.
void resample(oracle_sim& s) {
    const float maxd = s.maxd;
    const float re = s.d.density / maxd;   // flip.cpp:120
    const float dt = s.d.dt;
    const float springforce = s.d.spring;
    sort_particles(s, s.p, s.offsets, s.ids);
    const int64_t np = (int64_t)s.p.size();
    s.t.resize(np);
    s.t2.resize(np);
    s.accept.assign(np, 0);
    const uint32_t seed = s.d.seed;
    const uint32_t fr = (uint32_t)s.frame;
    pfor(*s.pool, 0, np, [&](int64_t n0) {
        V3 pp = s.p[n0];
        V3 spring = {0.0f, 0.0f, 0.0f};
        float x = gmax(0.0f, gmin(maxd, maxd * pp.x));
        float y = gmax(0.0f, gmin(maxd, maxd * pp.y));
        float z = gmax(0.0f, gmin(maxd, maxd * pp.z));
        for_cell_neighbors(s, V3{x, y, z}, 1, [&](int32_t q) {
            if (q == n0) return;
            V3 np_ = s.p[q];
            float dx = pp.x - np_.x, dy = pp.y - np_.y, dz = pp.z - np_.z;
            float dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            float w = springforce * 1.0f * smooth(dist * dist, re);
            if (dist > 0.1f * re) {
                spring.x += w * (pp.x - np_.x) / dist * re;
                spring.y += w * (pp.y - np_.y) / dist * re;
                spring.z += w * (pp.z - np_.z) / dist * re;
            } else {
                uint32_t hx = hash4(seed, 3u * fr + 0u, (uint32_t)n0, (uint32_t)q);
                uint32_t hy = hash4(seed, 3u * fr + 1u, (uint32_t)n0, (uint32_t)q);
                uint32_t hz = hash4(seed, 3u * fr + 2u, (uint32_t)n0, (uint32_t)q);
                spring.x += 0.01f * re / dt * (float)(int)(hx % 101u) / 100.0f;
                spring.y += 0.01f * re / dt * (float)(int)(hy % 101u) / 100.0f;
                spring.z += 0.01f * re / dt * (float)(int)(hz % 101u) / 100.0f;
            }
        });
        s.t[n0].x = pp.x + dt * spring.x;
        s.t[n0].y = pp.y + dt * spring.y;
        s.t[n0].z = pp.z + dt * spring.z;
    });
    pfor(*s.pool, 0, np, [&](int64_t n) {
        V3 tp = s.t[n];
        V3 uu = s.u[n];
        float wsum = 0.0f;
        V3 ru = {0.0f, 0.0f, 0.0f};
        float x = gmax(0.0f, gmin(maxd - 1, maxd * tp.x));
        float y = gmax(0.0f, gmin(maxd - 1, maxd * tp.y));
        float z = gmax(0.0f, gmin(maxd - 1, maxd * tp.z));
        for_cell_neighbors(s, V3{x, y, z}, 1, [&](int32_t q) {
            float dist2 = sqrlength(tp, s.p[q]);
            float w = 1.0f * sharpen(dist2, re);
            ru.x = ru.x + w * s.u[q].x;
            ru.y = ru.y + w * s.u[q].y;
            ru.z = ru.z + w * s.u[q].z;
            wsum += w;
        });
        if (wsum != 0.0f) {
            ru.x = ru.x / wsum; ru.y = ru.y / wsum; ru.z = ru.z / wsum;
        } else {
            ru = uu;
        }
        s.t2[n] = ru;
    });
    for (int64_t n = 0; n < np; ++n) {
        if (!inside_wall(s, s.t[n])) {
            s.accept[n] = 1;
            s.p[n] = s.t[n];
            s.u[n] = s.t2[n];
        }
    }
}

// Seeding (Scene::GenerateParticles, scene.cpp:257-360; AddLiquidParticle, 436-455)
// with E2 jitter and E10 order. Lattice spacing w = density * (1/maxd); m = 1/density
// lattice points per cell and axis; lattice point a sits at a*w + u*w with u = 0.5
// (reference lattice) or u ~ U[0,1) from hash4(seed, 3*l+axis, a*Nm+b, c).
void seed_particles(oracle_sim& s) {
    const int N = s.N;
    const float maxd = s.maxd;
    const float thickness = 1.0f / maxd;
    const float w = s.d.density * thickness;
    const int m = (int)std::lround(1.0 / s.d.density);
    const int Nm = N * m;
    const int wall = std::max(0, s.d.wall);
    s.p.clear(); s.u.clear(); s.rho.clear();
    for (int l = 0; l < s.d.n_liquids; ++l) {
        const oracle_liquid& lq = s.d.liquids[l];
        for (int ci = wall; ci < N - wall; ++ci)
            for (int cj = wall; cj < N - wall; ++cj)
                for (int ck = wall; ck < N - wall; ++ck)
                    for (int oa = 0; oa < m; ++oa)
                        for (int ob = 0; ob < m; ++ob)
                            for (int oc = 0; oc < m; ++oc) {
                                uint32_t a = (uint32_t)(ci * m + oa), b = (uint32_t)(cj * m + ob),
                                         c = (uint32_t)(ck * m + oc);
                                float jx = 0.5f, jy = 0.5f, jz = 0.5f;
                                if (s.d.jitter) {
                                    uint32_t key = a * (uint32_t)Nm + b;
                                    jx = hash_unit(hash4(s.d.seed, 3u * l + 0u, key, c));
                                    jy = hash_unit(hash4(s.d.seed, 3u * l + 1u, key, c));
                                    jz = hash_unit(hash4(s.d.seed, 3u * l + 2u, key, c));
                                }
                                V3 pos = {(float)a * w + jx * w, (float)b * w + jy * w,
                                          (float)c * w + jz * w};
                                bool inside;
                                if (lq.kind == 0) {
                                    inside = pos.x >= lq.lo[0] && pos.x < lq.hi[0] &&
                                             pos.y >= lq.lo[1] && pos.y < lq.hi[1] &&
                                             pos.z >= lq.lo[2] && pos.z < lq.hi[2];
                                } else {
                                    float dx = pos.x - lq.center[0], dy = pos.y - lq.center[1],
                                          dz = pos.z - lq.center[2];
                                    inside = dx * dx + dy * dy + dz * dz < lq.radius * lq.radius;
                                }
                                if (!inside) continue;
                                s.p.push_back(pos);
                                s.u.push_back(V3{lq.velocity[0], lq.velocity[1], lq.velocity[2]});
                                s.rho.push_back(10.0f);
                            }
    }
}

void max_divergence(oracle_sim& s) {
    const int N = s.N;
    const float h = 1.0f / s.maxd;
    double mx = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                if (s.a_at(i, j, k) != FLUID) continue;
                float dv = (s.uf[0][s.fidx(0, i + 1, j, k)] - s.uf[0][s.fidx(0, i, j, k)] +
                            s.uf[1][s.fidx(1, i, j + 1, k)] - s.uf[1][s.fidx(1, i, j, k)] +
                            s.uf[2][s.fidx(2, i, j, k + 1)] - s.uf[2][s.fidx(2, i, j, k)]) / h;
                mx = std::max(mx, (double)std::fabs(dv));
            }
    s.st.max_divergence = mx;
}

int count_fluid(const oracle_sim& s) {
    int n = 0;
    for (int64_t c = 0; c < s.nc(); ++c) n += s.A[c] == FLUID;
    return n;
}

void run_stage(oracle_sim& s, int stage, int mode) {
    double t0 = now_ms();
    switch (stage) {
        case 0: sort_particles(s, s.p, s.offsets, s.ids); break;
        case 1: compute_density(s, s.p, s.rho, s.max_density); apply_forces(s); break;
        case 2: mark_cell_types(s); build_sdf(s); break;
        case 3: splat_to_grid(s); store_previous(s); enforce_boundary(s); break;
        case 4: divergence(s); break;
        case 5: {
            double t1 = now_ms();
            if (mode == ORACLE_MODE_FAITHFUL) solve_faithful(s);
            else solve_extended(s);
            s.st.solve_ms = now_ms() - t1;
            s.st.fluid_cells = count_fluid(s);
            break;
        }
        case 6: subtract_gradient(s); enforce_boundary(s); max_divergence(s); break;
        case 7: extrapolate(s); subtract_previous(s); break;
        case 8: solve_pic_flip(s); advect(s); break;
        case 9:
            if (s.d.resample) resample(s);
            break;
        default: break;
    }
    s.st.stage_ms[stage] = now_ms() - t0;
}

}  // namespace

// ================================================================== C interface
extern "C" {

oracle_sim* oracle_create(const oracle_desc* d, int threads) {
    if (!d || d->dims[0] <= 2 || d->dims[0] != d->dims[1] || d->dims[0] != d->dims[2])
        return nullptr;
    oracle_sim* s = new oracle_sim();
    s->d = *d;
    s->N = d->dims[0];
    s->maxd = std::max(std::max((float)d->dims[0], (float)d->dims[2]), (float)d->dims[1]);
    s->pool = new Pool(threads);
    for (int c = 0; c < 3; ++c) { s->uf[c].assign(s->nf(), 0.0f); s->prev[c].assign(s->nf(), 0.0f); }
    s->P.assign(s->nc(), 0.0f);
    s->D.assign(s->nc(), 0.0f);
    s->L.assign(s->nc(), 1.6f);     // macgrid.inl:82
    s->A.assign(s->nc(), AIR);
    s->offsets.assign(s->nc() + 1, 0);
    return s;
}

void oracle_destroy(oracle_sim* s) {
    if (!s) return;
    delete s->pool;
    delete s;
}

// FlipSim::Init (flip.cpp:42-77)
int oracle_init(oracle_sim* s) {
    const float maxd = s->maxd;
    const float h = s->d.density / maxd;
    std::vector<V3> tmp;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
            for (int k = 0; k < 10; ++k)
                tmp.push_back(V3{((float)i + 0.5f) * h, ((float)j + 0.5f) * h, ((float)k + 0.5f) * h});
    sort_particles(*s, tmp, s->offsets, s->ids);
    std::vector<float> r;
    compute_density(*s, tmp, r, 1.0f);
    float md = 0.0f;
    for (float v : r) md = gmax(md, v);
    s->max_density = md;
    seed_particles(*s);
    sort_particles(*s, s->p, s->offsets, s->ids);
    mark_cell_types(*s);
    s->frame = 0;
    return 0;
}

// FlipSim::Step (flip.cpp:92-128) in the reference's order.
int oracle_step(oracle_sim* s, int mode) {
    double t0 = now_ms();
    s->frame++;
    auto timed = [&](int stage, const std::function<void()>& f) {
        double a = now_ms();
        f();
        s->st.stage_ms[stage] += now_ms() - a;
    };
    for (double& v : s->st.stage_ms) v = 0.0;
    timed(0, [&] { sort_particles(*s, s->p, s->offsets, s->ids); });
    timed(1, [&] { compute_density(*s, s->p, s->rho, s->max_density); apply_forces(*s); });
    timed(3, [&] { splat_to_grid(*s); });
    timed(2, [&] { mark_cell_types(*s); });
    timed(3, [&] { store_previous(*s); enforce_boundary(*s); });
    timed(4, [&] { divergence(*s); });
    timed(2, [&] { build_sdf(*s); });
    timed(5, [&] {
        double a = now_ms();
        if (mode == ORACLE_MODE_FAITHFUL) solve_faithful(*s); else solve_extended(*s);
        s->st.solve_ms = now_ms() - a;
        s->st.fluid_cells = count_fluid(*s);
    });
    timed(6, [&] { subtract_gradient(*s); enforce_boundary(*s); max_divergence(*s); });
    timed(7, [&] { extrapolate(*s); subtract_previous(*s); });
    timed(8, [&] { solve_pic_flip(*s); advect(*s); });
    timed(9, [&] { if (s->d.resample) resample(*s); });
    s->st.step_ms = now_ms() - t0;
    return 0;
}

int oracle_stage(oracle_sim* s, int stage, int mode) {
    if (stage < 0 || stage > 9) return -1;
    run_stage(*s, stage, mode);
    return 0;
}

int64_t oracle_num_particles(const oracle_sim* s) { return (int64_t)s->p.size(); }

void oracle_get_particles(const oracle_sim* s, float* pos, float* vel, float* density) {
    for (size_t i = 0; i < s->p.size(); ++i) {
        if (pos) { pos[3 * i] = s->p[i].x; pos[3 * i + 1] = s->p[i].y; pos[3 * i + 2] = s->p[i].z; }
        if (vel) { vel[3 * i] = s->u[i].x; vel[3 * i + 1] = s->u[i].y; vel[3 * i + 2] = s->u[i].z; }
        if (density) density[i] = s->rho[i];
    }
}

void oracle_set_particles(oracle_sim* s, const float* pos, const float* vel, const float* density,
                          int64_t n) {
    s->p.resize(n); s->u.resize(n); s->rho.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        s->p[i] = V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        s->u[i] = vel ? V3{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]} : V3{0, 0, 0};
        s->rho[i] = density ? density[i] : 10.0f;
    }
}

static std::vector<float>* field_ptr(oracle_sim* s, int field) {
    switch (field) {
        case 0: case 1: case 2: return &s->uf[field];
        case 3: case 4: case 5: return &s->prev[field - 3];
        case 6: return &s->P;
        case 7: return &s->D;
        case 8: return &s->L;
        default: return nullptr;
    }
}

int64_t oracle_field_size(const oracle_sim* s, int field) {
    return field <= 5 ? s->nf() : s->nc();
}

void oracle_get_field(const oracle_sim* s, int field, float* dst) {
    auto* v = field_ptr(const_cast<oracle_sim*>(s), field);
    if (v) std::copy(v->begin(), v->end(), dst);
}

void oracle_set_field(oracle_sim* s, int field, const float* src) {
    auto* v = field_ptr(s, field);
    if (v) std::copy(src, src + v->size(), v->begin());
}

void oracle_get_flags(const oracle_sim* s, uint8_t* dst) {
    for (int64_t c = 0; c < s->nc(); ++c) dst[c] = (uint8_t)s->A[c];
}
void oracle_set_flags(oracle_sim* s, const uint8_t* src) {
    for (int64_t c = 0; c < s->nc(); ++c) s->A[c] = src[c];
}
void oracle_get_sort(const oracle_sim* s, int32_t* ids, int32_t* offsets) {
    if (ids) std::copy(s->ids.begin(), s->ids.end(), ids);
    if (offsets) std::copy(s->offsets.begin(), s->offsets.end(), offsets);
}
void oracle_get_accept(const oracle_sim* s, uint8_t* dst) {
    std::copy(s->accept.begin(), s->accept.end(), dst);
}
float oracle_get_max_density(const oracle_sim* s) { return s->max_density; }
void oracle_set_max_density(oracle_sim* s, float v) { s->max_density = v; }
void oracle_get_stats(const oracle_sim* s, oracle_stats* st) {
    *st = s->st;
    st->frame = s->frame;
    st->particles = (int64_t)s->p.size();
    st->max_density = s->max_density;
}
int oracle_get_frame(const oracle_sim* s) { return s->frame; }
void oracle_set_frame(oracle_sim* s, int frame) { s->frame = frame; }

void oracle_apply_A(const oracle_sim* s, const float* x, float* y) {
    compute_ax(*s, x, y, *s->pool);
}

void oracle_apply_A_f64(const oracle_sim* s, const float* x, double* y) {
    compute_ax_f64(*s, x, y, *s->pool);
}

void oracle_apply_precond(oracle_sim* s, const float* r, float* z) {
    build_precond(*s);
    std::vector<float> zz(s->nc(), 0.0f);
    apply_precond(*s, r, zz.data());
    std::copy(zz.begin(), zz.end(), z);
}

void oracle_get_precond(const oracle_sim* s, float* dst) {
    std::copy(s->pc.begin(), s->pc.end(), dst);
}

// FP64 PCG with natural-order MIC(0) in double on the same operator. Parity
// pinned against scipy's sparse direct solve in tests on small grids.
int oracle_solve_fp64(const oracle_sim* s, double tol, int max_iters, double* p_out,
                      double* rel_residual) {
    const int N = s->N;
    const int64_t n = s->nc();
    const double scale = (double)N * (double)N;
    std::vector<double> diag(n, 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s->cidx(i, j, k);
                if (s->A[c] != FLUID) continue;
                double dg = 6.0;
                const int q[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                     {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                for (int m = 0; m < 6; ++m) {
                    int a = q[m][0], b = q[m][1], cc = q[m][2];
                    if (!s->in_range(a, b, cc) || s->a_at(a, b, cc) == SOLID) dg -= 1.0;
                    else if (s->a_at(a, b, cc) == AIR)
                        dg -= (double)s->L[s->cidx(a, b, cc)] / (double)gmin(1.0e-6f, s->L[c]);
                }
                diag[c] = dg;
            }
    auto fl = [&](int a, int b, int c) { return s->in_range(a, b, c) && s->a_at(a, b, c) == FLUID; };
    auto matvec = [&](const std::vector<double>& x, std::vector<double>& y) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    int64_t c = s->cidx(i, j, k);
                    if (s->A[c] != FLUID) { y[c] = 0.0; continue; }
                    double v = diag[c] * x[c];
                    for (int d = 0; d < 6; ++d) {
                        int a = i + DIRS[d][0], b = j + DIRS[d][1], cc = k + DIRS[d][2];
                        if (fl(a, b, cc)) v -= x[s->cidx(a, b, cc)];
                    }
                    y[c] = scale * v;
                }
    };
    // natural-order MIC(0) in double
    std::vector<double> pc(n, 0.0);
    const double tau = s->d.mic_tau, sigma = s->d.mic_sigma;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                int64_t c = s->cidx(i, j, k);
                if (s->A[c] != FLUID) continue;
                double e = diag[c];
                for (int d = 0; d < 3; ++d) {
                    int a = i + DIRS[d][0], b = j + DIRS[d][1], cc = k + DIRS[d][2];
                    if (!fl(a, b, cc)) continue;
                    int64_t jc = s->cidx(a, b, cc);
                    int nupj = 0;
                    for (int u = 3; u < 6; ++u)
                        if (fl(a + DIRS[u][0], b + DIRS[u][1], cc + DIRS[u][2])) nupj++;
                    e -= pc[jc] * pc[jc] * (1.0 + tau * (nupj - 1));
                }
                if (diag[c] > 0) {
                    if (e < sigma * diag[c]) e = diag[c];
                    pc[c] = 1.0 / std::sqrt(e);
                }
            }
    auto precond = [&](const std::vector<double>& r, std::vector<double>& z) {
        std::vector<double> q(n, 0.0);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    int64_t c = s->cidx(i, j, k);
                    if (s->A[c] != FLUID) continue;
                    double t = r[c];
                    for (int d = 0; d < 3; ++d) {
                        int a = i + DIRS[d][0], b = j + DIRS[d][1], cc = k + DIRS[d][2];
                        if (fl(a, b, cc)) { int64_t jc = s->cidx(a, b, cc); t += pc[jc] * q[jc]; }
                    }
                    q[c] = t * pc[c];
                }
        std::fill(z.begin(), z.end(), 0.0);
        for (int i = N - 1; i >= 0; --i)
            for (int j = N - 1; j >= 0; --j)
                for (int k = N - 1; k >= 0; --k) {
                    int64_t c = s->cidx(i, j, k);
                    if (s->A[c] != FLUID) continue;
                    double t = q[c];
                    for (int d = 3; d < 6; ++d) {
                        int a = i + DIRS[d][0], b = j + DIRS[d][1], cc = k + DIRS[d][2];
                        if (fl(a, b, cc)) t += pc[c] * z[s->cidx(a, b, cc)];
                    }
                    z[c] = t * pc[c];
                }
    };
    std::vector<double> x(n, 0.0), r(n, 0.0), z(n, 0.0), sdir(n, 0.0), az(n, 0.0);
    double bn = 0.0;
    for (int64_t c = 0; c < n; ++c)
        if (s->A[c] == FLUID) { r[c] = s->D[c]; bn += r[c] * r[c]; }
    bn = std::sqrt(bn);
    int it = 0;
    double rn = bn;
    if (bn > 0) {
        precond(r, z);
        sdir = z;
        double rho = 0.0;
        for (int64_t c = 0; c < n; ++c) rho += z[c] * r[c];
        for (it = 1; it <= max_iters; ++it) {
            matvec(sdir, az);
            double sg = 0.0;
            for (int64_t c = 0; c < n; ++c) sg += az[c] * sdir[c];
            double alpha = rho / sg;
            rn = 0.0;
            for (int64_t c = 0; c < n; ++c) {
                x[c] += alpha * sdir[c];
                r[c] -= alpha * az[c];
                rn += r[c] * r[c];
            }
            rn = std::sqrt(rn);
            if (rn <= tol * bn) break;
            precond(r, z);
            double rho2 = 0.0;
            for (int64_t c = 0; c < n; ++c) rho2 += z[c] * r[c];
            double beta = rho2 / rho;
            for (int64_t c = 0; c < n; ++c) sdir[c] = z[c] + beta * sdir[c];
            rho = rho2;
        }
    }
    if (p_out) std::copy(x.begin(), x.end(), p_out);
    if (rel_residual) *rel_residual = bn > 0 ? rn / bn : 0.0;
    return (bn == 0 || rn <= tol * bn) ? it : -1;
}

float oracle_smooth(float r2, float h) { return smooth(r2, h); }
float oracle_sharpen(float r2, float h) { return sharpen(r2, h); }
float oracle_sqrlength(const float* p0, const float* p1) {
    return sqrlength(V3{p0[0], p0[1], p0[2]}, V3{p1[0], p1[1], p1[2]});
}
uint32_t oracle_hash4(uint32_t seed, uint32_t a, uint32_t b, uint32_t c) {
    return hash4(seed, a, b, c);
}
void oracle_interpolate_velocity(const oracle_sim* s, int which_grid, const float* p, float* u) {
    V3 v = interpolate_velocity(*s, which_grid == 0 ? s->uf : s->prev, V3{p[0], p[1], p[2]});
    u[0] = v.x; u[1] = v.y; u[2] = v.z;
}

}  // extern "C"
