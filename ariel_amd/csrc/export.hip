// export.hip — frame export of the fluid (Scene::ExportParticles,
// src/scene/scene.cpp:34-101): the particle level set rasterised on the GPU,
// its zero surface as a polygon mesh (OBJ); the particle and VDB files are
// written by files.hip.
//
// Level set (LevelSet(particles, maxd), src/grid/levelset.cpp:101-111): nodes
// of a grid of voxel size 1 in grid units (a particle sits at p * maxd), node
// (i, j, k) at (i, j, k), 0 <= i, j, k <= N; phi = min over particles of
// |x - p| - r with r = 0.5, limited to the narrow band +3 voxels (OpenVDB's
// default half width, whose background value is 3). The reference rasterises
// velocity trails with OpenVDB's ParticlesToLevelSet (not in this image); the
// union of spheres here is its rasterizeSpheres variant (commented out at
// levelset.cpp:108).
//
// File formats (OpenVDB, Partio) are in files.hip.
//
// Mesh (LevelSet::WriteObjToFile, levelset.cpp:113-173, which calls OpenVDB's
// VolumeToMesh): a dual surface like VolumeToMesh's (one vertex per voxel
// that the zero surface crosses, one quad per crossed voxel edge), built with
// surface nets (Gibson, "Constrained elastic surface nets", MERL TR 99-24,
// 1999, without the relaxation): the vertex is the mean of the zero crossings
// on the voxel's edges. Faces wind so that normals point out of the liquid.
#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Cell of each particle (the simulation's binning rule) and its rank among
// the cell's particles (arrival order; the level set is a minimum, so the
// order inside a cell does not matter).
struct KSdfBin {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float4* pos;
        int64_t np;
        int N;
        float maxd;
        int32_t* cell_of;
        int32_t* rank;
        int32_t* counts;
    };
    static __device__ void run(const Args& a) {
        const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (s >= a.np) return;
        const float4 p = a.pos[s];
        const int32_t c = (int32_t)cidx(a.N, cell_coord(p.x, a.maxd), cell_coord(p.y, a.maxd),
                                        cell_coord(p.z, a.maxd));
        a.cell_of[s] = c;
        a.rank[s] = atomicAdd(a.counts + c, 1);
    }
};

// Particle positions grouped by cell, in grid units.
struct KSdfScatter {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float4* pos;
        int64_t np;
        float maxd;
        const int32_t* cell_of;
        const int32_t* rank;
        const int32_t* offsets;
        float4* out;
    };
    static __device__ void run(const Args& a) {
        const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (s >= a.np) return;
        const float4 p = a.pos[s];
        a.out[a.offsets[a.cell_of[s]] + a.rank[s]] =
            make_float4(p.x * a.maxd, p.y * a.maxd, p.z * a.maxd, 0.0f);
    }
};

// phi at every node. Cells are searched in shells around the node: shell s
// covers the cells [i-1-s, i+s] on each axis (only those not covered by shell
// s-1 are read), so any particle not yet seen is at least s+1 away; the search
// stops once the nearest distance found is below that (less a margin far above
// the rounding of the distances), or after shell 3, beyond which every
// particle is at least 4 > band + r away. Lanes take consecutive k, so they
// read the same columns' offsets.
struct KParticleSdf {
    static constexpr int kThreads = kBlock;
    static constexpr int kShells = 4;
    struct Args {
        const float4* sp;        // positions grouped by cell (grid units)
        const int32_t* offsets;  // N^3 + 1
        int N;
        float band, rad;
        float* phi;              // (N+1)^3
    };
    static __device__ float column_min(const Args& a, float best, int cx, int cy, int z0, int z1,
                                       float x, float y, float z) {
        if (z0 < 0) z0 = 0;
        if (z1 > a.N - 1) z1 = a.N - 1;
        if (z0 > z1) return best;
        const int64_t col = cidx(a.N, cx, cy, 0);
        const int32_t q0 = a.offsets[col + z0], q1 = a.offsets[col + z1 + 1];
        for (int32_t q = q0; q < q1; ++q) {
            const float4 p = a.sp[q];
            const float dx = x - p.x, dy = y - p.y, dz = z - p.z;
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);
            best = d < best ? d : best;
        }
        return best;
    }
    static __device__ void run(const Args& a) {
        const int M = a.N + 1;
        const int64_t n = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (n >= (int64_t)M * M * M) return;
        const int k = (int)(n % M), j = (int)((n / M) % M), i = (int)(n / ((int64_t)M * M));
        const float x = (float)i, y = (float)j, z = (float)k;
        float best = a.band + a.rad;   // at or beyond: phi is the band value
        for (int s = 0; s < kShells; ++s) {
            const int x0 = max(i - 1 - s, 0), x1 = min(i + s, a.N - 1);
            const int y0 = max(j - 1 - s, 0), y1 = min(j + s, a.N - 1);
            for (int cx = x0; cx <= x1; ++cx) {
                for (int cy = y0; cy <= y1; ++cy) {
                    const bool inner = s > 0 && cx >= i - s && cx <= i + s - 1 && cy >= j - s &&
                                       cy <= j + s - 1;
                    if (inner) {
                        best = column_min(a, best, cx, cy, k - 1 - s, k - 1 - s, x, y, z);
                        best = column_min(a, best, cx, cy, k + s, k + s, x, y, z);
                    } else {
                        best = column_min(a, best, cx, cy, k - 1 - s, k + s, x, y, z);
                    }
                }
            }
            if (best < (float)(s + 1) - 1.0e-3f) break;
        }
        const float v = best - a.rad;
        a.phi[n] = v < a.band ? v : a.band;
    }
};

// Velocity trails (LevelSet(particles, maxd) calls rasterizeTrails,
// levelset.cpp:109; OpenVDB's ParticlesToLevelSet, not in this image, restated
// from its published source). Per particle, radius R0 = 0.5 (getPosRadVel,
// levelset.hpp:55-67) at P0 = p * maxd, trail direction n = -v/|v| with v the
// particle velocity as given (unit-box units per second, unscaled, as the
// reference passes it): spheres at P, radius R, starting at (P0, R0), each
// followed by P += 0.5 R n, d = |P - P0|, R = R0 - (R0 - Rmin) d / |v|, while
// d <= |v| (Rmin = 0.01, raster.setRmin at levelset.cpp:104; a particle at rest
// gives the one sphere at P0). A sphere stamps every node c with |c - P|^2 <
// (R + w)^2 (w = 3, the narrow band) with float(|c - P| - R) when smaller (CSG
// union); a node at |c - P|^2 <= max(0, R - w)^2 becomes -w. Arithmetic in double
// like OpenVDB's (Vec3R, Real); nodes outside [0, N]^3 are not stored. The
// union is a minimum, taken with integer atomics on an order-preserving key of
// the float, so the result does not depend on the order of the particles.
__device__ __forceinline__ int32_t float_key(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float key_float(int32_t k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

struct KTrailSdf {
    static constexpr int kThreads = kBlock;
    struct Args {
        const float4* pos;
        const float4* vel;
        int64_t np;
        float maxd;
        int M;             // nodes per axis (N + 1)
        double band, rmin, r0;
        int32_t* key;      // M^3 keys of phi, initialised to the key of band
    };
    static __device__ void sphere(const Args& a, double px, double py, double pz, double R) {
        const double mx = R + a.band, max2 = mx * mx;
        const double inner = R - a.band > 0.0 ? R - a.band : 0.0, min2 = inner * inner;
        const int x0 = max((int)floor(px - mx), 0), x1 = min((int)ceil(px + mx), a.M - 1);
        const int y0 = max((int)floor(py - mx), 0), y1 = min((int)ceil(py + mx), a.M - 1);
        const int z0 = max((int)floor(pz - mx), 0), z1 = min((int)ceil(pz + mx), a.M - 1);
        for (int x = x0; x <= x1; ++x) {
            const double x2 = (x - px) * (x - px);
            for (int y = y0; y <= y1; ++y) {
                const double x2y2 = x2 + (y - py) * (y - py);
                for (int z = z0; z <= z1; ++z) {
                    const double d2 = x2y2 + (z - pz) * (z - pz);
                    if (d2 >= max2) continue;
                    const float v = d2 <= min2 ? (float)-a.band : (float)(sqrt(d2) - R);
                    const int32_t k = float_key(v);
                    int32_t* slot = a.key + ((int64_t)x * a.M + y) * a.M + z;
                    if (k < *slot) atomicMin(slot, k);
                }
            }
        }
    }
    static __device__ void run(const Args& a) {
        const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (s >= a.np) return;
        const float4 p = a.pos[s];
        const float4 u = a.vel[s];
        // positions scaled in float (the reference's Vec3f), then widened
        const double P0x = (double)(p.x * a.maxd), P0y = (double)(p.y * a.maxd), P0z = (double)(p.z * a.maxd);
        const double vx = (double)u.x, vy = (double)u.y, vz = (double)u.z;
        const double speed = sqrt(vx * vx + vy * vy + vz * vz), inv = 1.0 / speed;
        const double nx = -vx * inv, ny = -vy * inv, nz = -vz * inv;
        const double R0 = a.r0;
        if (R0 < a.rmin) return;
        double Px = P0x, Py = P0y, Pz = P0z, R = R0, d = 0.0;
        // a sphere per iteration; the trail ends once d exceeds the speed (at
        // rest the direction is NaN and the first step ends it)
        for (int m = 0; d <= speed && m < 1 << 20; ++m) {
            sphere(a, Px, Py, Pz, R);
            Px += 0.5 * R * nx;
            Py += 0.5 * R * ny;
            Pz += 0.5 * R * nz;
            const double ex = Px - P0x, ey = Py - P0y, ez = Pz - P0z;
            d = sqrt(ex * ex + ey * ey + ez * ez);
            R = R0 - (R0 - a.rmin) * d * inv;
        }
    }
};

struct KKeyToFloat {
    static constexpr int kThreads = kBlock;
    struct Args {
        int32_t* key;
        int64_t n;
    };
    static __device__ void run(const Args& a) {
        const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (i < a.n) a.key[i] = __float_as_int(key_float(a.key[i]));
    }
};

struct KFillKey {
    static constexpr int kThreads = kBlock;
    struct Args {
        int32_t* key;
        int64_t n;
        float v;
    };
    static __device__ void run(const Args& a) {
        const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
        if (i < a.n) a.key[i] = float_key(a.v);
    }
};

}  // namespace

void launch_particle_sdf(hipStream_t s, const float4* pos, int64_t np, int N, float maxd, float band,
                         float rad, SdfScratch& w, float* phi) {
    const int64_t nc = (int64_t)N * N * N;
    const int64_t M = N + 1;
    launch_fill_i32(s, w.counts, nc, 0);
    if (np > 0)
        launch<KSdfBin>(s, dim3(grid_for(np, kBlock)), dim3(kBlock), {pos, np, N, maxd, w.cell_of, w.rank, w.counts});
    launch_scan(s, w.counts, w.offsets, nc, w.scan_sums);
    if (np > 0)
        launch<KSdfScatter>(s, dim3(grid_for(np, kBlock)), dim3(kBlock),
                            {pos, np, maxd, w.cell_of, w.rank, w.offsets, w.sorted});
    launch<KParticleSdf>(s, dim3(grid_for(M * M * M, kBlock)), dim3(kBlock), {w.sorted, w.offsets, N, band, rad, phi});
}

void launch_trail_sdf(hipStream_t s, const float4* pos, const float4* vel, int64_t np, int N, float maxd,
                      float band, float rad, float rmin, float* phi) {
    const int64_t M = N + 1, nm = M * M * M;
    int32_t* key = reinterpret_cast<int32_t*>(phi);
    launch<KFillKey>(s, dim3(grid_for(nm, kBlock)), dim3(kBlock), {key, nm, band});
    if (np > 0)
        launch<KTrailSdf>(s, dim3(grid_for(np, kBlock)), dim3(kBlock),
                          {pos, vel, np, maxd, (int)M, (double)band, (double)rmin, (double)rad, key});
    launch<KKeyToFloat>(s, dim3(grid_for(nm, kBlock)), dim3(kBlock), {key, nm});
}

// ---------------------------------------------------------------- host side

void surface_nets(const float* phi, int M, std::vector<float>& verts, std::vector<int32_t>& quads) {
    verts.clear();
    quads.clear();
    if (M < 2) return;
    const int C = M - 1;   // voxels per axis
    auto node = [&](int i, int j, int k) { return phi[((int64_t)i * M + j) * M + k]; };
    std::vector<int32_t> vid((size_t)C * C * C, -1);
    // corner c of a voxel: offset (c >> 2 & 1, c >> 1 & 1, c & 1); the 12 edges as corner pairs
    static const int kEdge[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                     {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < C; ++j)
            for (int k = 0; k < C; ++k) {
                float v[8];
                int inside = 0;
                for (int c = 0; c < 8; ++c) {
                    v[c] = node(i + (c >> 2 & 1), j + (c >> 1 & 1), k + (c & 1));
                    inside += v[c] < 0.0f;
                }
                if (inside == 0 || inside == 8) continue;
                double sx = 0, sy = 0, sz = 0;
                int ncross = 0;
                for (const auto& e : kEdge) {
                    const float va = v[e[0]], vb = v[e[1]];
                    if ((va < 0.0f) == (vb < 0.0f)) continue;
                    const double t = (double)va / ((double)va - (double)vb);
                    const int a = e[0], b = e[1];
                    sx += (a >> 2 & 1) + t * ((b >> 2 & 1) - (a >> 2 & 1));
                    sy += (a >> 1 & 1) + t * ((b >> 1 & 1) - (a >> 1 & 1));
                    sz += (a & 1) + t * ((b & 1) - (a & 1));
                    ++ncross;
                }
                vid[((size_t)i * C + j) * C + k] = (int32_t)(verts.size() / 3);
                verts.push_back((float)(i + sx / ncross));
                verts.push_back((float)(j + sy / ncross));
                verts.push_back((float)(k + sz / ncross));
            }
    auto vx = [&](int i, int j, int k) { return vid[((size_t)i * C + j) * C + k]; };
    auto emit = [&](bool out_positive, int32_t a, int32_t b, int32_t c, int32_t d) {
        if (out_positive) quads.insert(quads.end(), {a, b, c, d});
        else quads.insert(quads.end(), {a, d, c, b});
    };
    // a node edge whose ends differ in sign is shared by 4 voxels, each holding a
    // vertex; the quad through them faces from the negative end to the positive
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j)
            for (int k = 0; k < M; ++k) {
                const bool in = node(i, j, k) < 0.0f;
                if (i + 1 < M && j >= 1 && k >= 1 && j < C && k < C && in != (node(i + 1, j, k) < 0.0f))
                    emit(in, vx(i, j - 1, k - 1), vx(i, j, k - 1), vx(i, j, k), vx(i, j - 1, k));
                if (j + 1 < M && i >= 1 && k >= 1 && i < C && k < C && in != (node(i, j + 1, k) < 0.0f))
                    emit(in, vx(i - 1, j, k - 1), vx(i - 1, j, k), vx(i, j, k), vx(i, j, k - 1));
                if (k + 1 < M && i >= 1 && j >= 1 && i < C && j < C && in != (node(i, j, k + 1) < 0.0f))
                    emit(in, vx(i - 1, j - 1, k), vx(i, j - 1, k), vx(i, j, k), vx(i - 1, j, k));
            }
}

bool write_obj(const char* path, const std::vector<float>& verts, const std::vector<int32_t>& quads) {
    FILE* f = std::fopen(path, "w");
    if (!f) return false;
    bool ok = true;
    // Obj::WriteObj (src/geom/obj/obj.cpp:177-227): "v x y z", then "f a b c d"
    // with 1-based vertex indices
    for (size_t v = 0; ok && v < verts.size(); v += 3)
        ok = std::fprintf(f, "v %.9g %.9g %.9g\n", verts[v], verts[v + 1], verts[v + 2]) > 0;
    for (size_t q = 0; ok && q < quads.size(); q += 4)
        ok = std::fprintf(f, "f %d %d %d %d\n", quads[q] + 1, quads[q + 1] + 1, quads[q + 2] + 1,
                          quads[q + 3] + 1) > 0;
    return std::fclose(f) == 0 && ok;
}

}  // namespace ariel
