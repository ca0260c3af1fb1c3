// scene_geom.hip — the reference's scene geometry on the host (src/geom,
// src/scene/scene.cpp:257-455): triangle/quad meshes read from OBJ files or made
// by the box and sphere generators, per-frame mesh and transform sequences with
// frame offset, interval and persistence, inside tests by ray parity, and the
// liquid particles each geom emits on a frame. Host code only; the emitted
// particles are uploaded and join the simulation on the GPU.
//
// Arithmetic follows the reference's float expressions:
//  - transforms T R S with R = Rz Ry Rx, angles in degrees
//    (utilities.inl:191-209, buildRotation :229-248), column-major products;
//  - the box (CubeGen::Tesselate, cubegen.cpp:22-111) and the 20 x 20 UV
//    sphere (SphereGen::Tesselate, spheregen.cpp) as the reference tessellates
//    them, vertices placed by T S in float;
//  - inside test (Scene::CheckPointInsideGeomByID, scene.cpp:362-378): a ray
//    from the point along +z, taken into object space by the inverse transform
//    (Ray::Transform: origin and origin + direction transformed, their
//    difference, normalised, the direction), counting every triangle or quad it meets
//    (Obj::RayTriangleTest, obj.cpp:354-400, quads as two triangles counted
//    once, obj.cpp:344-351) whose hit lies in front of the origin or at it
//    (bvh.inl:80-86); an odd count is inside. The reference decides "in front"
//    by acos(dot(normalize(hit - o), d)) < pi/2 or NaN. Every hit lies on the
//    ray, so that cosine is +-1 up to rounding, and a hit behind the origin
//    whose rounded cosine falls below -1 gives NaN and is counted: for rays
//    that are not axis-aligned in object space (rotated geoms) about half of
//    the hits behind the origin count, and points outside the mesh are
//    emitted. Here the test is what the reference's own comment there asks
//    for ("hit is actually in front of origin"): dot(hit - o, d) > 0 or
//    hit == o, the reference's outcome whenever its cosine rounds to +-1
//    exactly. Instead of a BVH the primitives are binned by their projection
//    along the ray direction (RayCull); both only skip primitives the ray
//    misses.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ariel.h"
#include "ariel_device.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

constexpr double kPi = 3.1415926535897932384626422832795028841971;

struct M4 {
    float m[4][4];   // m[column][row], as glm
};

M4 identity() {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = c == k ? 1.0f : 0.0f;
    return r;
}

// glm's mat4 product: column c of A*B is sum_k A[k] * B[c][k], in k order
M4 mul(const M4& A, const M4& B) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            float s = A.m[0][row] * B.m[c][0];
            s = s + A.m[1][row] * B.m[c][1];
            s = s + A.m[2][row] * B.m[c][2];
            s = s + A.m[3][row] * B.m[c][3];
            r.m[c][row] = s;
        }
    return r;
}

// utilityCore::multiply (utilities.inl:211-218)
void apply(const M4& m, const float v[4], float out[4]) {
    for (int row = 0; row < 4; ++row)
        out[row] = m.m[0][row] * v[0] + m.m[1][row] * v[1] + m.m[2][row] * v[2] + m.m[3][row] * v[3];
}

float to_radian(float degree) { return (float)(degree * kPi / 180.0f); }

M4 translation(const float t[3]) {
    M4 r = identity();
    r.m[3][0] = t[0];
    r.m[3][1] = t[1];
    r.m[3][2] = t[2];
    return r;
}
M4 scaling(const float s[3]) {
    M4 r = identity();
    r.m[0][0] = s[0];
    r.m[1][1] = s[1];
    r.m[2][2] = s[2];
    return r;
}
// buildRotation(radian, axis) for a unit coordinate axis
M4 rotation(float radian, float ax, float ay, float az) {
    const float c = std::cos(radian), s = std::sin(radian);
    M4 m = identity();
    m.m[0][0] = c + (1.0f - c) * ax * ax;
    m.m[0][1] = (1.0f - c) * ax * ay + s * az;
    m.m[0][2] = (1.0f - c) * ax * az - s * ay;
    m.m[1][0] = (1.0f - c) * ay * ax - s * az;
    m.m[1][1] = c + (1.0f - c) * ay * ay;
    m.m[1][2] = (1.0f - c) * ay * az + s * ax;
    m.m[2][0] = (1.0f - c) * az * ax + s * ay;
    m.m[2][1] = (1.0f - c) * az * ay - s * ax;
    m.m[2][2] = c + (1.0f - c) * az * az;
    return m;
}

M4 build_transform(const float t[3], const float r[3], const float s[3]) {
    M4 R = rotation(to_radian(r[2]), 0, 0, 1);
    R = mul(R, rotation(to_radian(r[1]), 0, 1, 0));
    R = mul(R, rotation(to_radian(r[0]), 1, 0, 0));
    return mul(mul(translation(t), R), scaling(s));
}
M4 build_inverse(const float t[3], const float r[3], const float s[3]) {
    const float nt[3] = {-t[0], -t[1], -t[2]};
    M4 R = rotation(to_radian(-r[0]), 1, 0, 0);
    R = mul(R, rotation(to_radian(-r[1]), 0, 1, 0));
    R = mul(R, rotation(to_radian(-r[2]), 0, 0, 1));
    const float is[3] = {1.0f / s[0], 1.0f / s[1], 1.0f / s[2]};
    return mul(mul(scaling(is), R), translation(nt));
}

void place(GeomMesh& g, const M4& T) {
    for (auto& v : g.v) {
        const float in[4] = {v[0], v[1], v[2], 1.0f};
        float o[4];
        apply(T, in, o);
        v = {o[0], o[1], o[2]};
    }
}

void bounds(GeomMesh& g) {
    g.lo = {1e30f, 1e30f, 1e30f};
    g.hi = {-1e30f, -1e30f, -1e30f};
    for (const auto& v : g.v)
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = std::min(g.lo[a], v[a]);
            g.hi[a] = std::max(g.hi[a], v[a]);
        }
}

// Obj::RayTriangleTest: hit with barycentrics in range (no test on t); the hit
// point o + t d
bool ray_triangle(const std::array<float, 3>& v0, const std::array<float, 3>& v1,
                  const std::array<float, 3>& v2, const float o[3], const float d[3], float hit[3]) {
    const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const float p[3] = {d[1] * e2[2] - e2[1] * d[2], d[2] * e2[0] - e2[2] * d[0], d[0] * e2[1] - e2[0] * d[1]};
    const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0f) return false;
    const float inv = 1.0f / det;
    const float tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const float bx = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * inv;
    const float q[3] = {tv[1] * e1[2] - e1[1] * tv[2], tv[2] * e1[0] - e1[2] * tv[0], tv[0] * e1[1] - e1[0] * tv[1]};
    const float by = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    const float t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    if (!(bx >= 0.0f && by >= 0.0f && (bx + by) <= 1.0f)) return false;
    for (int a = 0; a < 3; ++a) hit[a] = o[a] + t * d[a];
    return true;
}

}  // namespace

// ---------------------------------------------------------------- meshes
bool read_obj_mesh(const std::string& path, GeomMesh& g, std::string& err) {
    std::ifstream in(path);
    if (!in.is_open()) {
        err = "cannot open mesh file " + path;
        return false;
    }
    g.v.clear();
    g.f.clear();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            std::string a, b, c;
            ss >> a >> b >> c;
            g.v.push_back({(float)std::atof(a.c_str()), (float)std::atof(b.c_str()), (float)std::atof(c.c_str())});
        } else if (tag == "f") {
            // up to four vertices (Obj::ReadObj, obj.cpp:108-131); "a/b/c" takes a
            std::array<int, 4> f = {0, 0, 0, 0};
            std::string tok;
            int n = 0;
            while (n < 4 && (ss >> tok)) f[n++] = std::atoi(tok.c_str());
            if (n < 3) continue;
            g.f.push_back(f);
        }
    }
    for (const auto& f : g.f)
        for (int k = 0; k < 4; ++k)
            if (f[k] < 0 || f[k] > (int)g.v.size() || (k < 3 && f[k] == 0)) {
                err = "mesh file " + path + ": a face index is out of range";
                return false;
            }
    bounds(g);
    return true;
}

void cube_mesh(GeomMesh& g, const float p0[3], const float p1[3]) {
    static const float V[8][3] = {{-0.5f, -0.5f, 0.5f}, {0.5f, -0.5f, 0.5f}, {-0.5f, 0.5f, 0.5f},
                                  {0.5f, 0.5f, 0.5f},   {-0.5f, 0.5f, -0.5f}, {0.5f, 0.5f, -0.5f},
                                  {-0.5f, -0.5f, -0.5f}, {0.5f, -0.5f, -0.5f}};
    static const int F[6][4] = {{1, 2, 4, 3}, {3, 4, 6, 5}, {5, 6, 8, 7}, {7, 8, 2, 1}, {2, 8, 6, 4}, {7, 1, 3, 5}};
    g.v.clear();
    g.f.clear();
    for (const auto& v : V) g.v.push_back({v[0], v[1], v[2]});
    for (const auto& f : F) g.f.push_back({f[0], f[1], f[2], f[3]});
    const float sc[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const float ce[3] = {(p1[0] + p0[0]) / 2.0f, (p1[1] + p0[1]) / 2.0f, (p1[2] + p0[2]) / 2.0f};
    const float zero[3] = {0, 0, 0};
    place(g, build_transform(ce, zero, sc));
    bounds(g);
}

void sphere_mesh(GeomMesh& g, const float center[3], float radius, int subdivs) {
    const unsigned a = (unsigned)std::max(subdivs, 3), h = a;
    const unsigned vert_count = a * (h - 1) + 2, face_count = a * h;
    g.v.assign(vert_count, {0, 0, 0});
    g.f.assign(face_count, {0, 0, 0, 0});
    for (unsigned x = 1; x < h; ++x)
        for (unsigned y = 0; y < a; ++y) {
            const float a1 = float(x) * float(1.0 / a), a2 = float(y) * float(1.0 / h);
            const float px = (float)(std::sin(kPi * a1) * std::cos(2.0 * kPi * a2));
            const float py = (float)(std::sin(kPi * a1) * std::sin(2.0 * kPi * a2));
            const float pz = (float)std::cos(kPi * a1);
            // normalize(vec3(x, z, y)) / 2
            const float len = std::sqrt(px * px + pz * pz + py * py);
            g.v[(x - 1) * a + y] = {px / len / 2.0f, pz / len / 2.0f, py / len / 2.0f};
        }
    g.v[vert_count - 2] = {0.0f, -0.5f, 0.0f};
    g.v[vert_count - 1] = {0.0f, 0.5f, 0.0f};
    // sides
    for (unsigned x = 1; x < h - 1; ++x)
        for (unsigned y = 0; y < a; ++y) {
            const unsigned i1 = (x - 1) * a + y;
            unsigned i2 = (x - 1) * a + (y + 1), i3 = x * a + (y + 1);
            const unsigned i4 = x * a + y;
            if (y == a - 1) {
                i2 = (x - 1) * a;
                i3 = x * a;
            }
            g.f[i1] = {(int)i1 + 1, (int)i2 + 1, (int)i3 + 1, (int)i4 + 1};
        }
    // top pole: the first row's faces, as triangles with the top vertex
    for (unsigned x = 0; x < h; ++x) {
        std::array<int, 4> ind = g.f[x];
        ind[3] = 0;
        ind[2] = ind[0];
        ind[0] = (int)vert_count;
        g.f[face_count - a * 2 + x] = ind;
    }
    // bottom pole: the last row's faces with the bottom vertex
    for (unsigned x = 0; x < h; ++x) {
        std::array<int, 4> ind = g.f[(h - 1) * (a - 2) - 2 + x];
        ind[1] = ind[2];
        ind[0] = ind[3];
        ind[3] = 0;
        ind[2] = (int)vert_count - 1;
        g.f[face_count - a + x] = ind;
    }
    const float sc[3] = {radius * 2.0f, radius * 2.0f, radius * 2.0f};
    const float zero[3] = {0, 0, 0};
    place(g, build_transform(center, zero, sc));
    bounds(g);
}

// ---------------------------------------------------------------- geoms
// MeshContainer::IsInFrame / GetMeshFrame / GetTransforms (mesh.cpp:61-109)
bool geom_in_frame(const GeomNode& n, float frame) {
    const float cf = (frame - float(n.offset)) / float(n.interval);
    return !((cf < 0.0f && !n.pre_persist) || (cf >= (float)n.frames() && !n.post_persist));
}

bool geom_at(const SceneGeom& s, const GeomNode& n, float frame, int& mesh, float M[16], float Minv[16]) {
    if (!geom_in_frame(n, frame)) return false;
    float cf = (frame - float(n.offset)) / float(n.interval);
    cf = std::min(std::max(cf, 0.0f), float(n.frames() - 1));
    const unsigned upper = (unsigned)std::ceil(cf), lower = (unsigned)std::floor(cf);
    const float w = cf - float(lower);
    const GeomXform& A = s.xforms[n.xform[lower]];
    const GeomXform& B = s.xforms[n.xform[upper]];
    float t[3], r[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        t[k] = A.t[k] * (1.0f - w) + B.t[k] * w;
        r[k] = A.r[k] * (1.0f - w) + B.r[k] * w;
        sc[k] = A.s[k] * (1.0f - w) + B.s[k] * w;
    }
    const M4 f = build_transform(t, r, sc), b = build_inverse(t, r, sc);
    std::memcpy(M, f.m, sizeof(f.m));
    std::memcpy(Minv, b.m, sizeof(b.m));
    mesh = n.mesh[lower];
    return true;
}

// Point p (grid units) inside geom n at `frame` (CheckPointInsideGeomByID).
namespace {

// The inside test of one geom at one frame, prepared for many points: the
// inverse transform once, and the primitives binned by their projection on the
// plane normal to the object-space ray direction d0 (every point's ray is +z in
// world space, so in object space it runs along d0 up to rounding). A primitive
// can report a hit only for origins whose projection lies in its projected
// bounds widened by a margin that covers the rounding of the hit test
// (barycentrics are quotients by the projected area, so thin primitives get
// wide margins); primitives whose margin would cover much of the grid, or whose
// projection is nearly degenerate, are tested for every point. The count of
// hits, hence the result, is the one a test of every primitive gives.
struct RayCull {
    const GeomMesh* g = nullptr;
    M4 inv;
    double u[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    double x0 = 0, y0 = 0, inv_cs = 0;
    int nx = 0, ny = 0;
    std::vector<int> start, cand, always;
};

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

bool prepare_cull(const SceneGeom& s, const GeomNode& n, float frame, const float qlo[3],
                  const float qhi[3], RayCull& rc) {
    int mi;
    float Mf[16], Mi[16];
    if (!geom_at(s, n, frame, mi, Mf, Mi)) return false;
    std::memcpy(rc.inv.m, Mi, sizeof(rc.inv.m));
    const GeomMesh& g = s.meshes[mi];
    rc.g = &g;
    // object-space direction of world +z, and a basis (u, v) of its normal plane
    const float z0[4] = {0, 0, 0, 1}, z1[4] = {0, 0, 1, 1};
    float a0[4], a1[4];
    apply(rc.inv, z0, a0);
    apply(rc.inv, z1, a1);
    double d0[3] = {(double)a1[0] - a0[0], (double)a1[1] - a0[1], (double)a1[2] - a0[2]};
    const double dl = std::sqrt(dot3(d0, d0));
    for (double& c : d0) c /= dl;
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (std::fabs(d0[k]) < std::fabs(d0[ax])) ax = k;
    double e[3] = {0, 0, 0};
    e[ax] = 1.0;
    double* u = rc.u;
    double* v = rc.v;
    u[0] = d0[1] * e[2] - d0[2] * e[1];
    u[1] = d0[2] * e[0] - d0[0] * e[2];
    u[2] = d0[0] * e[1] - d0[1] * e[0];
    const double ul = std::sqrt(dot3(u, u));
    for (int k = 0; k < 3; ++k) u[k] /= ul;
    v[0] = d0[1] * u[2] - d0[2] * u[1];
    v[1] = d0[2] * u[0] - d0[0] * u[2];
    v[2] = d0[0] * u[1] - d0[1] * u[0];
    // D bounds the object-space coordinates of every origin and vertex
    double D = 0.0;
    for (int c = 0; c < 8; ++c) {
        const float w4[4] = {(c & 4) ? qhi[0] : qlo[0], (c & 2) ? qhi[1] : qlo[1], (c & 1) ? qhi[2] : qlo[2], 1.0f};
        float o4[4];
        apply(rc.inv, w4, o4);
        D = std::max(D, std::sqrt((double)o4[0] * o4[0] + (double)o4[1] * o4[1] + (double)o4[2] * o4[2]));
    }
    for (const auto& p : g.v) D = std::max(D, std::sqrt((double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2]));
    D = 2.0 * D + 1e-30;
    const size_t F = g.f.size();
    std::vector<double> box(4 * F);
    std::vector<char> wide(F, 0);
    double gx0 = 1e300, gy0 = 1e300, gx1 = -1e300, gy1 = -1e300;
    for (size_t fi = 0; fi < F; ++fi) {
        const auto& f = g.f[fi];
        const int nv = f[3] != 0 ? 4 : 3;
        double px[4], py[4];
        for (int k = 0; k < nv; ++k) {
            const double q[3] = {g.v[f[k] - 1][0], g.v[f[k] - 1][1], g.v[f[k] - 1][2]};
            px[k] = dot3(q, u);
            py[k] = dot3(q, v);
        }
        // the smaller projected area of the primitive's triangles (a quad is
        // (0, 1, 2) and (0, 3, 2)) and its longest projected edge
        auto area2 = [&](int i, int j, int k) {
            return std::fabs((px[j] - px[i]) * (py[k] - py[i]) - (px[k] - px[i]) * (py[j] - py[i]));
        };
        double A = area2(0, 1, 2);
        if (nv == 4) A = std::min(A, area2(0, 3, 2));
        double L2 = 0.0;
        for (int i = 0; i < nv; ++i)
            for (int j = i + 1; j < nv; ++j)
                L2 = std::max(L2, (px[j] - px[i]) * (px[j] - px[i]) + (py[j] - py[i]) * (py[j] - py[i]));
        double lo0 = px[0], hi0 = px[0], lo1 = py[0], hi1 = py[0];
        for (int k = 1; k < nv; ++k) {
            lo0 = std::min(lo0, px[k]);
            hi0 = std::max(hi0, px[k]);
            lo1 = std::min(lo1, py[k]);
            hi1 = std::max(hi1, py[k]);
        }
        if (!(A >= 1e-3 * L2) || !(L2 > 0.0)) {
            wide[fi] = 1;
            continue;
        }
        const double m = 1e-4 * D * (1.0 + L2 / A);
        box[4 * fi + 0] = lo0 - m;
        box[4 * fi + 1] = hi0 + m;
        box[4 * fi + 2] = lo1 - m;
        box[4 * fi + 3] = hi1 + m;
        gx0 = std::min(gx0, box[4 * fi + 0]);
        gx1 = std::max(gx1, box[4 * fi + 1]);
        gy0 = std::min(gy0, box[4 * fi + 2]);
        gy1 = std::max(gy1, box[4 * fi + 3]);
    }
    const int side = std::max(1, std::min(512, (int)std::sqrt((double)F * 2.0)));
    const double span = std::max(gx1 - gx0, gy1 - gy0);
    if (!(span > 0.0) || !std::isfinite(span)) {
        for (size_t fi = 0; fi < F; ++fi) rc.always.push_back((int)fi);
        return true;
    }
    rc.x0 = gx0;
    rc.y0 = gy0;
    rc.inv_cs = side / span;
    rc.nx = std::max(1, (int)std::ceil((gx1 - gx0) * rc.inv_cs));
    rc.ny = std::max(1, (int)std::ceil((gy1 - gy0) * rc.inv_cs));
    const int64_t ncell = (int64_t)rc.nx * rc.ny;
    auto cell_range = [&](size_t fi, int& i0, int& i1, int& j0, int& j1) {
        i0 = std::max(0, (int)std::floor((box[4 * fi + 0] - rc.x0) * rc.inv_cs));
        i1 = std::min(rc.nx - 1, (int)std::floor((box[4 * fi + 1] - rc.x0) * rc.inv_cs));
        j0 = std::max(0, (int)std::floor((box[4 * fi + 2] - rc.y0) * rc.inv_cs));
        j1 = std::min(rc.ny - 1, (int)std::floor((box[4 * fi + 3] - rc.y0) * rc.inv_cs));
    };
    std::vector<int> count(ncell + 1, 0);
    for (size_t fi = 0; fi < F; ++fi) {
        if (wide[fi]) continue;
        int i0, i1, j0, j1;
        cell_range(fi, i0, i1, j0, j1);
        if ((int64_t)(i1 - i0 + 1) * (j1 - j0 + 1) * 4 > ncell) {
            wide[fi] = 1;
            continue;
        }
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) ++count[(int64_t)i * rc.ny + j];
    }
    rc.start.assign(ncell + 1, 0);
    for (int64_t c = 0; c < ncell; ++c) rc.start[c + 1] = rc.start[c] + count[c];
    rc.cand.resize(rc.start[ncell]);
    std::vector<int> fill(rc.start.begin(), rc.start.end() - 1);
    for (size_t fi = 0; fi < F; ++fi) {
        if (wide[fi]) {
            rc.always.push_back((int)fi);
            continue;
        }
        int i0, i1, j0, j1;
        cell_range(fi, i0, i1, j0, j1);
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) rc.cand[fill[(int64_t)i * rc.ny + j]++] = (int)fi;
    }
    return true;
}

bool cull_contains(const RayCull& rc, const float p[3]) {
    const float o4[4] = {p[0], p[1], p[2], 1.0f}, e4[4] = {p[0], p[1], p[2] + 1.0f, 1.0f};
    float o[4], e[4];
    apply(rc.inv, o4, o);
    apply(rc.inv, e4, e);
    // the Ray constructor normalises its direction (ray.cpp:24-30)
    float d[3] = {e[0] - o[0], e[1] - o[1], e[2] - o[2]};
    const float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (float& c : d) c = c / dl;
    const GeomMesh& g = *rc.g;
    // in front of the origin, or at it: n = hit - o points along d, or is zero
    auto front = [&](const float hit[3]) {
        const float n3[3] = {hit[0] - o[0], hit[1] - o[1], hit[2] - o[2]};
        if (n3[0] == 0.0f && n3[1] == 0.0f && n3[2] == 0.0f) return true;
        return n3[0] * d[0] + n3[1] * d[1] + n3[2] * d[2] > 0.0f;
    };
    auto hits_face = [&](int fi) {
        const auto& f = g.f[fi];
        float hit[3];
        const auto& v0 = g.v[f[0] - 1];
        const auto& v2 = g.v[f[2] - 1];
        bool h = ray_triangle(v0, g.v[f[1] - 1], v2, o, d, hit);
        if (!h && f[3] != 0) h = ray_triangle(v0, g.v[f[3] - 1], v2, o, d, hit);
        return h && front(hit);
    };
    unsigned hits = 0;
    for (int fi : rc.always) hits += hits_face(fi) ? 1u : 0u;
    if (rc.nx > 0) {
        const double od[3] = {o[0], o[1], o[2]};
        const double pu = dot3(od, rc.u), pv = dot3(od, rc.v);
        const double fx = std::floor((pu - rc.x0) * rc.inv_cs), fy = std::floor((pv - rc.y0) * rc.inv_cs);
        if (fx >= 0.0 && fy >= 0.0 && fx < rc.nx && fy < rc.ny) {
            const int64_t c = (int64_t)fx * rc.ny + (int64_t)fy;
            for (int k = rc.start[c]; k < rc.start[c + 1]; ++k) hits += hits_face(rc.cand[k]) ? 1u : 0u;
        }
    }
    return hits % 2 == 1;
}

}  // namespace

bool geom_contains(const SceneGeom& s, const GeomNode& n, float frame, const float p[3]) {
    RayCull rc;
    if (!prepare_cull(s, n, frame, p, p, rc)) return false;
    return cull_contains(rc, p);
}

// World-space bounds of geom n at `frame` (MeshContainer::GetAabb, Aabb::Transform).
bool geom_bounds(const SceneGeom& s, const GeomNode& n, float frame, float lo[3], float hi[3]) {
    int mi;
    float Mf[16], Mi[16];
    if (!geom_at(s, n, frame, mi, Mf, Mi)) return false;
    M4 T;
    std::memcpy(T.m, Mf, sizeof(T.m));
    const GeomMesh& g = s.meshes[mi];
    for (int a = 0; a < 3; ++a) {
        lo[a] = 1e30f;
        hi[a] = -1e30f;
    }
    for (int c = 0; c < 8; ++c) {
        const float in[4] = {(c & 4) ? g.hi[0] : g.lo[0], (c & 2) ? g.hi[1] : g.lo[1], (c & 1) ? g.hi[2] : g.lo[2], 1.0f};
        float o[4];
        apply(T, in, o);
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], o[a]);
            hi[a] = std::max(hi[a], o[a]);
        }
    }
    return true;
}

// The liquid particles the scene's liquid geoms emit on `frame`
// (Scene::GenerateParticles + AddLiquidParticle, scene.cpp:257-300, 436-455),
// in the deterministic order of the built-in liquids: liquids in scene order,
// interior cells of the geom's bounds in i, j, k order, lattice octants
// lexicographically, a lattice point at a*w + u*w with u = 0.5 or the seeded
// octant jitter (salt 3l + 97 frame: at frame 0 the built-in liquids' jitter);
// a point is kept when the geom contains it.
void emit_geom_liquids(const SceneGeom& s, const ariel_scene_desc& d, int N, int frame,
                       std::vector<float4>& pos, std::vector<float4>& vel) {
    const float maxd = (float)N;
    const float w = d.density * (1.0f / maxd);
    const int m = (int)std::lround(1.0 / d.density);
    const uint32_t Nm = (uint32_t)(N * m);
    const int wall = d.wall;
    for (size_t l = 0; l < s.liquids.size(); ++l) {
        const GeomLiquid& lq = s.liquids[l];
        const GeomNode& gn = s.geoms[lq.geom];
        float lo[3], hi[3];
        if (!geom_bounds(s, gn, (float)frame, lo, hi)) continue;
        int c0[3], c1[3];
        for (int a = 0; a < 3; ++a) {
            c0[a] = std::max(wall, (int)std::floor(lo[a]) - 1);
            c1[a] = std::min(N - wall, (int)std::ceil(hi[a]) + 1);
        }
        const int span = std::max(0, c1[0] - c0[0]);
        if (span == 0 || c1[1] <= c0[1] || c1[2] <= c0[2]) continue;
        const float qlo[3] = {(float)c0[0], (float)c0[1], (float)c0[2]};
        const float qhi[3] = {(float)c1[0], (float)c1[1], (float)c1[2]};
        RayCull rc;
        if (!prepare_cull(s, gn, (float)frame, qlo, qhi, rc)) continue;
        const unsigned hw = std::thread::hardware_concurrency();
        const int nth = std::max(1, std::min<int>(std::min(hw, 16u), span));
        std::vector<std::vector<float4>> P(nth), V(nth);
        std::vector<std::thread> ts;
        for (int t = 0; t < nth; ++t) {
            const int i0 = c0[0] + span * t / nth, i1 = c0[0] + span * (t + 1) / nth;
            ts.emplace_back([&, t, i0, i1] {
                for (int ci = i0; ci < i1; ++ci)
                    for (int cj = c0[1]; cj < c1[1]; ++cj)
                        for (int ck = c0[2]; ck < c1[2]; ++ck)
                            for (int oa = 0; oa < m; ++oa)
                                for (int ob = 0; ob < m; ++ob)
                                    for (int oc = 0; oc < m; ++oc) {
                                        const uint32_t a = (uint32_t)(ci * m + oa), b = (uint32_t)(cj * m + ob),
                                                       c = (uint32_t)(ck * m + oc);
                                        float jx = 0.5f, jy = 0.5f, jz = 0.5f;
                                        if (d.jitter) {
                                            const uint32_t key = a * Nm + b;
                                            const uint32_t salt = 3u * (uint32_t)l + 97u * (uint32_t)frame;
                                            jx = (float)(hash4(d.seed, salt + 0u, key, c) >> 8) * (1.0f / 16777216.0f);
                                            jy = (float)(hash4(d.seed, salt + 1u, key, c) >> 8) * (1.0f / 16777216.0f);
                                            jz = (float)(hash4(d.seed, salt + 2u, key, c) >> 8) * (1.0f / 16777216.0f);
                                        }
                                        const float x = (float)a * w + jx * w, y = (float)b * w + jy * w,
                                                    z = (float)c * w + jz * w;
                                        const float world[3] = {x * maxd, y * maxd, z * maxd};
                                        if (!cull_contains(rc, world)) continue;
                                        P[t].push_back(make_float4(x, y, z, 0.0f));
                                        V[t].push_back(make_float4(lq.velocity[0], lq.velocity[1], lq.velocity[2], 10.0f));
                                    }
            });
        }
        for (auto& th : ts) th.join();
        for (int t = 0; t < nth; ++t) {
            pos.insert(pos.end(), P[t].begin(), P[t].end());
            vel.insert(vel.end(), V[t].begin(), V[t].end());
        }
    }
}

}  // namespace ariel
