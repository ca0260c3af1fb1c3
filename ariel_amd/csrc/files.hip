// files.hip — host-side writers of the export formats the reference produces
// through third-party libraries (Scene::ExportParticles, src/scene/scene.cpp:
// 34-101): the level set as an OpenVDB file (LevelSet::WriteVDBGridToFile,
// src/grid/levelset.cpp:195-201) and the particles through Partio::write
// (scene.cpp:50-82), whose format follows the file extension. Neither library
// is in this image; the formats are written from their published layouts.
//
// VDB (OpenVDB file format version 222, "node mask compression"): one FloatGrid
// of tree type Tree_float_5_4_3 (root -> 32^3 internal -> 16^3 internal ->
// 8^3 leaf), level-set class, uniform voxel size 1 (grid units, a particle at
// p * maxd). The dense node volume becomes leaves wherever a node is active
// (|phi| < background, the narrow band); every other node is an inactive
// background tile. Per-grid compression is COMPRESS_ACTIVE_MASK: a node whose
// inactive values all equal the background stores its active values only.
//
// Partio: ".geo" Houdini classic ASCII geometry, ".bgeo" its big-endian binary
// form, each optionally gzipped (".geo.gz", ".bgeo.gz"), with the attributes
// of ExportParticles: position (grid units), "v" (vector 3), "id" (int 1).
#include <zlib.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ariel.h"
#include "ariel_internal.h"

namespace ariel {

namespace {

// ------------------------------------------------------------------ output
// A byte sink over a plain file or a gzip stream.
class Sink {
  public:
    explicit Sink(const std::string& path) : gz_(ends_with(path, ".gz")) {
        if (gz_) z_ = gzopen(path.c_str(), "wb6");
        else f_ = std::fopen(path.c_str(), "wb");
    }
    ~Sink() { close(); }
    bool open() const { return gz_ ? z_ != nullptr : f_ != nullptr; }
    void put(const void* p, size_t n) {
        if (!ok_ || n == 0) return;
        if (gz_) ok_ = gzwrite(z_, p, (unsigned)n) == (int)n;
        else ok_ = std::fwrite(p, 1, n, f_) == n;
    }
    bool close() {
        if (gz_ && z_) {
            ok_ = gzclose(z_) == Z_OK && ok_;
            z_ = nullptr;
        } else if (f_) {
            ok_ = std::fclose(f_) == 0 && ok_;
            f_ = nullptr;
        }
        return ok_;
    }
    static bool ends_with(const std::string& s, const char* suf) {
        const size_t n = std::strlen(suf);
        return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
    }

  private:
    bool gz_;
    FILE* f_ = nullptr;
    gzFile z_ = nullptr;
    bool ok_ = true;
};

// little-endian buffer (OpenVDB writes the host's byte order)
struct Buf {
    std::vector<unsigned char> b;
    void raw(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    template <class T>
    void put(T v) { raw(&v, sizeof(T)); }
    void str(const std::string& s) {   // io::writeString: uint32 length, then the bytes
        put<uint32_t>((uint32_t)s.size());
        raw(s.data(), s.size());
    }
    size_t pos() const { return b.size(); }
    template <class T>
    void patch(size_t at, T v) { std::memcpy(&b[at], &v, sizeof(T)); }
};

// ------------------------------------------------------------------- VDB
constexpr int64_t kVdbMagic = 0x56444220;        // "VDB " (OPENVDB_MAGIC)
constexpr uint32_t kVdbFileVersion = 222;        // node-mask compression
constexpr uint32_t kCompressActiveMask = 0x2;    // io::COMPRESS_ACTIVE_MASK
constexpr int8_t kNoMaskOrInactiveVals = 0;      // every inactive value is the background
constexpr int8_t kNoMaskAndAllVals = 6;          // every value stored

// node extents of Tree_float_5_4_3 (log2 of the children per axis)
constexpr int kLeafLog2 = 3, kLowerLog2 = 4, kUpperLog2 = 5;
constexpr int kLeafDim = 1 << kLeafLog2;                          // 8 voxels
constexpr int kLowerDim = kLeafDim << kLowerLog2;                 // 128 voxels
constexpr int kUpperDim = kLowerDim << kUpperLog2;                // 4096 voxels

struct Leaf {
    int x, y, z;                 // origin (voxels)
    uint64_t mask[8];            // active voxels, bit (x&7)<<6 | (y&7)<<3 | (z&7)
    float v[512];
};

void mask_set(uint64_t* m, int n) { m[n >> 6] |= 1ull << (n & 63); }

// io::writeCompressedValues with COMPRESS_ACTIVE_MASK: a metadata byte, then the
// active values only when every inactive value (outside child slots) is the
// background, else every value.
void write_values(Buf& o, const float* vals, int count, const uint64_t* vmask, const uint64_t* cmask,
                  float bg) {
    bool plain = true;
    for (int n = 0; n < count && plain; ++n) {
        const bool active = (vmask[n >> 6] >> (n & 63)) & 1ull;
        const bool child = cmask && ((cmask[n >> 6] >> (n & 63)) & 1ull);
        if (!active && !child && std::memcmp(&vals[n], &bg, 4) != 0) plain = false;
    }
    if (plain) {
        o.put<int8_t>(kNoMaskOrInactiveVals);
        for (int n = 0; n < count; ++n)
            if ((vmask[n >> 6] >> (n & 63)) & 1ull) o.put<float>(vals[n]);
    } else {
        o.put<int8_t>(kNoMaskAndAllVals);
        o.raw(vals, sizeof(float) * (size_t)count);
    }
}

void meta_str(Buf& o, const char* name, const std::string& v) {
    o.str(name);
    o.str("string");
    o.put<uint32_t>((uint32_t)v.size());
    o.raw(v.data(), v.size());
}
void meta_vec3i(Buf& o, const char* name, int a, int b, int c) {
    o.str(name);
    o.str("vec3i");
    o.put<uint32_t>(12);
    o.put<int32_t>(a);
    o.put<int32_t>(b);
    o.put<int32_t>(c);
}
void meta_i64(Buf& o, const char* name, int64_t v) {
    o.str(name);
    o.str("int64");
    o.put<uint32_t>(8);
    o.put<int64_t>(v);
}
void meta_bool(Buf& o, const char* name, bool v) {
    o.str(name);
    o.str("bool");
    o.put<uint32_t>(1);
    o.put<uint8_t>(v ? 1 : 0);
}

}  // namespace

// Dense node values phi[(i*M + j)*M + k] at index-space (i, j, k), 0 <= i,j,k < M
// <= 4096; nodes with |phi| < bg are active.
bool write_vdb(const char* path, const float* phi, int M, float bg, const char* grid_name,
               std::string& err) {
    if (M < 1 || M > kUpperDim) {
        err = "vdb: grid edge out of range";
        return false;
    }
    // leaves with at least one active node, in (x, y, z) order of origins: the
    // order in which the internal nodes list their children
    std::vector<Leaf> leaves;
    int64_t nactive = 0;
    int bmin[3] = {M, M, M}, bmax[3] = {-1, -1, -1};
    const int L = (M + kLeafDim - 1) / kLeafDim;
    for (int lx = 0; lx < L; ++lx)
        for (int ly = 0; ly < L; ++ly)
            for (int lz = 0; lz < L; ++lz) {
                Leaf lf;
                lf.x = lx * kLeafDim;
                lf.y = ly * kLeafDim;
                lf.z = lz * kLeafDim;
                std::memset(lf.mask, 0, sizeof(lf.mask));
                bool any = false;
                for (int a = 0; a < kLeafDim; ++a)
                    for (int b = 0; b < kLeafDim; ++b)
                        for (int c = 0; c < kLeafDim; ++c) {
                            const int n = (a << 6) | (b << 3) | c;
                            const int i = lf.x + a, j = lf.y + b, k = lf.z + c;
                            float v = bg;
                            if (i < M && j < M && k < M) v = phi[((int64_t)i * M + j) * M + k];
                            lf.v[n] = v;
                            // a leaf is kept for its active nodes, and for inactive
                            // nodes off the background (stored with all its values)
                            if (std::memcmp(&v, &bg, 4) != 0) any = true;
                            if (i < M && j < M && k < M && std::fabs(v) < bg) {
                                mask_set(lf.mask, n);
                                any = true;
                                ++nactive;
                                const int q[3] = {i, j, k};
                                for (int d = 0; d < 3; ++d) {
                                    bmin[d] = std::min(bmin[d], q[d]);
                                    bmax[d] = std::max(bmax[d], q[d]);
                                }
                            }
                        }
                if (any) leaves.push_back(lf);
            }
    // lower internal nodes (128^3) that hold a leaf, by origin
    std::vector<int> lower_of(leaves.size());
    std::vector<std::array<int, 3>> lowers;
    for (size_t n = 0; n < leaves.size(); ++n) {
        const std::array<int, 3> o = {leaves[n].x / kLowerDim * kLowerDim, leaves[n].y / kLowerDim * kLowerDim,
                                      leaves[n].z / kLowerDim * kLowerDim};
        auto it = std::find(lowers.begin(), lowers.end(), o);
        if (it == lowers.end()) {
            lowers.push_back(o);
            it = lowers.end() - 1;
        }
        lower_of[n] = (int)(it - lowers.begin());
    }
    // children are listed by their linear offset inside the parent (x-major)
    auto lower_rank = [](const std::array<int, 3>& o) {
        return ((o[0] / kLowerDim) << (2 * kUpperLog2)) | ((o[1] / kLowerDim) << kUpperLog2) | (o[2] / kLowerDim);
    };
    std::vector<int> lorder(lowers.size());
    for (size_t n = 0; n < lorder.size(); ++n) lorder[n] = (int)n;
    std::sort(lorder.begin(), lorder.end(), [&](int a, int b) { return lower_rank(lowers[a]) < lower_rank(lowers[b]); });
    auto leaf_rank = [](const Leaf& l) {
        const int x = l.x % kLowerDim / kLeafDim, y = l.y % kLowerDim / kLeafDim, z = l.z % kLowerDim / kLeafDim;
        return (x << (2 * kLowerLog2)) | (y << kLowerLog2) | z;
    };
    std::vector<std::vector<int>> kids(lowers.size());
    for (size_t n = 0; n < leaves.size(); ++n) kids[lower_of[n]].push_back((int)n);
    for (auto& k : kids)
        std::sort(k.begin(), k.end(), [&](int a, int b) { return leaf_rank(leaves[a]) < leaf_rank(leaves[b]); });

    Buf o;
    // archive header (io::Archive::writeHeader)
    o.put<int64_t>(kVdbMagic);
    o.put<uint32_t>(kVdbFileVersion);
    o.put<uint32_t>(3);   // library major
    o.put<uint32_t>(0);   // library minor
    o.put<uint8_t>(1);    // grid offsets present
    char uuid[40];
    std::snprintf(uuid, sizeof(uuid), "%08x-%04x-4%03x-8%03x-%012llx", (unsigned)M, (unsigned)(nactive & 0xffff),
                  (unsigned)(leaves.size() & 0xfff), (unsigned)(lowers.size() & 0xfff),
                  (unsigned long long)(nactive & 0xffffffffffffull));
    o.raw(uuid, 36);
    o.put<uint32_t>(0);   // file metadata: none
    o.put<int32_t>(1);    // one grid
    // grid descriptor (io::GridDescriptor::writeHeader, then its stream offsets)
    o.str(grid_name);
    o.str("Tree_float_5_4_3");
    o.str("");            // instance parent
    const size_t at_offsets = o.pos();
    o.put<int64_t>(0);
    o.put<int64_t>(0);
    o.put<int64_t>(0);
    const int64_t grid_pos = (int64_t)o.pos();
    o.put<uint32_t>(kCompressActiveMask);
    // grid metadata
    o.put<uint32_t>(7);
    meta_str(o, "class", "level set");
    meta_vec3i(o, "file_bbox_max", bmax[0], bmax[1], bmax[2]);
    meta_vec3i(o, "file_bbox_min", bmin[0], bmin[1], bmin[2]);
    meta_i64(o, "file_mem_bytes", (int64_t)(leaves.size() * sizeof(Leaf)));
    meta_i64(o, "file_voxel_count", nactive);
    meta_bool(o, "is_saved_as_half_float", false);
    meta_str(o, "name", grid_name);
    // transform (math::Transform::write): UniformScaleMap, voxel size 1
    o.str("UniformScaleMap");
    const double sv[5] = {1.0, 1.0, 1.0, 1.0, 0.5};   // scale, voxel size, 1/scale, 1/scale^2, 1/(2 scale)
    for (double s : sv)
        for (int d = 0; d < 3; ++d) o.put<double>(s);
    // topology (TreeBase::writeTopology, RootNode / InternalNode / LeafNode)
    o.put<int32_t>(1);   // buffer count
    o.put<float>(bg);
    o.put<uint32_t>(0);                                   // root tiles
    o.put<uint32_t>(leaves.empty() ? 0u : 1u);            // root children
    std::vector<float> zeros;
    if (!leaves.empty()) {
        o.put<int32_t>(0);
        o.put<int32_t>(0);
        o.put<int32_t>(0);
        // upper internal node at the origin: 32^3 slots
        constexpr int kUpperN = 1 << (3 * kUpperLog2);
        std::vector<uint64_t> cmask(kUpperN / 64, 0), vmask(kUpperN / 64, 0);
        for (int l : lorder) mask_set(cmask.data(), lower_rank(lowers[l]));
        std::vector<float> vals(kUpperN, bg);
        for (int l : lorder) vals[lower_rank(lowers[l])] = 0.0f;   // child slots hold zero
        o.raw(cmask.data(), cmask.size() * 8);
        o.raw(vmask.data(), vmask.size() * 8);
        write_values(o, vals.data(), kUpperN, vmask.data(), cmask.data(), bg);
        for (int l : lorder) {
            constexpr int kLowerN = 1 << (3 * kLowerLog2);
            std::vector<uint64_t> lc(kLowerN / 64, 0), lv(kLowerN / 64, 0);
            std::vector<float> lvals(kLowerN, bg);
            for (int n : kids[l]) {
                mask_set(lc.data(), leaf_rank(leaves[n]));
                lvals[leaf_rank(leaves[n])] = 0.0f;
            }
            o.raw(lc.data(), lc.size() * 8);
            o.raw(lv.data(), lv.size() * 8);
            write_values(o, lvals.data(), kLowerN, lv.data(), lc.data(), bg);
            for (int n : kids[l]) o.raw(leaves[n].mask, sizeof(leaves[n].mask));
        }
    }
    const int64_t block_pos = (int64_t)o.pos();
    // buffers (leaf values, in the same order)
    for (int l : lorder)
        for (int n : kids[l]) {
            o.raw(leaves[n].mask, sizeof(leaves[n].mask));
            write_values(o, leaves[n].v, 512, leaves[n].mask, nullptr, bg);
        }
    const int64_t end_pos = (int64_t)o.pos();
    o.patch<int64_t>(at_offsets, grid_pos);
    o.patch<int64_t>(at_offsets + 8, block_pos);
    o.patch<int64_t>(at_offsets + 16, end_pos);
    Sink s(path);
    if (!s.open()) {
        err = std::string("cannot open ") + path;
        return false;
    }
    s.put(o.b.data(), o.b.size());
    if (!s.close()) {
        err = std::string("write failed: ") + path;
        return false;
    }
    return true;
}

// ----------------------------------------------------------------- Partio
namespace {

// ASCII classic geometry (the body of Partio's GEO writer)
bool geo_text(Sink& s, int64_t n, const float* pos, const float* vel) {
    char head[512];
    const int hl = std::snprintf(head, sizeof(head),
                                 "PGEOMETRY V5\n"
                                 "NPoints %lld NPrims 0\n"
                                 "NPointGroups 0 NPrimGroups 0\n"
                                 "NPointAttrib 2 NVertexAttrib 0 NPrimAttrib 0 NAttrib 0\n"
                                 "PointAttrib\n"
                                 "v 3 vector 0 0 0\n"
                                 "id 1 int 0\n",
                                 (long long)n);
    s.put(head, (size_t)hl);
    // the point lines are formatted by several threads, a block of points each,
    // and written in order
    const int64_t kBlockPoints = 1 << 16;
    const int64_t nblocks = (n + kBlockPoints - 1) / kBlockPoints;
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int64_t b0 = 0; b0 < nblocks; b0 += nthreads) {
        const int64_t b1 = std::min<int64_t>(nblocks, b0 + nthreads);
        std::vector<std::string> text((size_t)(b1 - b0));
        std::vector<std::thread> pool;
        for (int64_t b = b0; b < b1; ++b)
            pool.emplace_back([&, b] {
                std::string& t = text[(size_t)(b - b0)];
                char line[256];
                for (int64_t p = b * kBlockPoints; p < std::min(n, (b + 1) * kBlockPoints); ++p) {
                    const int len = std::snprintf(line, sizeof(line), "%.9g %.9g %.9g 1 (%.9g %.9g %.9g %lld)\n",
                                                  pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], vel[3 * p],
                                                  vel[3 * p + 1], vel[3 * p + 2], (long long)p);
                    t.append(line, (size_t)len);
                }
            });
        for (auto& th : pool) th.join();
        for (const std::string& t : text) s.put(t.data(), t.size());
    }
    static const char tail[] = "beginExtra\nendExtra\n";
    s.put(tail, sizeof(tail) - 1);
    return true;
}

void be32(unsigned char* d, uint32_t v) {
    d[0] = (unsigned char)(v >> 24);
    d[1] = (unsigned char)(v >> 16);
    d[2] = (unsigned char)(v >> 8);
    d[3] = (unsigned char)v;
}

// Binary classic geometry (Partio's BGEO writer): big-endian. Header: magic
// "Bgeo", 'V', version 5, point/prim/group/attribute counts; each point
// attribute as a 16-bit name length, the name, a 16-bit size, the Houdini type
// (0 float, 1 int, 5 vector) and size default values; then per point x, y, z, w
// = 1 and the attribute values; then the beginExtra / endExtra bytes 0x00 0xff.
void bgeo_binary(Sink& s, int64_t n, const float* pos, const float* vel) {
    std::vector<unsigned char> h;
    auto u32 = [&](uint32_t v) {
        unsigned char d[4];
        be32(d, v);
        h.insert(h.end(), d, d + 4);
    };
    auto u16 = [&](uint16_t v) {
        h.push_back((unsigned char)(v >> 8));
        h.push_back((unsigned char)v);
    };
    u32(((((uint32_t)'B' << 8 | 'g') << 8 | 'e') << 8) | 'o');
    h.push_back('V');
    u32(5);
    u32((uint32_t)n);
    u32(0);   // prims
    u32(0);   // point groups
    u32(0);   // prim groups
    u32(2);   // point attributes (v, id)
    u32(0);   // vertex attributes
    u32(0);   // prim attributes
    u32(0);   // detail attributes
    const char* names[2] = {"v", "id"};
    const uint16_t sizes[2] = {3, 1};
    const uint32_t types[2] = {5, 1};
    for (int a = 0; a < 2; ++a) {
        u16((uint16_t)std::strlen(names[a]));
        h.insert(h.end(), names[a], names[a] + std::strlen(names[a]));
        u16(sizes[a]);
        u32(types[a]);
        for (int c = 0; c < sizes[a]; ++c) u32(0);
    }
    s.put(h.data(), h.size());
    // 32 bytes per point: position, w, velocity, id
    const int64_t kChunk = 1 << 16;
    std::vector<unsigned char> body((size_t)std::min<int64_t>(n, kChunk) * 32);
    for (int64_t p0 = 0; p0 < n; p0 += kChunk) {
        const int64_t p1 = std::min(n, p0 + kChunk);
        for (int64_t p = p0; p < p1; ++p) {
            unsigned char* d = &body[(size_t)(p - p0) * 32];
            const float w = 1.0f;
            const float f[7] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], w, vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
            for (int c = 0; c < 7; ++c) {
                uint32_t u;
                std::memcpy(&u, &f[c], 4);
                be32(d + 4 * c, u);
            }
            be32(d + 28, (uint32_t)(int32_t)p);
        }
        s.put(body.data(), (size_t)(p1 - p0) * 32);
    }
    const unsigned char extra[2] = {0x00, 0xff};
    s.put(extra, 2);
}

}  // namespace

// Partio::write's dispatch on the extension (".gz" appended means gzipped):
// 1 ASCII, 2 binary, 0 not a format built here.
int partio_format(const char* path) {
    const std::string p(path);
    const std::string base = Sink::ends_with(p, ".gz") ? p.substr(0, p.size() - 3) : p;
    return Sink::ends_with(base, ".geo") ? 1 : Sink::ends_with(base, ".bgeo") ? 2 : 0;
}

int write_partio(const char* path, int64_t n, const float* pos, const float* vel, std::string& err) {
    const std::string p(path);
    const int fmt = partio_format(path);
    const bool geo = fmt == 1;
    if (fmt == 0) {
        err = "particle export: the extension must be .geo or .bgeo (optionally .gz)";
        return ARIEL_ERR_UNSUPPORTED;
    }
    Sink s(p);
    if (!s.open()) {
        err = std::string("cannot open ") + path;
        return ARIEL_ERR_IO;
    }
    if (geo) geo_text(s, n, pos, vel);
    else bgeo_binary(s, n, pos, vel);
    if (!s.close()) {
        err = std::string("write failed: ") + path;
        return ARIEL_ERR_IO;
    }
    return ARIEL_OK;
}

}  // namespace ariel
