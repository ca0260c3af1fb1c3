"""Readers of the files the library writes: the OpenVDB level-set file of
ariel_export_levelset (".vdb") and the Partio particle files of
ariel_export_particles (".geo", ".bgeo", either gzipped as ".gz").

The VDB reader follows the OpenVDB file layout (format versions 222-224, the
Tree_float_5_4_3 FloatGrid, per-node active-mask compression without zip) far
enough to read back such files into a dense array; it is no general OpenVDB
reader.
"""
import gzip
import struct

import numpy as np

VDB_MAGIC = 0x56444220
_LEAF_LOG2, _LOWER_LOG2, _UPPER_LOG2 = 3, 4, 5


class _Reader:
    def __init__(self, data):
        self.d = data
        self.p = 0

    def take(self, n):
        b = self.d[self.p:self.p + n]
        if len(b) != n:
            raise ValueError("truncated file")
        self.p += n
        return b

    def unpack(self, fmt):
        return struct.unpack("<" + fmt, self.take(struct.calcsize("<" + fmt)))

    def u32(self):
        return self.unpack("I")[0]

    def string(self):
        return self.take(self.u32()).decode()


def _mask(r, nbits):
    words = np.frombuffer(r.take(nbits // 8), dtype="<u8")
    return np.unpackbits(words.view(np.uint8), bitorder="little").astype(bool)


def _values(r, count, vmask, bg, cmask=None):
    """io::readCompressedValues for an uncompressed (no zip) stream with the
    active-mask flag: metadata byte, optional inactive values, then values."""
    meta = r.unpack("b")[0]
    inactive = np.full(count, bg, np.float32)
    if meta == 6:   # every value
        return np.frombuffer(r.take(4 * count), dtype="<f4").copy()
    if meta in (1,):
        inactive[:] = -bg
    elif meta in (2, 4):
        inactive[:] = r.unpack("f")[0]
    elif meta == 5:
        a, b = r.unpack("ff")
        sel = _mask(r, count)
        inactive[:] = np.where(sel, b, a)
    elif meta == 3:
        sel = _mask(r, count)
        inactive[:] = np.where(sel, -bg, bg)
    elif meta != 0:
        raise ValueError(f"unknown value metadata {meta}")
    n_on = int(vmask.sum())
    act = np.frombuffer(r.take(4 * n_on), dtype="<f4")
    out = inactive
    out[vmask] = act
    return out


def read_vdb(path, shape=None):
    """Returns (phi, background, active, meta): phi and active dense over [0, M)^3
    (M from `shape` or the grid's bounding box), x-major; inactive nodes hold
    their stored inactive values (the background for tiles)."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    magic, version = r.unpack("qI")
    if magic != VDB_MAGIC:
        raise ValueError("not a VDB file")
    r.unpack("II")          # library version
    r.unpack("B")           # grid offsets flag
    r.take(36)              # uuid
    for _ in range(r.u32()):   # file metadata
        r.string(), r.string()
        r.take(r.u32())
    ngrids = r.unpack("i")[0]
    if ngrids != 1:
        raise ValueError("expected one grid")
    name, gtype, _parent = r.string(), r.string(), r.string()
    if gtype != "Tree_float_5_4_3":
        raise ValueError(f"grid type {gtype}")
    grid_pos, block_pos, end_pos = r.unpack("qqq")
    assert r.p == grid_pos, (r.p, grid_pos)
    compression = r.u32()
    if compression & 0x5:
        raise ValueError("zip/blosc compressed grids are not read here")
    meta = {}
    for _ in range(r.u32()):
        key, typ = r.string(), r.string()
        raw = r.take(r.u32())
        meta[key] = {"string": lambda b: b.decode(), "vec3i": lambda b: struct.unpack("<3i", b),
                     "int64": lambda b: struct.unpack("<q", b)[0],
                     "bool": lambda b: bool(b[0])}.get(typ, lambda b: b)(raw)
    xform = r.string()
    xvals = r.unpack("15d")
    buffers = r.unpack("i")[0]
    bg = r.unpack("f")[0]
    ntiles, nchildren = r.unpack("II")
    if ntiles:
        raise ValueError("root tiles are not read here")
    if shape is None:
        hi = meta.get("file_bbox_max", (-1, -1, -1))
        shape = max(hi) + 1
    M = int(shape)
    phi = np.full((M, M, M), bg, np.float32)
    active = np.zeros((M, M, M), bool)
    leaves = []   # (origin, mask) in topology order
    upper_n = 1 << (3 * _UPPER_LOG2)
    lower_n = 1 << (3 * _LOWER_LOG2)
    for _ in range(nchildren):
        ox, oy, oz = r.unpack("3i")
        cm = _mask(r, upper_n)
        vm = _mask(r, upper_n)
        _values(r, upper_n, vm, bg, cm)
        for u in np.nonzero(cm)[0]:
            ux, uy, uz = u >> 10, (u >> 5) & 31, u & 31
            lo = (ox + (ux << 7), oy + (uy << 7), oz + (uz << 7))
            lcm = _mask(r, lower_n)
            lvm = _mask(r, lower_n)
            _values(r, lower_n, lvm, bg, lcm)
            for l in np.nonzero(lcm)[0]:
                lx, ly, lz = l >> 8, (l >> 4) & 15, l & 15
                origin = (lo[0] + (lx << 3), lo[1] + (ly << 3), lo[2] + (lz << 3))
                leaves.append((origin, _mask(r, 512)))
    assert r.p == block_pos, (r.p, block_pos)
    for origin, tmask in leaves:
        vm = _mask(r, 512)
        if not np.array_equal(vm, tmask):
            raise ValueError("leaf mask differs between topology and buffers")
        vals = _values(r, 512, vm, bg).reshape(8, 8, 8)
        am = vm.reshape(8, 8, 8)
        x, y, z = origin
        sx, sy, sz = min(8, M - x), min(8, M - y), min(8, M - z)
        if sx > 0 and sy > 0 and sz > 0:
            phi[x:x + sx, y:y + sy, z:z + sz] = vals[:sx, :sy, :sz]
            active[x:x + sx, y:y + sy, z:z + sz] = am[:sx, :sy, :sz]
    assert r.p == end_pos, (r.p, end_pos)
    meta.update({"name": name, "version": version, "transform": xform, "transform_values": xvals,
                 "buffers": buffers, "leaves": len(leaves), "compression": compression})
    return phi, bg, active, meta


def _open(path):
    path = str(path)
    return gzip.open(path, "rb") if path.endswith(".gz") else open(path, "rb")


def read_partio(path):
    """Returns (pos, vel, id) of a ".geo"/".bgeo" file (optionally ".gz") as the
    library writes them: position (grid units), "v" and "id"."""
    p = str(path)
    base = p[:-3] if p.endswith(".gz") else p
    with _open(p) as f:
        data = f.read()
    if base.endswith(".bgeo"):
        r = _Reader(data)
        head = struct.unpack(">IcI", r.take(9))
        if head[0] != 0x4267656F or head[1] != b"V" or head[2] != 5:
            raise ValueError("not a classic bgeo file")
        npts, nprim, npg, nprg, npa, nva, nprima, nattr = struct.unpack(">8I", r.take(32))
        attrs = []
        for _ in range(npa):
            nlen = struct.unpack(">H", r.take(2))[0]
            name = r.take(nlen).decode()
            size = struct.unpack(">H", r.take(2))[0]
            typ = struct.unpack(">I", r.take(4))[0]
            r.take(4 * size)
            attrs.append((name, size, typ))
        width = 4 + sum(a[1] for a in attrs)
        body = np.frombuffer(r.take(4 * width * npts), dtype=">u4").reshape(npts, width)
        extra = r.take(2)
        if extra != b"\x00\xff" or r.p != len(data):
            raise ValueError("bad bgeo trailer")
        f32 = body.view(">f4")
        pos = f32[:, :3].astype(np.float32)
        col = 4
        out = {}
        for name, size, typ in attrs:
            blk = body[:, col:col + size]
            out[name] = blk.view(">i4").astype(np.int64) if typ == 1 else blk.view(">f4").astype(np.float32)
            col += size
        return pos, out["v"], out["id"][:, 0]
    lines = data.decode().splitlines()
    if lines[0] != "PGEOMETRY V5":
        raise ValueError("not a classic geo file")
    n = int(lines[1].split()[1])
    body = lines[7:7 + n]
    vals = np.array([ln.replace("(", " ").replace(")", " ").split() for ln in body], np.float64).reshape(n, 8)
    return vals[:, 0:3].astype(np.float32), vals[:, 4:7].astype(np.float32), vals[:, 7].astype(np.int64)
