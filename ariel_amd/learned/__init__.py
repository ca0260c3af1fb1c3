"""Learned preconditioner for the pressure solve: weight blob format and location.

The network (a 4-level convolutional U-Net, BF16 weights) runs inside the
library's flexible PCG; this package holds its packed weights
(``weights.bin``) and the blob reader/writer. ``trainer`` drives the library's
own training kernels, ``train`` is the training script, and ``net`` is the
PyTorch restatement of the network that tests check the kernels against
(torch is needed only for those three).

Blob layout: ``b"ARIELNN1"``, int32 layer count (11), int32 zero, int64 value
count, then that many BF16 values (little endian) in the layer order and
layouts listed in LAYERS.
"""
import os
import struct

import numpy as np

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights.bin")
MAGIC = b"ARIELNN1"

# (name, taps, stride, cin, cout, relu); kernel layout per layer:
#   "in":     [16][64]         k = 2*tap + channel, zero for k >= 54
#   3x3x3:    [27][cout][cin]  tap = 9*(di+1) + 3*(dj+1) + (dk+1)
#   1x1x1:    [cout][cin]
#   "out":    [27][16][16]     only output column 0 is used
LAYERS = (
    ("in", 27, 1, 2, 16, True),
    ("e1", 27, 2, 16, 32, True),
    ("e2", 27, 2, 32, 48, True),
    ("e3", 27, 2, 48, 64, True),
    ("p2", 1, 1, 64, 48, False),
    ("d2", 27, 1, 48, 48, True),
    ("p1", 1, 1, 48, 32, False),
    ("d1", 27, 1, 32, 32, True),
    ("p0", 1, 1, 32, 16, False),
    ("d0", 27, 1, 16, 16, True),
    ("out", 27, 1, 16, 1, False),
)


def packed_size(name, taps, cin, cout):
    if name == "in":
        return 16 * 64
    if name == "out":
        return 27 * 16 * 16
    return taps * cout * cin


def layout():
    """Offsets (BF16 elements) of each layer in the blob and the total count."""
    off, acc = [], 0
    for name, taps, _s, cin, cout, _r in LAYERS:
        off.append(acc)
        acc += packed_size(name, taps, cin, cout)
    return off, acc


def f32_to_bf16_bits(a):
    """Round-to-nearest-even FP32 -> BF16 bit patterns (finite inputs)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_bits_to_f32(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def layers_to_flat(layers):
    """11 float32 arrays in torch conv layout ([cout][cin][k][k][k]) -> the packed
    FP32 vector (the blob's layout, no rounding)."""
    if len(layers) != len(LAYERS):
        raise ValueError("expected %d layers" % len(LAYERS))
    parts = []
    for (name, taps, _s, cin, cout, _r), w in zip(LAYERS, layers):
        w = np.asarray(w, np.float32)
        k = 3 if taps == 27 else 1
        if w.shape != (cout, cin, k, k, k):
            raise ValueError("layer %s: shape %s" % (name, w.shape))
        if name == "in":
            # [co][c][tap] -> [co][2*tap + c], padded to 64
            t = w.reshape(cout, cin, 27).transpose(0, 2, 1).reshape(cout, 54)
            m = np.zeros((16, 64), np.float32)
            m[:, :54] = t
        elif name == "out":
            m = np.zeros((27, 16, 16), np.float32)
            m[:, 0, :] = w.reshape(1, cin, 27)[0].T
        elif taps == 27:
            m = w.reshape(cout, cin, 27).transpose(2, 0, 1)
        else:
            m = w.reshape(cout, cin)
        parts.append(np.ascontiguousarray(m, np.float32).reshape(-1))
    flat = np.concatenate(parts)
    assert flat.size == layout()[1]
    return flat


def flat_to_layers(flat):
    """The packed vector -> 11 float32 arrays in torch conv layout."""
    flat = np.asarray(flat, np.float32)
    off, want = layout()
    if flat.size != want:
        raise ValueError("packed weights: %d values, want %d" % (flat.size, want))
    out = []
    for (name, taps, _s, cin, cout, _r), o in zip(LAYERS, off):
        m = flat[o:o + packed_size(name, taps, cin, cout)]
        if name == "in":
            w = m.reshape(16, 64)[:, :54].reshape(cout, 27, cin).transpose(0, 2, 1)
            w = w.reshape(cout, cin, 3, 3, 3)
        elif name == "out":
            w = m.reshape(27, 16, 16)[:, 0, :].T.reshape(1, cin, 3, 3, 3)
        elif taps == 27:
            w = m.reshape(27, cout, cin).transpose(1, 2, 0).reshape(cout, cin, 3, 3, 3)
        else:
            w = m.reshape(cout, cin, 1, 1, 1)
        out.append(np.ascontiguousarray(w, np.float32))
    return out


def pack_flat(flat):
    """Packed FP32 vector -> blob bytes (BF16, round to nearest even)."""
    flat = np.asarray(flat, np.float32)
    _off, total = layout()
    if flat.size != total:
        raise ValueError("packed weights: %d values, want %d" % (flat.size, total))
    vals = f32_to_bf16_bits(flat)
    return MAGIC + struct.pack("<iiq", len(LAYERS), 0, total) + vals.astype("<u2").tobytes()


def pack(layers):
    """layers: list of 11 float32 arrays in torch conv layout [cout][cin][3][3][3]
    (or [cout][cin][1][1][1]) in LAYERS order -> blob bytes."""
    return pack_flat(layers_to_flat(layers))


def unpack_flat(blob):
    """Blob bytes -> the packed vector of its BF16 values as float32."""
    if blob[:8] != MAGIC:
        raise ValueError("bad magic")
    nl, _z, total = struct.unpack("<iiq", blob[8:24])
    _off, want = layout()
    if nl != len(LAYERS) or total != want or len(blob) != 24 + 2 * total:
        raise ValueError("layout mismatch")
    return bf16_bits_to_f32(np.frombuffer(blob, "<u2", total, 24))


def unpack(blob):
    """Blob bytes -> list of float32 arrays in torch conv layout (BF16 values)."""
    return flat_to_layers(unpack_flat(blob))
