"""TEST INFRASTRUCTURE ONLY: CPU restatement of the particle-file format, the
checker for the library's scenario readers and writers.

Loaded by tests/ only, as the checker and never as the product path.

Restated from:
  - the xxHash XXH32 specification (xxh32 below; pinned in the tests against
    the xxhash package and published known answers);
  - the lz4 frame format and block format specifications (lz4 library,
    vcpkg port at builtin-baseline 8ff168270bc5977384b008bec6e879e15089d850,
    which the reference links through LZ4F_decompress / LZ4F_compressUpdate in
    src/libFluid/serialization/helpers/Lz4CompressedStream.cpp:7-187);
  - ParticleSerializer::deserialize (src/libFluid/serialization/ParticleSerializer.cpp:91-141)
    for the record layout: u32 version, u32 component count, u32 ids,
    u64 particle count, then the components of each particle in id order.
Pure Python and numpy, written for clarity rather than speed.
"""
import struct

import numpy as np

P1, P2, P3, P4, P5 = 2654435761, 2246822519, 3266489917, 668265263, 374761393
M32 = 0xFFFFFFFF


def _rotl(x, r):
    return ((x << r) | (x >> (32 - r))) & M32


def xxh32(data, seed=0):
    data = bytes(data)
    n = len(data)
    i = 0
    if n >= 16:
        v = [(seed + P1 + P2) & M32, (seed + P2) & M32, seed & M32, (seed - P1) & M32]
        while i + 16 <= n:
            for k in range(4):
                lane = struct.unpack_from("<I", data, i + 4 * k)[0]
                v[k] = (_rotl((v[k] + lane * P2) & M32, 13) * P1) & M32
            i += 16
        h = (_rotl(v[0], 1) + _rotl(v[1], 7) + _rotl(v[2], 12) + _rotl(v[3], 18)) & M32
    else:
        h = (seed + P5) & M32
    h = (h + n) & M32
    while i + 4 <= n:
        h = (_rotl((h + struct.unpack_from("<I", data, i)[0] * P3) & M32, 17) * P4) & M32
        i += 4
    while i < n:
        h = (_rotl((h + data[i] * P5) & M32, 11) * P1) & M32
        i += 1
    h ^= h >> 15
    h = (h * P2) & M32
    h ^= h >> 13
    h = (h * P3) & M32
    h ^= h >> 16
    return h


class FormatError(ValueError):
    pass


def lz4_block_decode(block, out, window_start):
    """Appends the decoded LZ4 block to out (a bytearray); matches may reach
    back to out[window_start]."""
    i, n = 0, len(block)
    while True:
        if i >= n:
            raise FormatError("missing token")
        token = block[i]
        i += 1
        lit = token >> 4
        if lit == 15:
            while True:
                b = block[i]
                i += 1
                lit += b
                if b != 255:
                    break
        if i + lit > n:
            raise FormatError("literals past the block")
        out += block[i : i + lit]
        i += lit
        if i == n:
            return
        off = block[i] | (block[i + 1] << 8)
        i += 2
        if off == 0 or off > len(out) - window_start:
            raise FormatError("offset out of range")
        ml = token & 15
        if ml == 15:
            while True:
                b = block[i]
                i += 1
                ml += b
                if b != 255:
                    break
        ml += 4
        start = len(out) - off
        for k in range(ml):
            out.append(out[start + k])


def lz4_frame_decode(data):
    data = bytes(data)
    out = bytearray()
    pos = 0
    while pos < len(data):
        magic = struct.unpack_from("<I", data, pos)[0]
        pos += 4
        if magic & 0xFFFFFFF0 == 0x184D2A50:
            size = struct.unpack_from("<I", data, pos)[0]
            pos += 4 + size
            continue
        if magic != 0x184D2204:
            raise FormatError("bad magic")
        flg, bd = data[pos], data[pos + 1]
        if flg >> 6 != 1:
            raise FormatError("version")
        indep, bchk, csize, cchk, dictid = (flg >> 5) & 1, (flg >> 4) & 1, (flg >> 3) & 1, (flg >> 2) & 1, flg & 1
        dlen = 2 + 8 * csize + 4 * dictid
        if (xxh32(data[pos : pos + dlen]) >> 8) & 0xFF != data[pos + dlen]:
            raise FormatError("header checksum")
        if dictid:
            raise FormatError("dictionary frames")
        block_max = 1 << (8 + 2 * ((bd >> 4) & 7))
        content_size = struct.unpack_from("<Q", data, pos + 2)[0] if csize else None
        pos += dlen + 1
        frame_start = len(out)
        while True:
            word = struct.unpack_from("<I", data, pos)[0]
            pos += 4
            if word == 0:
                break
            size = word & 0x7FFFFFFF
            if size > block_max:
                raise FormatError("block too large")
            block = data[pos : pos + size]
            if bchk and xxh32(block) != struct.unpack_from("<I", data, pos + size)[0]:
                raise FormatError("block checksum")
            if word & 0x80000000:
                out += block
            else:
                lz4_block_decode(block, out, len(out) if indep else frame_start)
            pos += size + 4 * bchk
        if cchk:
            if xxh32(out[frame_start:]) != struct.unpack_from("<I", data, pos)[0]:
                raise FormatError("content checksum")
            pos += 4
        if content_size is not None and content_size != len(out) - frame_start:
            raise FormatError("content size")
    return bytes(out)


COMPONENT_FIELDS = {
    1: [("position", "<f4", (2,)), ("velocity", "<f4", (2,)), ("acceleration", "<f4", (2,))],
    2: [("position", "<f4", (3,)), ("velocity", "<f4", (3,)), ("acceleration", "<f4", (3,))],
    3: [("type", "u1"), ("tag", "<u4")],
    4: [("density", "<f4"), ("mass", "<f4"), ("pressure", "<f4")],
    5: [("external", "<f4", (2,))],
    6: [("external", "<f4", (3,))],
}


def parse_particle_file(content):
    """(version, component ids, records) of decoded particle-file content;
    records is a packed numpy structured array in file layout."""
    version, nc = struct.unpack_from("<II", content, 0)
    ids = list(struct.unpack_from(f"<{nc}I", content, 8))
    (count,) = struct.unpack_from("<Q", content, 8 + 4 * nc)
    fields = []
    for cid in sorted(set(ids)):
        fields += COMPONENT_FIELDS[cid]
    dt = np.dtype(fields)  # packed: numpy structured dtypes carry no padding unless asked
    start = 16 + 4 * nc
    records = np.frombuffer(content, dtype=dt, count=count, offset=start)
    return version, ids, records
