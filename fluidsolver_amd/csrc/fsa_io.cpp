// fsa_io.cpp — scenario files on the host: the LZ4-framed particle .data format
// and the JSON version-2 scenario.
//
// Reference: ParticleSerializer.cpp:41-141 (records), Lz4CompressedStream.cpp:7-187
// (one LZ4 frame over the whole file, written with the lz4 library's frame API
// and default preferences), EndianSafeBinaryStream.hpp:29-30 (little endian),
// MainSerializer.cpp:114-223, ScenarioSerializer.cpp:43-79,
// SolverSerializer.cpp:137-259, TimestepGeneratorSerializer.cpp:47-72.
//
// The LZ4 frame and block formats and xxHash32 are restated here from their
// published specifications (lz4 "Frame format" and "Block format" documents,
// xxHash "XXH32" description); no lz4 library is linked.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "fsa_host.hpp"

namespace fsa {
namespace io {

// ---- byte helpers (little endian) ---------------------------------------------------
inline uint32_t rd32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
inline uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }
inline void wr32(std::vector<uint8_t>& v, uint32_t x) {
    for (int k = 0; k < 4; k++) v.push_back((uint8_t)(x >> (8 * k)));
}
inline void wr64(std::vector<uint8_t>& v, uint64_t x) {
    wr32(v, (uint32_t)x);
    wr32(v, (uint32_t)(x >> 32));
}
inline void wrf(std::vector<uint8_t>& v, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    wr32(v, u);
}
inline float rdf(const uint8_t* p) {
    const uint32_t u = rd32(p);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

[[noreturn]] void io_error(const std::string& m) { throw FsaError(FSA_E_IO, m); }

// ---- xxHash32 -------------------------------------------------------------------------
constexpr uint32_t kP1 = 2654435761u, kP2 = 2246822519u, kP3 = 3266489917u, kP4 = 668265263u, kP5 = 374761393u;
inline uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
inline uint32_t xround(uint32_t acc, uint32_t in) { return rotl(acc + in * kP2, 13) * kP1; }

uint32_t xxh32(const uint8_t* p, size_t n, uint32_t seed) {
    const uint8_t* const end = p + n;
    uint32_t h;
    if (n >= 16) {
        uint32_t v1 = seed + kP1 + kP2, v2 = seed + kP2, v3 = seed, v4 = seed - kP1;
        const uint8_t* const limit = end - 16;
        do {
            v1 = xround(v1, rd32(p));
            v2 = xround(v2, rd32(p + 4));
            v3 = xround(v3, rd32(p + 8));
            v4 = xround(v4, rd32(p + 12));
            p += 16;
        } while (p <= limit);
        h = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
    } else {
        h = seed + kP5;
    }
    h += (uint32_t)n;
    while (p + 4 <= end) {
        h = rotl(h + rd32(p) * kP3, 17) * kP4;
        p += 4;
    }
    while (p < end) {
        h = rotl(h + (*p) * kP5, 11) * kP1;
        p++;
    }
    h ^= h >> 15;
    h *= kP2;
    h ^= h >> 13;
    h *= kP3;
    h ^= h >> 16;
    return h;
}

// ---- LZ4 block format -------------------------------------------------------------------
// Sequences of token (literal length << 4 | match length - 4), literal length
// extension bytes (255 continues), literals, a 16-bit little-endian offset,
// match length extension bytes.  The last sequence carries literals only.
// Decodes src into out (appending); matches may reach back to out[window_start].
void lz4_block_decode(const uint8_t* src, size_t n, std::vector<uint8_t>& out, size_t window_start,
                      size_t max_out) {
    const uint8_t* ip = src;
    const uint8_t* const iend = src + n;
    const size_t out_start = out.size();
    auto length_ext = [&](size_t len) {
        uint8_t b;
        do {
            if (ip >= iend) io_error("LZ4 block: truncated length");
            b = *ip++;
            len += b;
        } while (b == 255);
        return len;
    };
    if (n == 0) io_error("LZ4 block: empty compressed block");
    while (true) {
        if (ip >= iend) io_error("LZ4 block: missing token");
        const uint8_t token = *ip++;
        size_t lit = token >> 4;
        if (lit == 15) lit = length_ext(lit);
        if ((size_t)(iend - ip) < lit) io_error("LZ4 block: literals run past the block");
        if (out.size() - out_start + lit > max_out) io_error("LZ4 block: output exceeds the block size");
        out.insert(out.end(), ip, ip + lit);
        ip += lit;
        if (ip == iend) break;  // last sequence
        if (iend - ip < 2) io_error("LZ4 block: truncated offset");
        const size_t offset = (size_t)ip[0] | ((size_t)ip[1] << 8);
        ip += 2;
        if (offset == 0 || offset > out.size() - window_start) io_error("LZ4 block: match offset out of range");
        size_t ml = token & 15u;
        if (ml == 15) ml = length_ext(ml);
        ml += 4;
        if (out.size() - out_start + ml > max_out) io_error("LZ4 block: output exceeds the block size");
        const size_t from = out.size() - offset;
        out.reserve(out.size() + ml);
        for (size_t k = 0; k < ml; k++) {
            const uint8_t byte = out[from + k];  // overlapping copies repeat the last offset bytes
            out.push_back(byte);
        }
    }
}

// Greedy LZ4 block compressor (hash of 4-byte sequences, one candidate per
// slot).  Keeps the format's end rules: a match starts at least 12 bytes and
// ends at least 5 bytes before the end of the block.  Returns the compressed
// size, or 0 when the block does not shrink.
size_t lz4_block_encode(const uint8_t* src, size_t n, std::vector<uint8_t>& dst) {
    constexpr int kHashBits = 12;
    constexpr size_t kMfLimit = 12, kLastLiterals = 5;
    dst.clear();
    std::vector<int32_t> table((size_t)1 << kHashBits, -1);
    size_t anchor = 0, ip = 0;
    auto emit = [&](size_t lit_end, size_t offset, size_t ml) {
        const size_t lit = lit_end - anchor;
        const size_t mlc = ml >= 4 ? ml - 4 : 0;
        uint8_t token = (uint8_t)(std::min<size_t>(lit, 15) << 4);
        if (ml) token |= (uint8_t)std::min<size_t>(mlc, 15);
        dst.push_back(token);
        if (lit >= 15) {
            size_t r = lit - 15;
            for (; r >= 255; r -= 255) dst.push_back(255);
            dst.push_back((uint8_t)r);
        }
        dst.insert(dst.end(), src + anchor, src + lit_end);
        if (ml) {
            dst.push_back((uint8_t)(offset & 0xff));
            dst.push_back((uint8_t)(offset >> 8));
            if (mlc >= 15) {
                size_t r = mlc - 15;
                for (; r >= 255; r -= 255) dst.push_back(255);
                dst.push_back((uint8_t)r);
            }
        }
    };
    if (n > kMfLimit) {
        const size_t limit = n - kMfLimit, match_end_limit = n - kLastLiterals;
        while (ip < limit) {
            const uint32_t seq = rd32(src + ip);
            const uint32_t h = (seq * kP1) >> (32 - kHashBits);
            const int32_t ref = table[h];
            table[h] = (int32_t)ip;
            if (ref >= 0 && ip - (size_t)ref <= 65535 && rd32(src + ref) == seq) {
                size_t ml = 4;
                while (ip + ml < match_end_limit && src[ref + ml] == src[ip + ml]) ml++;
                emit(ip, ip - (size_t)ref, ml);
                ip += ml;
                anchor = ip;
            } else {
                ip++;
            }
        }
    }
    emit(n, 0, 0);  // last literals
    return dst.size() < n ? dst.size() : 0;
}

// ---- LZ4 frame format --------------------------------------------------------------------
constexpr uint32_t kFrameMagic = 0x184D2204u;
constexpr uint32_t kSkippableMask = 0xFFFFFFF0u, kSkippableMagic = 0x184D2A50u;
constexpr size_t kEncodeBlock = 64 * 1024;

std::vector<uint8_t> lz4_frame_decode(const uint8_t* p, size_t n) {
    std::vector<uint8_t> out;
    size_t pos = 0;
    auto need = [&](size_t k) {
        if (n - pos < k) io_error("LZ4 frame: truncated");
    };
    bool any = false;
    while (pos < n) {
        need(4);
        const uint32_t magic = rd32(p + pos);
        pos += 4;
        if ((magic & kSkippableMask) == kSkippableMagic) {
            need(4);
            const uint32_t sz = rd32(p + pos);
            pos += 4;
            need(sz);
            pos += sz;
            continue;
        }
        if (magic != kFrameMagic) io_error("not an LZ4 frame (bad magic number)");
        any = true;
        need(2);
        const uint8_t flg = p[pos], bd = p[pos + 1];
        if ((flg >> 6) != 1) io_error("LZ4 frame: unsupported version");
        if (flg & 0x02) io_error("LZ4 frame: reserved FLG bit set");
        if ((bd & 0x80) || (bd & 0x0f)) io_error("LZ4 frame: reserved BD bits set");
        const bool indep = flg & 0x20, bchk = flg & 0x10, csize = flg & 0x08, cchk = flg & 0x04, dict = flg & 0x01;
        const uint32_t bsid = (bd >> 4) & 7u;
        if (bsid < 4) io_error("LZ4 frame: invalid block maximum size");
        const size_t block_max = (size_t)1 << (8 + 2 * bsid);
        const size_t desc_len = 2 + (csize ? 8 : 0) + (dict ? 4 : 0);
        need(desc_len + 1);
        const uint8_t hc = p[pos + desc_len];
        if (((xxh32(p + pos, desc_len, 0) >> 8) & 0xffu) != hc) io_error("LZ4 frame: header checksum mismatch");
        uint64_t content_size = 0;
        if (csize) content_size = rd64(p + pos + 2);
        if (dict) io_error("LZ4 frame: dictionary frames are not supported");
        pos += desc_len + 1;
        const size_t frame_start = out.size();
        while (true) {
            need(4);
            const uint32_t bsize = rd32(p + pos);
            pos += 4;
            if (bsize == 0) break;  // end mark
            const bool raw = bsize & 0x80000000u;
            const size_t sz = bsize & 0x7fffffffu;
            if (sz > block_max) io_error("LZ4 frame: block larger than the frame's maximum");
            need(sz + (bchk ? 4 : 0));
            if (bchk && xxh32(p + pos, sz, 0) != rd32(p + pos + sz)) io_error("LZ4 frame: block checksum mismatch");
            const size_t block_start = out.size();
            if (raw) {
                out.insert(out.end(), p + pos, p + pos + sz);
            } else {
                lz4_block_decode(p + pos, sz, out, indep ? block_start : frame_start, block_max);
            }
            pos += sz + (bchk ? 4 : 0);
        }
        if (cchk) {
            need(4);
            if (xxh32(out.data() + frame_start, out.size() - frame_start, 0) != rd32(p + pos))
                io_error("LZ4 frame: content checksum mismatch");
            pos += 4;
        }
        if (csize && content_size != out.size() - frame_start) io_error("LZ4 frame: content size mismatch");
    }
    if (!any) io_error("no LZ4 frame in the input");
    return out;
}

std::vector<uint8_t> lz4_frame_encode(const uint8_t* p, size_t n) {
    std::vector<uint8_t> out;
    wr32(out, kFrameMagic);
    const uint8_t desc[2] = {0x64, 0x40};  // version 01, independent blocks, content checksum; 64 KiB blocks
    out.push_back(desc[0]);
    out.push_back(desc[1]);
    out.push_back((uint8_t)((xxh32(desc, 2, 0) >> 8) & 0xffu));
    std::vector<uint8_t> buf;
    for (size_t off = 0; off < n; off += kEncodeBlock) {
        const size_t len = std::min(kEncodeBlock, n - off);
        const size_t c = lz4_block_encode(p + off, len, buf);
        if (c) {
            wr32(out, (uint32_t)c);
            out.insert(out.end(), buf.begin(), buf.begin() + c);
        } else {
            wr32(out, (uint32_t)len | 0x80000000u);
            out.insert(out.end(), p + off, p + off + len);
        }
    }
    wr32(out, 0);
    wr32(out, xxh32(p, n, 0));
    return out;
}

// ---- files -----------------------------------------------------------------------------------
std::vector<uint8_t> read_file(const char* path) {
    if (!path) throw FsaError(FSA_E_INVALID, "path is NULL");
    FILE* f = std::fopen(path, "rb");
    if (!f) io_error(std::string("cannot open ") + path);
    std::vector<uint8_t> v;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) v.insert(v.end(), buf, buf + k);
    const bool err = std::ferror(f);
    std::fclose(f);
    if (err) io_error(std::string("cannot read ") + path);
    return v;
}

void write_file(const char* path, const uint8_t* p, size_t n) {
    if (!path) throw FsaError(FSA_E_INVALID, "path is NULL");
    FILE* f = std::fopen(path, "wb");
    if (!f) io_error(std::string("cannot create ") + path);
    const size_t k = n ? std::fwrite(p, 1, n, f) : 0;
    const bool err = std::fclose(f) != 0;
    if (k != n || err) io_error(std::string("cannot write ") + path);
}

// ---- particle data files ---------------------------------------------------------------------
constexpr uint32_t kComponentBytes[7] = {0, 24, 36, 5, 12, 8, 12};

struct DataFile {
    fsa_data_header h{};
    std::vector<uint8_t> bytes;  // decompressed content
    size_t records = 0;          // offset of the first record
    size_t record_size = 0;
};

DataFile parse_data(const char* path) {
    DataFile d;
    const std::vector<uint8_t> raw = read_file(path);
    d.bytes = lz4_frame_decode(raw.data(), raw.size());
    const std::vector<uint8_t>& b = d.bytes;
    if (b.size() < 8) io_error("particle file: truncated header");
    d.h.version = rd32(b.data());
    if (d.h.version != 1) io_error("Unsupported particle data format version");
    const uint32_t nc = rd32(b.data() + 4);
    if (nc > 64 || b.size() < 8 + 4ull * nc + 8) io_error("particle file: truncated header");
    for (uint32_t k = 0; k < nc; k++) {
        const uint32_t id = rd32(b.data() + 8 + 4 * k);
        if (id < 1 || id > 6) io_error("invalid component id");
        d.h.components |= 1u << (id - 1);
    }
    d.h.count = rd64(b.data() + 8 + 4 * nc);
    d.records = 8 + 4 * nc + 8;
    for (uint32_t id = 1; id <= 6; id++)
        if (d.h.components & (1u << (id - 1))) d.record_size += kComponentBytes[id];
    const bool m2 = d.h.components & 1u, m3 = d.h.components & 2u;
    if (m2 && m3) io_error("particle file: both 2D and 3D movement components");
    d.h.dimension = m3 ? 3 : (m2 ? 2 : 0);
    if (d.record_size && d.h.count > (b.size() - d.records) / d.record_size)
        io_error("particle file: fewer records than the particle count");
    return d;
}

void read_into(const DataFile& d, const fsa_particle_view* v) {
    if (!v) throw FsaError(FSA_E_INVALID, "view is NULL");
    if (v->count != d.h.count) throw FsaError(FSA_E_INVALID, "view count differs from the file's particle count");
    const uint32_t dim = d.h.dimension ? d.h.dimension : ((d.h.components & 16u) ? 2u : 3u);
    const uint32_t c = d.h.components;
    const uint8_t* r = d.bytes.data() + d.records;
    for (uint64_t i = 0; i < d.h.count; i++, r += d.record_size) {
        const uint8_t* q = r;
        float* mv = v->movement ? (float*)((char*)v->movement + i * v->movement_stride) : nullptr;
        if (c & 3u) {
            for (uint32_t k = 0; k < 3 * dim; k++)
                if (mv) mv[k] = rdf(q + 4 * k);
            q += 12 * dim;
        } else if (mv) {
            for (uint32_t k = 0; k < 3 * dim; k++) mv[k] = 0.0f;
        }
        uint8_t* info = v->info ? (uint8_t*)v->info + i * v->info_stride : nullptr;
        if (c & 4u) {
            if (info) {
                const uint32_t tag = rd32(q + 1);
                std::memcpy(info, &tag, 4);
                info[4] = q[0];
            }
            q += 5;
        } else if (info) {
            std::memset(info, 0, 5);
        }
        float* pd = v->particle_data ? (float*)((char*)v->particle_data + i * v->particle_data_stride) : nullptr;
        if (c & 8u) {
            if (pd) {  // file: density, mass, pressure; record: mass, pressure, density
                pd[2] = rdf(q);
                pd[0] = rdf(q + 4);
                pd[1] = rdf(q + 8);
            }
            q += 12;
        } else if (pd) {
            pd[0] = pd[1] = pd[2] = 0.0f;
        }
        float* ex = v->external ? (float*)((char*)v->external + i * v->external_stride) : nullptr;
        const uint32_t edim = (c & 16u) ? 2u : ((c & 32u) ? 3u : 0u);
        if (edim) {
            for (uint32_t k = 0; k < edim; k++)
                if (ex) ex[k] = rdf(q + 4 * k);
            q += 4 * edim;
        } else if (ex) {
            for (uint32_t k = 0; k < dim; k++) ex[k] = 0.0f;
        }
    }
}

std::vector<uint8_t> serialize_data(uint32_t dim, const fsa_particle_view* v) {
    if (!v) throw FsaError(FSA_E_INVALID, "view is NULL");
    if (dim != 2 && dim != 3) throw FsaError(FSA_E_INVALID, "dimension must be 2 or 3");
    if (!v->movement) throw FsaError(FSA_E_INVALID, "the movement component is required");
    std::vector<uint32_t> ids;
    ids.push_back(dim == 2 ? 1u : 2u);
    if (v->info) ids.push_back(3);
    if (v->particle_data) ids.push_back(4);
    if (v->external) ids.push_back(dim == 2 ? 5u : 6u);
    size_t rec = 0;
    for (uint32_t id : ids) rec += kComponentBytes[id];
    std::vector<uint8_t> b;
    b.reserve(16 + 4 * ids.size() + rec * v->count);
    wr32(b, 1u);
    wr32(b, (uint32_t)ids.size());
    for (uint32_t id : ids) wr32(b, id);
    wr64(b, v->count);
    for (uint64_t i = 0; i < v->count; i++) {
        const float* mv = (const float*)((const char*)v->movement + i * v->movement_stride);
        for (uint32_t k = 0; k < 3 * dim; k++) wrf(b, mv[k]);
        if (v->info) {
            const uint8_t* info = (const uint8_t*)v->info + i * v->info_stride;
            uint32_t tag;
            std::memcpy(&tag, info, 4);
            b.push_back(info[4]);
            wr32(b, tag);
        }
        if (v->particle_data) {
            const float* pd = (const float*)((const char*)v->particle_data + i * v->particle_data_stride);
            wrf(b, pd[2]);
            wrf(b, pd[0]);
            wrf(b, pd[1]);
        }
        if (v->external) {
            const float* ex = (const float*)((const char*)v->external + i * v->external_stride);
            for (uint32_t k = 0; k < dim; k++) wrf(b, ex[k]);
        }
    }
    return b;
}

// ---- JSON ------------------------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> fields;

    const Json* find(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& f : fields)
            if (f.first == key) return &f.second;
        return nullptr;
    }
};

class JsonParser {
  public:
    JsonParser(const char* p, size_t n) : p_(p), end_(p + n) {}
    Json parse() {
        Json v = value(0);
        ws();
        if (p_ != end_) fail("trailing characters");
        return v;
    }

  private:
    const char* p_;
    const char* end_;
    [[noreturn]] void fail(const char* what) { io_error(std::string("JSON: ") + what); }
    void ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
    }
    bool lit(const char* s) {
        const size_t k = std::strlen(s);
        if ((size_t)(end_ - p_) >= k && std::memcmp(p_, s, k) == 0) {
            p_ += k;
            return true;
        }
        return false;
    }
    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xc0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3f));
        } else if (cp < 0x10000) {
            s += (char)(0xe0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3f));
            s += (char)(0x80 | (cp & 0x3f));
        } else {
            s += (char)(0xf0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3f));
            s += (char)(0x80 | ((cp >> 6) & 0x3f));
            s += (char)(0x80 | (cp & 0x3f));
        }
    }
    uint32_t hex4() {
        if (end_ - p_ < 4) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        std::string s;
        p_++;  // opening quote
        while (true) {
            if (p_ >= end_) fail("unterminated string");
            const char c = *p_++;
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= end_) fail("unterminated escape");
            const char e = *p_++;
            switch (e) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xd800 && cp < 0xdc00) {
                        if (!lit("\\u")) fail("unpaired surrogate");
                        const uint32_t lo = hex4();
                        if (lo < 0xdc00 || lo >= 0xe000) fail("unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    } else if (cp >= 0xdc00 && cp < 0xe000) {
                        fail("unpaired surrogate");
                    }
                    utf8(s, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return s;
    }
    Json value(int depth) {
        if (depth > 200) fail("nesting too deep");
        ws();
        if (p_ >= end_) fail("unexpected end of input");
        Json v;
        const char c = *p_;
        if (c == '{') {
            v.kind = Json::Object;
            p_++;
            ws();
            if (p_ < end_ && *p_ == '}') {
                p_++;
                return v;
            }
            while (true) {
                ws();
                if (p_ >= end_ || *p_ != '"') fail("expected a key");
                std::string k = string();
                ws();
                if (p_ >= end_ || *p_ != ':') fail("expected ':'");
                p_++;
                v.fields.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == '}') {
                    p_++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::Array;
            p_++;
            ws();
            if (p_ < end_ && *p_ == ']') {
                p_++;
                return v;
            }
            while (true) {
                v.items.push_back(value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == ']') {
                    p_++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Json::Bool;
        } else if (lit("null")) {
            v.kind = Json::Null;
        } else {
            // number: -?(0|[1-9][0-9]*)(.[0-9]+)?([eE][+-]?[0-9]+)?
            const char* s = p_;
            if (p_ < end_ && *p_ == '-') p_++;
            if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) fail("unexpected character");
            while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                                 *p_ == '+' || *p_ == '-'))
                p_++;
            v.kind = Json::Number;
            const auto r = std::from_chars(s, p_, v.num);
            if (r.ec != std::errc() || r.ptr != p_) fail("malformed number");
        }
        return v;
    }
};

// the reference reads settings with nlohmann's get<T>() and fails on missing or mistyped keys
const Json& need(const Json& node, const char* key, const std::string& where) {
    const Json* v = node.find(key);
    if (!v) io_error("scenario: missing \"" + where + key + "\"");
    return *v;
}
float get_float(const Json& node, const char* key, const std::string& where) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::Number) io_error("scenario: \"" + where + key + "\" is not a number");
    return (float)v.num;
}
uint32_t get_count(const Json& node, const char* key, const std::string& where) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::Number || v.num < 0 || v.num != std::floor(v.num) || v.num > 4294967295.0)
        io_error("scenario: \"" + where + key + "\" is not a non-negative integer");
    return (uint32_t)v.num;
}
bool get_bool(const Json& node, const char* key, const std::string& where) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::Bool) io_error("scenario: \"" + where + key + "\" is not a boolean");
    return v.b;
}
std::string get_string(const Json& node, const char* key, const std::string& where) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::String) io_error("scenario: \"" + where + key + "\" is not a string");
    return v.str;
}
const Json& get_object(const Json& node, const char* key, const std::string& where) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::Object) io_error("scenario: \"" + where + key + "\" is not an object");
    return v;
}

void get_vec3(const Json& node, const char* key, const std::string& where, float out[3]) {
    const Json& v = need(node, key, where);
    if (v.kind != Json::Array || v.items.size() != 3) io_error("scenario: \"" + where + key + "\" is not a 3-vector");
    for (int k = 0; k < 3; k++) {
        if (v.items[k].kind != Json::Number) io_error("scenario: \"" + where + key + "\" is not a 3-vector");
        out[k] = (float)v.items[k].num;
    }
}

std::string parent_dir(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k);
}
bool file_exists(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

void copy_str(char* dst, size_t cap, const std::string& s, const char* what) {
    if (s.size() + 1 > cap) io_error(std::string("scenario: ") + what + " is too long");
    std::memcpy(dst, s.c_str(), s.size() + 1);
}

void parse_scenario(const char* json_path, fsa_scenario* out) {
    if (!out) throw FsaError(FSA_E_INVALID, "out is NULL");
    const std::vector<uint8_t> raw = read_file(json_path);
    const Json root = JsonParser((const char*)raw.data(), raw.size()).parse();
    if (root.kind != Json::Object) io_error("scenario: the root is not an object");
    const Json* ver = root.find("version");
    if (!ver || ver->kind != Json::Number || ver->num != 2.0) io_error("Invalid file version number!");

    const Json& sc = get_object(root, "scenario", "");
    const Json& so = get_object(root, "solver", "");
    const Json& tg = get_object(root, "timestep-generator", "");
    const std::string type = get_string(so, "type", "solver.");
    const std::string search = get_string(get_object(so, "neighborhood-search", "solver."), "type",
                                          "solver.neighborhood-search.");
    const std::string kernel = get_string(get_object(so, "kernel", "solver."), "type", "solver.kernel.");

    // the solver / search / kernel combinations SolverSerializer::deserialize builds
    uint32_t dim = 0, solver = 0, srch = 0;
    if (kernel == "cubic-spline-kernel" && (type == "sesph" || type == "iisph")) {
        dim = 2;
        if (search == "quadratic-dynamic-allocated") srch = FSA_SEARCH_QUADRATIC_2D;
        else if (search == "hashed") srch = FSA_SEARCH_HASHED_2D;
    } else if (kernel == "cubic-spline-kernel-3d" && (type == "sesph-3d" || type == "iisph-3d")) {
        dim = 3;
        if (search == "quadratic-dynamic-allocated-3d") srch = FSA_SEARCH_QUADRATIC_3D;
        else if (search == "hashed-3d") srch = FSA_SEARCH_HASHED_3D;
        else if (search == "compressed-3d") srch = FSA_SEARCH_COMPRESSED_3D;
    }
    if (!dim || !srch) io_error("Encountered unknown solver, neighborhood search, kernel combination!");
    solver = (type == "iisph" || type == "iisph-3d") ? FSA_SOLVER_IISPH : FSA_SOLVER_SESPH;

    fsa_scenario_defaults(out, dim, solver);
    fsa_scene_desc& d = out->desc;
    d.search = srch;
    const std::string particles = get_string(sc, "particles", "scenario.");
    d.gravity = get_float(sc, "gravity", "scenario.");
    d.particle_size = get_float(sc, "particle-size", "scenario.");
    d.rest_density = get_float(sc, "rest-density", "scenario.");
    if (const Json* e = sc.find("entities")) {
        if (e->kind != Json::Array) io_error("scenario: \"scenario.entities\" is not an array");
        out->entities = (uint32_t)e->items.size();
        for (size_t k = 0; k < e->items.size(); k++) {
            const Json& en = e->items[k];
            const std::string where = "scenario.entities[" + std::to_string(k) + "].";
            if (en.kind != Json::Object) io_error("scenario: an entity is not an object");
            const std::string t = get_string(en, "type", where);
            fsa_entity ent{};
            if (t == "particle-remover-3d") {
                const Json& vol = get_object(en, "volume", where);
                get_vec3(vol, "center", where + "volume.", ent.center);
                get_vec3(vol, "distance-from-center", where + "volume.", ent.distance_from_center);
                ent.type = FSA_ENTITY_PARTICLE_REMOVER_3D;
                ent.remove_if_outside = get_bool(en, "remove-if-outside", where) ? 1u : 0u;
            } else if (t == "boundary-preprocessor") {
                ent.type = FSA_ENTITY_BOUNDARY_PREPROCESSOR;
            } else if (t == "velocity-alteration-by-tag") {
                ent.type = FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG;
                ent.tag = get_count(en, "tag", where);
                get_vec3(en, "velocity", where, ent.velocity);
            } else if (t == "particle-remover" || t == "particle-spawner") {
                continue;  // 2D entities: reported through the entity count only
            } else {
                io_error("Encountered invalid entity type!");
            }
            if (out->entity_count >= sizeof(out->entity_list) / sizeof(out->entity_list[0]))
                io_error("scenario: more than 32 entities");
            out->entity_list[out->entity_count++] = ent;
        }
    }
    if (const Json* s = sc.find("sensors")) out->sensors = s->kind == Json::Array ? (uint32_t)s->items.size() : 0;

    if (solver == FSA_SOLVER_SESPH) {
        d.sesph_stiffness_k = get_float(so, "stiffness", "solver.");
        d.sesph_viscosity = get_float(so, "viscosity", "solver.");
        if (dim == 3) {
            d.sesph_single_layer_boundary = get_bool(so, "single-layer-boundary-enabled", "solver.");
            if (d.sesph_single_layer_boundary) {
                const Json& sl = get_object(so, "single-layer-settings", "solver.");
                d.sesph_single_layer_gamma_1 = get_float(sl, "gamma-1", "solver.single-layer-settings.");
                d.sesph_single_layer_gamma_2 = get_float(sl, "gamma-2", "solver.single-layer-settings.");
            }
        }
    } else {
        d.iisph_gamma = get_float(so, "gamma", "solver.");
        d.iisph_omega = get_float(so, "omega", "solver.");
        d.iisph_max_density_error_allowed = get_float(so, "max-density-error", "solver.");
        d.iisph_max_iterations = get_count(so, "max-iterations", "solver.");
        d.iisph_min_iterations = get_count(so, "min-iterations", "solver.");
        d.iisph_viscosity = get_float(so, "viscosity", "solver.");
        if (dim == 3) {
            d.iisph_single_layer_boundary = get_bool(so, "single-layer-boundary-enabled", "solver.");
            if (d.iisph_single_layer_boundary) {
                const Json& sl = get_object(so, "single-layer-settings", "solver.");
                d.iisph_single_layer_gamma_1 = get_float(sl, "gamma-1", "solver.single-layer-settings.");
            }
        }
    }

    const std::string tgt = get_string(tg, "type", "timestep-generator.");
    if (tgt == "constant") {
        d.timestep_generator = FSA_TIMESTEP_CONSTANT;
        d.timestep = get_float(tg, "timestep", "timestep-generator.");
    } else if (tgt == "dynamic-cfl") {
        d.timestep_generator = FSA_TIMESTEP_DYNAMIC_CFL;
        d.max_timestep = get_float(tg, "max-timestep", "timestep-generator.");
        d.min_timestep = get_float(tg, "min-timestep", "timestep-generator.");
        d.cfl_number = get_float(tg, "cfl", "timestep-generator.");
    } else {
        io_error("Encountered unknown timestep generator type!");
    }

    copy_str(out->particles, sizeof(out->particles), particles, "scenario.particles");
    // MainSerializer::get_full_particle_data_path: beside the JSON file, else as written
    const std::string dir = parent_dir(json_path);
    std::string full = particles;
    if (!particles.empty() && particles[0] != '/') {
        const std::string beside = dir.empty() ? particles : dir + "/" + particles;
        full = file_exists(beside) ? beside : particles;
    }
    copy_str(out->particles_path, sizeof(out->particles_path), full, "the particle file path");
}

// floats are written with 9 significant digits, which read back to the same float
std::string num(float f) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.9g", (double)f);
    return buf;
}
std::string quoted(const std::string& s) {
    std::string r = "\"";
    for (unsigned char c : s) {
        if (c == '"' || c == '\\') {
            r += '\\';
            r += (char)c;
        } else if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof(buf), "\\u%04x", c);
            r += buf;
        } else {
            r += (char)c;
        }
    }
    return r + "\"";
}

std::string format_scenario(const fsa_scenario* in) {
    if (!in) throw FsaError(FSA_E_INVALID, "scenario is NULL");
    const fsa_scene_desc& d = in->desc;
    if (d.dimension != 2 && d.dimension != 3) throw FsaError(FSA_E_INVALID, "dimension must be 2 or 3");
    if (d.solver != FSA_SOLVER_IISPH && d.solver != FSA_SOLVER_SESPH) throw FsaError(FSA_E_INVALID, "unknown solver");
    const char* search = nullptr;
    const uint32_t s = d.search ? d.search : (d.dimension == 2 ? FSA_SEARCH_HASHED_2D : FSA_SEARCH_COMPRESSED_3D);
    if (d.dimension == 2) {
        if (s == FSA_SEARCH_QUADRATIC_2D) search = "quadratic-dynamic-allocated";
        else if (s == FSA_SEARCH_HASHED_2D) search = "hashed";
    } else {
        if (s == FSA_SEARCH_QUADRATIC_3D) search = "quadratic-dynamic-allocated-3d";
        else if (s == FSA_SEARCH_HASHED_3D) search = "hashed-3d";
        else if (s == FSA_SEARCH_COMPRESSED_3D) search = "compressed-3d";
    }
    if (!search) throw FsaError(FSA_E_INVALID, "search does not match the dimension");
    const bool iisph = d.solver == FSA_SOLVER_IISPH;
    const std::string I1 = "    ", I2 = I1 + I1, I3 = I2 + I1;
    std::string o = "{\n";
    o += I1 + "\"version\": 2,\n";
    o += I1 + "\"scenario\": {\n";
    size_t plen = 0;
    while (plen < sizeof(in->particles) && in->particles[plen]) plen++;
    o += I2 + "\"particles\": " + quoted(std::string(in->particles, plen)) + ",\n";
    o += I2 + "\"gravity\": " + num(d.gravity) + ",\n";
    o += I2 + "\"particle-size\": " + num(d.particle_size) + ",\n";
    o += I2 + "\"rest-density\": " + num(d.rest_density);
    if (in->entity_count) {
        if (in->entity_count > sizeof(in->entity_list) / sizeof(in->entity_list[0]))
            throw FsaError(FSA_E_INVALID, "entity_count exceeds the entity list");
        auto vec = [&](const float* v) { return "[" + num(v[0]) + ", " + num(v[1]) + ", " + num(v[2]) + "]"; };
        o += ",\n" + I2 + "\"entities\": [";
        for (uint32_t k = 0; k < in->entity_count; k++) {
            const fsa_entity& e = in->entity_list[k];
            o += k ? ",\n" : "\n";
            if (e.type == FSA_ENTITY_PARTICLE_REMOVER_3D) {
                o += I3 + "{\"type\": \"particle-remover-3d\", \"volume\": {\"center\": " + vec(e.center) +
                     ", \"distance-from-center\": " + vec(e.distance_from_center) + "}, \"remove-if-outside\": " +
                     (e.remove_if_outside ? "true" : "false") + "}";
            } else if (e.type == FSA_ENTITY_BOUNDARY_PREPROCESSOR) {
                o += I3 + "{\"type\": \"boundary-preprocessor\"}";
            } else if (e.type == FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG) {
                o += I3 + "{\"type\": \"velocity-alteration-by-tag\", \"tag\": " + std::to_string(e.tag) +
                     ", \"velocity\": " + vec(e.velocity) + "}";
            } else {
                throw FsaError(FSA_E_INVALID, "unknown entity type");
            }
        }
        o += "\n" + I2 + "]";
    }
    o += "\n" + I1 + "},\n";
    o += I1 + "\"solver\": {\n";
    o += I2 + "\"type\": \"" + std::string(iisph ? "iisph" : "sesph") + (d.dimension == 3 ? "-3d" : "") + "\",\n";
    o += I2 + "\"neighborhood-search\": {\n" + I3 + "\"type\": \"" + search + "\"\n" + I2 + "},\n";
    o += I2 + "\"kernel\": {\n" + I3 + "\"type\": \"cubic-spline-kernel" + (d.dimension == 3 ? "-3d" : "") + "\"\n" + I2 +
         "},\n";
    if (iisph) {
        o += I2 + "\"gamma\": " + num(d.iisph_gamma) + ",\n";
        o += I2 + "\"omega\": " + num(d.iisph_omega) + ",\n";
        o += I2 + "\"max-density-error\": " + num(d.iisph_max_density_error_allowed) + ",\n";
        o += I2 + "\"max-iterations\": " + std::to_string(d.iisph_max_iterations) + ",\n";
        o += I2 + "\"min-iterations\": " + std::to_string(d.iisph_min_iterations) + ",\n";
        o += I2 + "\"viscosity\": " + num(d.iisph_viscosity);
        if (d.dimension == 3) {
            o += ",\n" + I2 + "\"single-layer-boundary-enabled\": " + (d.iisph_single_layer_boundary ? "true" : "false");
            if (d.iisph_single_layer_boundary)
                o += ",\n" + I2 + "\"single-layer-settings\": {\n" + I3 + "\"gamma-1\": " +
                     num(d.iisph_single_layer_gamma_1) + "\n" + I2 + "}";
        }
    } else {
        o += I2 + "\"stiffness\": " + num(d.sesph_stiffness_k) + ",\n";
        o += I2 + "\"viscosity\": " + num(d.sesph_viscosity);
        if (d.dimension == 3) {
            o += ",\n" + I2 + "\"single-layer-boundary-enabled\": " + (d.sesph_single_layer_boundary ? "true" : "false");
            if (d.sesph_single_layer_boundary)
                o += ",\n" + I2 + "\"single-layer-settings\": {\n" + I3 + "\"gamma-1\": " +
                     num(d.sesph_single_layer_gamma_1) + ",\n" + I3 + "\"gamma-2\": " +
                     num(d.sesph_single_layer_gamma_2) + "\n" + I2 + "}";
        }
    }
    o += "\n" + I1 + "},\n";
    o += I1 + "\"timestep-generator\": {\n";
    if (d.timestep_generator == FSA_TIMESTEP_DYNAMIC_CFL) {
        o += I2 + "\"type\": \"dynamic-cfl\",\n";
        o += I2 + "\"max-timestep\": " + num(d.max_timestep) + ",\n";
        o += I2 + "\"min-timestep\": " + num(d.min_timestep) + ",\n";
        o += I2 + "\"cfl\": " + num(d.cfl_number) + "\n";
    } else if (d.timestep_generator == FSA_TIMESTEP_CONSTANT) {
        o += I2 + "\"type\": \"constant\",\n";
        o += I2 + "\"timestep\": " + num(d.timestep) + "\n";
    } else {
        throw FsaError(FSA_E_INVALID, "unknown timestep generator");
    }
    o += I1 + "},\n";
    o += I1 + "\"visualizer\": {\n" + I2 + "\"type\": \"no-visualizer\"\n" + I1 + "}\n";
    o += "}\n";
    return o;
}

}  // namespace io
}  // namespace fsa

using namespace fsa;

FSA_API void fsa_scenario_defaults(fsa_scenario* out, uint32_t dimension, uint32_t solver) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    fsa_scene_desc_defaults(&out->desc, dimension, solver);
}

FSA_API int32_t fsa_data_read_header(const char* path, fsa_data_header* out) {
    return guarded([&] {
        if (!out) throw FsaError(FSA_E_INVALID, "out is NULL");
        *out = io::parse_data(path).h;
    });
}

FSA_API int32_t fsa_data_read(const char* path, const fsa_particle_view* view) {
    return guarded([&] { io::read_into(io::parse_data(path), view); });
}

FSA_API int32_t fsa_data_write(const char* path, uint32_t dimension, const fsa_particle_view* view) {
    return guarded([&] {
        const std::vector<uint8_t> b = io::serialize_data(dimension, view);
        const std::vector<uint8_t> z = io::lz4_frame_encode(b.data(), b.size());
        io::write_file(path, z.data(), z.size());
    });
}

FSA_API int32_t fsa_scenario_read(const char* json_path, fsa_scenario* out) {
    return guarded([&] { io::parse_scenario(json_path, out); });
}

FSA_API int32_t fsa_scenario_write(const char* json_path, const fsa_scenario* in) {
    return guarded([&] {
        const std::string s = io::format_scenario(in);
        io::write_file(json_path, (const uint8_t*)s.data(), s.size());
    });
}

FSA_API int32_t fsa_lz4_frame_decode(const void* src, uint64_t src_size, void* dst, uint64_t capacity,
                                     uint64_t* out_size) {
    return guarded([&] {
        if (!out_size || (!src && src_size)) throw FsaError(FSA_E_INVALID, "NULL argument");
        const std::vector<uint8_t> v = io::lz4_frame_decode((const uint8_t*)src, src_size);
        *out_size = v.size();
        if (v.size() > capacity) throw FsaError(FSA_E_CAPACITY, "output buffer too small");
        if (!v.empty()) std::memcpy(dst, v.data(), v.size());
    });
}

FSA_API int32_t fsa_lz4_frame_encode(const void* src, uint64_t src_size, void* dst, uint64_t capacity,
                                     uint64_t* out_size) {
    return guarded([&] {
        if (!out_size || (!src && src_size)) throw FsaError(FSA_E_INVALID, "NULL argument");
        const std::vector<uint8_t> v = io::lz4_frame_encode((const uint8_t*)src, src_size);
        *out_size = v.size();
        if (v.size() > capacity) throw FsaError(FSA_E_CAPACITY, "output buffer too small");
        std::memcpy(dst, v.data(), v.size());
    });
}

FSA_API uint32_t fsa_xxh32(const void* data, uint64_t size, uint32_t seed) {
    if (!data && size) return 0;
    return io::xxh32((const uint8_t*)data, size, seed);
}
