// rc_dump.cpp -- binary dump of G-buffers, cascades and irradiance (host only).
//
// The reference keeps no GI data on disk (its only assets are OBJ/MTL files
// and textures, src/primitives.rs); this container carries named planar fp32
// arrays between machines (fixtures, captured frames, rows for the trainer)
// with a checksum:
//
//   offset  size          field
//   0       8             magic "RCDUMP01"
//   8       4             version (1)
//   12      4             n_planes
//   16      8             payload bytes
//   24      32            SHA-256 (FIPS 180-4) of everything after this
//                         field: the plane records, then the payload
//   56      48 * n        plane records: name[16] (NUL-padded), channels,
//                         height, width, reserved (u32), offset (u64, from
//                         the payload start), count (u64, floats)
//   ..      payload       the planes' fp32 values, little-endian, in order
//
// All integers little-endian.  Readers check magic, version, that the sizes
// account for the file exactly (before allocating), and the checksum, before
// handing out any value.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/rc_gi.h"

namespace {

// ---- SHA-256 (FIPS 180-4) -------------------------------------------------------
constexpr uint32_t kK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

struct Sha256 {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint8_t buf[64];
    size_t fill = 0;
    uint64_t total = 0;

    void block(const uint8_t *p)
    {
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) | ((uint32_t)p[4 * i + 2] << 8) | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kK[i] + w[i];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a;
        h[1] += b;
        h[2] += c;
        h[3] += d;
        h[4] += e;
        h[5] += f;
        h[6] += g;
        h[7] += hh;
    }
    void update(const void *data, size_t n)
    {
        const uint8_t *p = (const uint8_t *)data;
        total += n;
        while (n) {
            const size_t take = n < 64 - fill ? n : 64 - fill;
            memcpy(buf + fill, p, take);
            fill += take;
            p += take;
            n -= take;
            if (fill == 64) {
                block(buf);
                fill = 0;
            }
        }
    }
    void finish(uint8_t out[32])
    {
        const uint64_t bits = total * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; ++i) len[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(len, 8);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h[i] >> (24 - 8 * j));
    }
};

constexpr char kMagic[8] = {'R', 'C', 'D', 'U', 'M', 'P', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 56, kRecord = 48;

void put32(uint8_t *p, uint32_t v)
{
    for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
void put64(uint8_t *p, uint64_t v)
{
    for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
uint32_t get32(const uint8_t *p)
{
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)p[i] << (8 * i);
    return v;
}
uint64_t get64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

thread_local std::string g_dump_err;

rc_status dfail(rc_status s, const std::string &m)
{
    g_dump_err = m;
    return s;
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

// Reads and checks header, records and checksum; leaves the payload in `payload`.
rc_status load(const char *path, std::vector<rc_plane> &planes, std::vector<std::string> &names,
               std::vector<uint8_t> &payload)
{
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return dfail(RC_ERR_INVALID, std::string("rc_dump: cannot open ") + path);
    uint8_t hd[kHeader];
    if (fread(hd, 1, kHeader, fh.f) != kHeader) return dfail(RC_ERR_INVALID, "rc_dump: truncated header");
    if (memcmp(hd, kMagic, 8) != 0) return dfail(RC_ERR_INVALID, "rc_dump: not an RCDUMP01 file");
    if (get32(hd + 8) != kVersion) return dfail(RC_ERR_INVALID, "rc_dump: unsupported version");
    const uint32_t n = get32(hd + 12);
    const uint64_t bytes = get64(hd + 16);
    if (n > 4096 || bytes > (1ull << 40)) return dfail(RC_ERR_INVALID, "rc_dump: implausible sizes");
    // the header's sizes must account for the file exactly before anything
    // is allocated from them
    if (fseek(fh.f, 0, SEEK_END) != 0) return dfail(RC_ERR_INVALID, "rc_dump: cannot seek");
    const long long fsize = ftell(fh.f);
    const unsigned long long want = kHeader + kRecord * (unsigned long long)n + bytes;
    if (fsize < 0 || (unsigned long long)fsize < want) return dfail(RC_ERR_INVALID, "rc_dump: truncated file");
    if ((unsigned long long)fsize > want) return dfail(RC_ERR_INVALID, "rc_dump: trailing bytes");
    if (fseek(fh.f, (long)kHeader, SEEK_SET) != 0) return dfail(RC_ERR_INVALID, "rc_dump: cannot seek");
    std::vector<uint8_t> rec(kRecord * (size_t)n);
    if (n && fread(rec.data(), 1, rec.size(), fh.f) != rec.size()) return dfail(RC_ERR_INVALID, "rc_dump: truncated records");
    payload.resize((size_t)bytes);
    if (bytes && fread(payload.data(), 1, (size_t)bytes, fh.f) != (size_t)bytes)
        return dfail(RC_ERR_INVALID, "rc_dump: truncated payload");
    Sha256 sh;
    sh.update(rec.data(), rec.size());
    sh.update(payload.data(), payload.size());
    uint8_t dg[32];
    sh.finish(dg);
    if (memcmp(dg, hd + 24, 32) != 0) return dfail(RC_ERR_INVALID, "rc_dump: checksum mismatch");
    planes.assign(n, rc_plane{});
    names.assign(n, std::string());
    for (uint32_t i = 0; i < n; ++i) {
        const uint8_t *r = rec.data() + kRecord * i;
        char nm[17] = {0};
        memcpy(nm, r, 16);
        names[i] = nm;
        rc_plane &p = planes[i];
        memcpy(p.name, r, 16);
        p.channels = get32(r + 16);
        p.height = get32(r + 20);
        p.width = get32(r + 24);
        p.reserved = 0;
        const uint64_t off = get64(r + 32), cnt = get64(r + 40);
        // every product and sum is checked without wrapping: a record whose
        // sizes overflow 64 bits, or whose values run past the payload, is refused
        uint64_t prod = 0;
        const bool wraps = __builtin_mul_overflow((uint64_t)p.channels, (uint64_t)p.height, &prod) ||
                           __builtin_mul_overflow(prod, (uint64_t)p.width, &prod);
        if (wraps || cnt != prod || off % 4 || off > bytes || cnt > (bytes - off) / 4)
            return dfail(RC_ERR_INVALID, "rc_dump: bad plane record " + names[i]);
        p.data = (const float *)(uintptr_t)off;  // offset until handed out
    }
    return RC_OK;
}

}  // namespace

extern "C" {

void rc_sha256(const void *data, uint64_t n, uint8_t *out32)
{
    Sha256 sh;
    sh.update(data, (size_t)n);
    sh.finish(out32);
}

const char *rc_dump_last_error(void) { return g_dump_err.c_str(); }

rc_status rc_dump_write(const char *path, const rc_plane *planes, uint32_t n)
{
    if (!path || (n && !planes)) return dfail(RC_ERR_NULL, "rc_dump_write: NULL argument");
    if (n > 4096) return dfail(RC_ERR_INVALID, "rc_dump_write: too many planes");
    std::vector<uint8_t> rec(kRecord * (size_t)n, 0);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const rc_plane &p = planes[i];
        uint64_t cnt = 0;
        if (__builtin_mul_overflow((uint64_t)p.channels, (uint64_t)p.height, &cnt) ||
            __builtin_mul_overflow(cnt, (uint64_t)p.width, &cnt) || cnt > (1ull << 38) || off + 4 * cnt > (1ull << 40))
            return dfail(RC_ERR_INVALID, "rc_dump_write: plane too large");
        if (cnt && !p.data) return dfail(RC_ERR_NULL, "rc_dump_write: plane without data");
        for (uint32_t j = 0; j < i; ++j)
            if (strncmp(planes[j].name, p.name, 16) == 0) return dfail(RC_ERR_INVALID, "rc_dump_write: duplicate plane name");
        uint8_t *r = rec.data() + kRecord * i;
        strncpy((char *)r, p.name, 16);
        put32(r + 16, p.channels);
        put32(r + 20, p.height);
        put32(r + 24, p.width);
        put64(r + 32, off);
        put64(r + 40, cnt);
        off += 4 * cnt;
    }
    Sha256 sh;
    sh.update(rec.data(), rec.size());
    for (uint32_t i = 0; i < n; ++i)
        sh.update(planes[i].data, (size_t)4 * planes[i].channels * planes[i].height * planes[i].width);
    uint8_t hd[kHeader];
    memcpy(hd, kMagic, 8);
    put32(hd + 8, kVersion);
    put32(hd + 12, n);
    put64(hd + 16, off);
    sh.finish(hd + 24);
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) return dfail(RC_ERR_INVALID, std::string("rc_dump_write: cannot create ") + path);
    bool ok = fwrite(hd, 1, kHeader, fh.f) == kHeader;
    if (ok && n) ok = fwrite(rec.data(), 1, rec.size(), fh.f) == rec.size();
    for (uint32_t i = 0; ok && i < n; ++i) {
        const size_t b = (size_t)4 * planes[i].channels * planes[i].height * planes[i].width;
        if (b) ok = fwrite(planes[i].data, 1, b, fh.f) == b;
    }
    if (!ok) return dfail(RC_ERR_INVALID, "rc_dump_write: write failed");
    return RC_OK;
}

rc_status rc_dump_info(const char *path, rc_plane *planes, uint32_t cap, uint32_t *n_out)
{
    if (!path || !n_out) return dfail(RC_ERR_NULL, "rc_dump_info: NULL argument");
    std::vector<rc_plane> ps;
    std::vector<std::string> names;
    std::vector<uint8_t> payload;
    const rc_status st = load(path, ps, names, payload);
    if (st != RC_OK) return st;
    *n_out = (uint32_t)ps.size();
    for (uint32_t i = 0; planes && i < cap && i < ps.size(); ++i) {
        planes[i] = ps[i];
        planes[i].data = nullptr;
    }
    return RC_OK;
}

rc_status rc_dump_read(const char *path, const char *name, float *dst, uint64_t count)
{
    if (!path || !name || !dst) return dfail(RC_ERR_NULL, "rc_dump_read: NULL argument");
    std::vector<rc_plane> ps;
    std::vector<std::string> names;
    std::vector<uint8_t> payload;
    const rc_status st = load(path, ps, names, payload);
    if (st != RC_OK) return st;
    for (size_t i = 0; i < ps.size(); ++i) {
        if (names[i] != std::string(name).substr(0, 16)) continue;
        const uint64_t cnt = (uint64_t)ps[i].channels * ps[i].height * ps[i].width;
        if (cnt != count) return dfail(RC_ERR_INVALID, "rc_dump_read: size mismatch for " + names[i]);
        memcpy(dst, payload.data() + (uintptr_t)ps[i].data, (size_t)(4 * cnt));
        return RC_OK;
    }
    return dfail(RC_ERR_INVALID, std::string("rc_dump_read: no plane ") + name);
}

}  // extern "C"
