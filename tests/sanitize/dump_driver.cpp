// The dump reader and SHA-256 (radiancecascade_amd/csrc/rc_dump.cpp, host
// only) under AddressSanitizer / UndefinedBehaviorSanitizer: known answers,
// a round trip, and every single-byte change and every truncation of a small
// file (the reader parses untrusted input: it must refuse them, never read
// out of bounds).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/rc_gi.h"

static std::string hex(const uint8_t *d)
{
    static const char *h = "0123456789abcdef";
    std::string s;
    for (int i = 0; i < 32; ++i) s += h[d[i] >> 4], s += h[d[i] & 15];
    return s;
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    int bad = 0;
    uint8_t d[32];
    rc_sha256("abc", 3, d);
    bad |= hex(d) != "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad";
    rc_sha256("", 0, d);
    bad |= hex(d) != "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855";

    std::vector<float> a(3 * 5 * 7), b(2 * 3);
    for (size_t i = 0; i < a.size(); ++i) a[i] = 0.25f * (float)i;
    for (size_t i = 0; i < b.size(); ++i) b[i] = -1.0f - (float)i;
    rc_plane pl[2] = {};
    strncpy(pl[0].name, "emissive", 16);
    pl[0].channels = 3, pl[0].height = 5, pl[0].width = 7, pl[0].data = a.data();
    strncpy(pl[1].name, "sixteen_chars_ab", 16);
    pl[1].channels = 1, pl[1].height = 2, pl[1].width = 3, pl[1].data = b.data();
    const std::string path = dir + "/drv.rcdump";
    if (rc_dump_write(path.c_str(), pl, 2) != RC_OK) {
        printf("dump driver: FAILED (write: %s)\n", rc_dump_last_error());
        return 1;
    }
    std::vector<float> back(a.size());
    bad |= rc_dump_read(path.c_str(), "emissive", back.data(), back.size()) != RC_OK || back != a;
    std::vector<float> back2(b.size());
    bad |= rc_dump_read(path.c_str(), "sixteen_chars_ab", back2.data(), back2.size()) != RC_OK || back2 != b;

    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return printf("dump driver: FAILED (reopen)\n"), 1;
    std::vector<uint8_t> raw;
    for (int c; (c = fgetc(f)) != EOF;) raw.push_back((uint8_t)c);
    fclose(f);
    const std::string mut = dir + "/mut.rcdump";
    int accepted = 0;
    auto try_file = [&](const std::vector<uint8_t> &bytes) {
        FILE *g = fopen(mut.c_str(), "wb");
        if (!g) return true;  // counted as a failure
        if (!bytes.empty()) fwrite(bytes.data(), 1, bytes.size(), g);
        fclose(g);
        rc_plane info[4];
        uint32_t n = 0;
        const rc_status s1 = rc_dump_info(mut.c_str(), info, 4, &n);
        const rc_status s2 = rc_dump_read(mut.c_str(), "emissive", back.data(), back.size());
        return s1 == RC_OK || s2 == RC_OK;
    };
    for (size_t i = 0; i < raw.size(); ++i) {
        std::vector<uint8_t> m = raw;
        m[i] ^= 0x5A;
        if (try_file(m)) ++accepted;
    }
    for (size_t len = 0; len < raw.size(); ++len)
        if (try_file(std::vector<uint8_t>(raw.begin(), raw.begin() + len))) ++accepted;
    bad |= accepted != 0;

    // Crafted files whose checksum is valid but whose plane record has sizes
    // that wrap in 64 bits or run past the payload: the checksum cannot
    // refuse them, the record checks must.
    struct Crafted {
        uint32_t channels, height, width;
        uint64_t off, cnt;
    };
    const Crafted crafted[] = {
        {233u, 236248574u, 2429587879u, ~0ull - 3, (1ull << 62) + 2},  // c*h*w and 4*cnt both wrap
        {1u, 1u, 2u, 0, 2},                                             // past an 4-byte payload
        {1u, 1u, 1u, 4, 1},                                             // offset at the payload's end
        {1u, 1u, 1u, 8, 1},                                             // offset beyond the payload
        {1u, 1u, 1u << 31, 0, 1ull << 31},                              // count past the payload
        {1u << 31, 1u << 31, 4u, 0, 0},                                 // product wraps to 0
    };
    int crafted_accepted = 0;
    for (const Crafted &cr : crafted) {
        std::vector<uint8_t> rec(48, 0), payload(4, 0x11);
        memcpy(rec.data(), "emissive", 8);
        auto put32 = [&](size_t at, uint32_t v) { for (int k = 0; k < 4; ++k) rec[at + k] = (uint8_t)(v >> (8 * k)); };
        auto put64 = [&](std::vector<uint8_t> &v8, size_t at, uint64_t v) {
            for (int k = 0; k < 8; ++k) v8[at + k] = (uint8_t)(v >> (8 * k));
        };
        put32(16, cr.channels);
        put32(20, cr.height);
        put32(24, cr.width);
        put64(rec, 32, cr.off);
        put64(rec, 40, cr.cnt);
        std::vector<uint8_t> body = rec;
        body.insert(body.end(), payload.begin(), payload.end());
        std::vector<uint8_t> file(56, 0);
        memcpy(file.data(), "RCDUMP01", 8);
        file[8] = 1;   // version
        file[12] = 1;  // one plane
        put64(file, 16, payload.size());
        rc_sha256(body.data(), body.size(), file.data() + 24);
        file.insert(file.end(), body.begin(), body.end());
        if (try_file(file)) ++crafted_accepted;
    }
    bad |= crafted_accepted != 0;
    accepted += crafted_accepted;
    printf(bad ? "dump driver: FAILED (%d accepted)\n" : "dump driver: ok\n", accepted);
    return bad;
}
