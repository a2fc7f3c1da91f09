// State snapshots: a binary format for a fluid State (state.h:13-17:
// positions, velocities, colors) and its time in flicks, for checkpoints and
// served results. The reference has no such format; this one is defined here.
//
// Layout (little-endian, 48-byte header, then the payload):
//   0  char[8]  magic "EMSPHST" + NUL
//   8  u32      version (1)
//  12  u32      flags: bit 0 velocities present, bit 1 colors present
//  16  u64      particle count n
//  24  i64      time in flicks (1/705600000 s)
//  32  i32      method (emsph_method, or -1)
//  36  u32      reserved (0)
//  40  u64      FNV-1a 64-bit hash of the payload bytes
//  48  f32      positions [n][2], then velocities [n][2] and colors [n][4]
//               when their flags are set
// Floats are stored as their IEEE bit patterns, so a round trip is exact.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emerald_sph.h"

namespace emsph_snapshot {

constexpr char MAGIC[8] = {'E', 'M', 'S', 'P', 'H', 'S', 'T', '\0'};
constexpr uint32_t VERSION = 1;
constexpr size_t HEADER = 48;

uint64_t fnv1a(unsigned char const* p, size_t n) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

void put32(unsigned char* p, uint32_t v) {
    for (int k = 0; k < 4; ++k) { p[k] = static_cast<unsigned char>(v >> (8 * k)); }
}
void put64(unsigned char* p, uint64_t v) {
    for (int k = 0; k < 8; ++k) { p[k] = static_cast<unsigned char>(v >> (8 * k)); }
}
uint32_t get32(unsigned char const* p) {
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) { v |= static_cast<uint32_t>(p[k]) << (8 * k); }
    return v;
}
uint64_t get64(unsigned char const* p) {
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) { v |= static_cast<uint64_t>(p[k]) << (8 * k); }
    return v;
}

size_t payload_floats(uint64_t n, uint32_t flags) {
    size_t f = 2;
    if (flags & 1u) { f += 2; }
    if (flags & 2u) { f += 4; }
    return static_cast<size_t>(n) * f;
}

void put_floats(unsigned char* p, float const* src, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        uint32_t b;
        std::memcpy(&b, src + i, 4);
        put32(p + 4 * i, b);
    }
}
void get_floats(unsigned char const* p, float* dst, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        uint32_t const b = get32(p + 4 * i);
        std::memcpy(dst + i, &b, 4);
    }
}

}  // namespace emsph_snapshot

// Set by the engine; the snapshot functions report errors through it.
emsph_status emsph_set_error(emsph_status s, char const* msg);

using namespace emsph_snapshot;

extern "C" {

size_t emsph_snapshot_bytes(size_t count, uint32_t flags) {
    return HEADER + 4 * payload_floats(count, flags & 3u);
}

emsph_status emsph_snapshot_encode(const float* pos_xy, const float* vel_xy,
                                   const float* colors_rgba, size_t count, int64_t time_flicks,
                                   int32_t method, void* out, size_t capacity, size_t* written) {
    if ((count && !pos_xy) || !out) {
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: null buffer");
    }
    uint32_t const flags = (vel_xy ? 1u : 0u) | (colors_rgba ? 2u : 0u);
    size_t const need = emsph_snapshot_bytes(count, flags);
    if (capacity < need) {
        return emsph_set_error(EMSPH_ERR_CAPACITY, "snapshot: output buffer too small");
    }
    auto* p = static_cast<unsigned char*>(out);
    unsigned char* q = p + HEADER;
    put_floats(q, pos_xy, 2 * count);
    q += 8 * count;
    if (vel_xy) {
        put_floats(q, vel_xy, 2 * count);
        q += 8 * count;
    }
    if (colors_rgba) { put_floats(q, colors_rgba, 4 * count); }
    std::memcpy(p, MAGIC, 8);
    put32(p + 8, VERSION);
    put32(p + 12, flags);
    put64(p + 16, count);
    put64(p + 24, static_cast<uint64_t>(time_flicks));
    put32(p + 32, static_cast<uint32_t>(method));
    put32(p + 36, 0u);
    put64(p + 40, fnv1a(p + HEADER, need - HEADER));
    if (written) { *written = need; }
    return EMSPH_OK;
}

emsph_status emsph_snapshot_info(const void* buf, size_t len, emsph_snapshot_header* info) {
    if (!buf || !info) { return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: null"); }
    auto const* p = static_cast<unsigned char const*>(buf);
    if (len < HEADER || std::memcmp(p, MAGIC, 8) != 0) {
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: not a snapshot");
    }
    if (get32(p + 8) != VERSION) {
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: unknown version");
    }
    uint32_t const flags = get32(p + 12);
    uint64_t const n = get64(p + 16);
    if ((flags & ~3u) || n > (static_cast<uint64_t>(1) << 40)) {
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: bad header");
    }
    size_t const need = emsph_snapshot_bytes(static_cast<size_t>(n), flags);
    if (len < need) { return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: truncated"); }
    if (fnv1a(p + HEADER, need - HEADER) != get64(p + 40)) {
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: checksum mismatch");
    }
    info->count = n;
    info->time_flicks = static_cast<int64_t>(get64(p + 24));
    info->method = static_cast<int32_t>(get32(p + 32));
    info->flags = flags;
    info->bytes = need;
    return EMSPH_OK;
}

emsph_status emsph_snapshot_decode(const void* buf, size_t len, float* pos_xy, float* vel_xy,
                                   float* colors_rgba) {
    emsph_snapshot_header h{};
    emsph_status const s = emsph_snapshot_info(buf, len, &h);
    if (s != EMSPH_OK) { return s; }
    size_t const n = static_cast<size_t>(h.count);
    auto const* q = static_cast<unsigned char const*>(buf) + HEADER;
    if (pos_xy) { get_floats(q, pos_xy, 2 * n); }
    q += 8 * n;
    if (h.flags & 1u) {
        if (vel_xy) { get_floats(q, vel_xy, 2 * n); }
        q += 8 * n;
    } else if (vel_xy) {
        std::memset(vel_xy, 0, 8 * n);
    }
    if (h.flags & 2u) {
        if (colors_rgba) { get_floats(q, colors_rgba, 4 * n); }
    } else if (colors_rgba) {
        std::memset(colors_rgba, 0, 16 * n);
    }
    return EMSPH_OK;
}

emsph_status emsph_snapshot_write_file(const char* path, const void* buf, size_t len) {
    if (!path || !buf) { return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: null"); }
    std::string const tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) { return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: cannot open file"); }
    size_t const w = std::fwrite(buf, 1, len, f);
    int const c = std::fclose(f);
    if (w != len || c != 0 || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return emsph_set_error(EMSPH_ERR_INVALID_ARGUMENT, "snapshot: write failed");
    }
    return EMSPH_OK;
}

}  // extern "C"

// Whole-file read for the load entry points.
bool emsph_snapshot_read_file(char const* path, std::vector<unsigned char>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { return false; }
    out.clear();
    unsigned char chunk[1 << 16];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof(chunk), f)) > 0) { out.insert(out.end(), chunk, chunk + r); }
    bool const ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}
