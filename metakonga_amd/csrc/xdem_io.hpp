// xdem_io.hpp — byte buffers, whole-file reads/writes and the checkpoint
// container (magic, kind, version, payload size, payload, FNV-1a hash) shared
// by the state-file entry points of the simulation and the batch.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "xdem_common.hpp"

namespace xdem {
namespace io {

constexpr char CKPT_MAGIC[8] = {'X', 'D', 'E', 'M', 'C', 'K', 'P', '1'};
constexpr uint32_t CKPT_VERSION = 1;
constexpr uint32_t CKPT_SIM = 1, CKPT_BATCH = 2;
constexpr size_t CKPT_HEAD = 8 + 4 + 4 + 8, CKPT_TAIL = 8;

inline uint64_t fnv1a(const uint8_t* p, size_t n) {
    uint64_t h = 14695981039346656037ull;  // FNV-1a 64 offset basis
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct Writer {
    std::vector<uint8_t> buf;
    void bytes(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        buf.insert(buf.end(), b, b + n);
    }
    template <class T>
    void put(const T& v) {
        bytes(&v, sizeof(T));
    }
    template <class T>
    void arr(const T* p, size_t n) {
        if (n) bytes(p, sizeof(T) * n);
    }
};

struct Reader {
    const uint8_t* p = nullptr;
    size_t n = 0, off = 0;
    bool bytes(void* dst, size_t k) {
        if (k > n - off) return false;
        if (k) std::memcpy(dst, p + off, k);
        off += k;
        return true;
    }
    template <class T>
    bool get(T& v) {
        return bytes(&v, sizeof(T));
    }
    template <class T>
    bool arr(T* dst, size_t k) {
        return bytes(dst, sizeof(T) * k);
    }
    bool done() const { return off == n; }
};

inline int32_t write_file(const char* path, const std::vector<uint8_t>& data) {
    XDEM_REQUIRE(path, "null path");
    FILE* f = std::fopen(path, "wb");
    XDEM_REQUIRE(f, "cannot open %s for writing", path);
    size_t w = data.empty() ? 0 : std::fwrite(data.data(), 1, data.size(), f);
    int c = std::fclose(f);
    XDEM_REQUIRE(w == data.size() && c == 0, "short write to %s", path);
    return XDEM_OK;
}

inline int32_t read_file(const char* path, std::vector<uint8_t>& out) {
    XDEM_REQUIRE(path, "null path");
    FILE* f = std::fopen(path, "rb");
    XDEM_REQUIRE(f, "cannot open %s", path);
    out.clear();
    uint8_t tmp[1 << 16];
    size_t r;
    while ((r = std::fread(tmp, 1, sizeof(tmp), f)) > 0) out.insert(out.end(), tmp, tmp + r);
    bool err = std::ferror(f) != 0;
    std::fclose(f);
    XDEM_REQUIRE(!err, "read error on %s", path);
    return XDEM_OK;
}

inline int32_t write_checkpoint(const char* path, uint32_t kind, const Writer& payload) {
    Writer w;
    w.bytes(CKPT_MAGIC, 8);
    w.put(kind);
    w.put(CKPT_VERSION);
    w.put((uint64_t)payload.buf.size());
    w.arr(payload.buf.data(), payload.buf.size());
    w.put(fnv1a(payload.buf.data(), payload.buf.size()));
    return write_file(path, w.buf);
}

// Reads and checks a checkpoint; on success `file` holds it and the returned
// reader spans its payload.  want_kind = 0 accepts either kind.
inline int32_t read_checkpoint(const char* path, uint32_t want_kind, std::vector<uint8_t>& file,
                               Reader& payload, uint32_t* kind_out) {
    XDEM_TRY(read_file(path, file));
    XDEM_REQUIRE(file.size() >= CKPT_HEAD + CKPT_TAIL, "%s: too short for a checkpoint", path);
    XDEM_REQUIRE(std::memcmp(file.data(), CKPT_MAGIC, 8) == 0, "%s: not a checkpoint", path);
    uint32_t kind, version;
    uint64_t size;
    std::memcpy(&kind, file.data() + 8, 4);
    std::memcpy(&version, file.data() + 12, 4);
    std::memcpy(&size, file.data() + 16, 8);
    XDEM_REQUIRE(version == CKPT_VERSION, "%s: checkpoint version %u (expected %u)", path, version,
                 CKPT_VERSION);
    XDEM_REQUIRE(kind == CKPT_SIM || kind == CKPT_BATCH, "%s: unknown checkpoint kind %u", path,
                 kind);
    XDEM_REQUIRE(want_kind == 0 || kind == want_kind, "%s: checkpoint kind %u (expected %u)", path,
                 kind, want_kind);
    XDEM_REQUIRE(size == file.size() - CKPT_HEAD - CKPT_TAIL,
                 "%s: payload size %llu does not match the file (%zu bytes): truncated?", path,
                 (unsigned long long)size, file.size());
    uint64_t h;
    std::memcpy(&h, file.data() + CKPT_HEAD + size, 8);
    XDEM_REQUIRE(h == fnv1a(file.data() + CKPT_HEAD, size), "%s: checkpoint hash mismatch", path);
    payload.p = file.data() + CKPT_HEAD;
    payload.n = size;
    payload.off = 0;
    if (kind_out) *kind_out = kind;
    return XDEM_OK;
}

}  // namespace io
}  // namespace xdem
