// Content keys of the 176-byte triangle records (include/rm_hip.h): records
// are hashed in chunks of kKeyChunk, each chunk with three mixers, over the
// vertex bytes (0..35), the normal bytes (36..95) and the material bytes
// (96..175), so an upload can tell which of the three changed. Used by the
// backend on a caller's records and by the marshal while it writes records,
// which then hands the chunk values to the upload instead of having the
// records hashed a second time; both give the same keys.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>

namespace rmkeys {

constexpr size_t kKeyChunk = size_t(1) << 15;  // records per chunk
constexpr uint64_t kMul = 0x100000001B3ULL;

// Four independent multiply-xor lanes over 8-byte words, so consecutive words
// do not wait for each other's multiply. Each words() call starts at lane 0.
struct Mix4 {
    uint64_t h[4];
    explicit Mix4(uint64_t seed) {
        for (int i = 0; i < 4; ++i) h[i] = seed + 0x9E3779B97F4A7C15ULL * uint64_t(i + 1);
    }
    void words(const unsigned char* p, size_t nwords) {
        for (size_t k = 0; k < nwords; ++k) {
            uint64_t w;
            std::memcpy(&w, p + 8 * k, 8);
            const uint64_t x = (h[k & 3] ^ w) * kMul;
            h[k & 3] = x ^ (x >> 29);
        }
    }
    uint64_t done() const { return fold(fold(fold(h[0], h[1]), h[2]), h[3]); }
    static uint64_t fold(uint64_t a, uint64_t b) {
        const uint64_t x = (a ^ b) * 0xBF58476D1CE4E5B9ULL;
        return x ^ (x >> 31);
    }
};

// The three mixers of chunk `chunk`, fed one record at a time in order.
struct RecordMix {
    Mix4 p, n, m;
    explicit RecordMix(size_t chunk) : p(0x5EED01 ^ chunk), n(0x5EED02 ^ chunk), m(0x5EED03 ^ chunk) {}
    void add(const unsigned char* r) {
        uint32_t split[2];  // bytes 32..35 (v2.z) and 36..39 (n.x)
        std::memcpy(split, r + 32, 8);
        const uint64_t w_pos = split[0], w_nrm = split[1];
        p.words(r, 4);  // bytes 0..31
        p.words(reinterpret_cast<const unsigned char*>(&w_pos), 1);
        n.words(reinterpret_cast<const unsigned char*>(&w_nrm), 1);
        n.words(r + 40, 7);   // bytes 40..95
        m.words(r + 96, 10);  // bytes 96..175
    }
};

// Per-chunk values of a record array (one entry per chunk of kKeyChunk
// records, in order): RecordMix(c)'s p / n / m done() over chunk c.
struct RecordKeys {
    const uint64_t* kp = nullptr;
    const uint64_t* kn = nullptr;
    const uint64_t* km = nullptr;
    size_t chunks = 0;
};

inline size_t key_chunks(size_t ntri) { return (ntri + kKeyChunk - 1) / kKeyChunk; }

}  // namespace rmkeys
