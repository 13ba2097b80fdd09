// Node keys and the open-addressing hash shared by host and device.
//
// A key is two 64-bit words: hi = (depth+1) << 32 | ox, lo = oy << 32 | oz,
// where depth = scale - root scale (0..30, MRCPP's MaxDepth, api/constants.h:33)
// and o = translation - corner * 2^depth (0 <= o < 2^32, which holds for up to
// three root boxes per direction at depth 30).  A hash slot stores its key's
// two words side by side; hi = 0 marks an empty slot.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MWG_HD __host__ __device__ inline
#else
#define MWG_HD inline
#endif

namespace mwgkey {

constexpr int kMaxDepth = 30;

struct Key {
    uint64_t hi, lo;
};

MWG_HD bool operator==(const Key &a, const Key &b) { return a.hi == b.hi && a.lo == b.lo; }

MWG_HD Key pack(int depth, uint32_t ox, uint32_t oy, uint32_t oz) {
    return Key{((uint64_t)(depth + 1) << 32) | (uint64_t)ox, ((uint64_t)oy << 32) | (uint64_t)oz};
}

// multiplicative hash of both words (top log2cap bits)
MWG_HD uint32_t slot(const Key &k, int log2cap) {
    uint64_t h = k.hi * 0x9E3779B97F4A7C15ull;
    h ^= (k.lo + (h >> 31)) * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 29;
    h *= 0x9E3779B97F4A7C15ull;
    return (uint32_t)(h >> (64 - log2cap));
}

} // namespace mwgkey
