// Packed node keys and the open-addressing hash shared by host and device.
//
// key = (depth+1) << 57 | ox << 38 | oy << 19 | oz, where depth = scale - root
// scale and o = translation - corner * 2^depth.  Translations must satisfy
// 0 <= o < 2^19, i.e. depth <= 18 for a box of two root nodes per direction;
// deeper trees are rejected by the host before anything reaches the device.
// Key 0 marks an empty hash slot.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MWG_HD __host__ __device__ inline
#else
#define MWG_HD inline
#endif

namespace mwgkey {

constexpr int kOffsetBits = 19;
constexpr int kMaxDepth = 18;

MWG_HD uint64_t pack(int depth, int ox, int oy, int oz) {
    return ((uint64_t)(depth + 1) << 57) | ((uint64_t)ox << 38) | ((uint64_t)oy << 19) | (uint64_t)oz;
}

MWG_HD uint32_t slot(uint64_t key, int log2cap) { return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - log2cap)); }

} // namespace mwgkey
