// Matrix-core building blocks of the band sweeps on gfx950.
//
// Every 1D sweep is a product  out[row][n] = sum_k A[row][k] B[k][n]  with
// rows = the (k+1)^2 lines of a block (padded to tiles of 16), k = the
// contracted index (padded to steps of 4), n = the two gt bits of one
// direction side by side, [O_(gbit 0) | O_(gbit 1)] (2(k+1) columns padded to
// tiles of 16).  One 16x16x4 instruction per (row tile, column tile, k step).
//
// Operand layouts (checked on the device):
//   A: lane i + 16 q holds A[i][q];  B: lane j + 16 q holds B[q][j]
//   D (f32): lane j + 16 q, register r holds D[4 q + r][j]
//   D (f64): lane j + 16 q, register r holds D[q + 4 r][j]
#pragma once

#include <hip/hip_runtime.h>

namespace mwgdev {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

// Tile counts of one sweep for N1 = k+1.
template <int N1> struct SweepShape {
    static constexpr int N2 = N1 * N1, N3 = N2 * N1;
    static constexpr int MT = (N2 + 15) / 16;     // row tiles
    static constexpr int KS = (N1 + 3) / 4;       // k steps
    static constexpr int NT = (2 * N1 + 15) / 16; // column tiles
    static constexpr int KP = 4 * KS;             // padded contraction length
    static constexpr int PM = 16 * MT + 4;        // LDS row pitch (elements)
    static constexpr int FRAG = KS * NT * 64;     // operand elements per (matrix block, f bit)
};

template <typename T> struct Mfma;

template <> struct Mfma<double> {
    using C = dvec4;
    __device__ static inline C mma(double a, double b, C c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <> struct Mfma<float> {
    using C = fvec4;
    __device__ static inline C mma(float a, float b, C c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Operand table element of (matrix block b4, f bit f, k step ks, column tile
// nt, lane): P[k][n] = O_(c = 2 (n / N1) + f)[k][n % N1] for k = 4 ks + lane/16,
// n = 16 nt + lane%16, zero in the padding.  Matrices are column-major
// O(i_f, j_g) at mats + (4 b4 + c) N1^2.
__host__ __device__ inline size_t frag_index(long long b4, int f, int ks, int nt, int lane, int KS, int NT) {
    return ((((size_t)b4 * 2 + f) * KS + ks) * NT + nt) * 64 + lane;
}

} // namespace mwgdev
