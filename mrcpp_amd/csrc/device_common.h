// Device helpers shared by the gfx950 kernels: error plumbing of the C ABI,
// per-kernel timing, key lookup, block scans and multiwavelet filter passes.
//
// Node layout: 8 blocks (block t = wavelet bits t, x -> bit 0), element
// i0 + N1 i1 + N1^2 i2 with N1 = k+1.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mwg.h"
#include "keys.h"

namespace mwgdev {

constexpr int kThreads = 256;
constexpr int kMaxTerms = 1000; // MRCPP's MaxSepRank (api/constants.h:36)
constexpr double kMachineZero = 1.0e-14;
// relative slack of the screening bounds against rounding
constexpr double kBoundSlack = 1.0 + 1.0e-9;

mwg_status fail(mwg_status s, const std::string &msg);
mwg_status need_init();
// Device address of the allocation-free gradient's scratch on `device`
// (current), looked up once per device.
mwg_status mlp_scratch_setup(int device);

// Per-kernel timing with HIP events on the launch stream (mwg_profile).
struct ProfileScope {
    int kind;
    hipEvent_t start{}, stop{};
    explicit ProfileScope(int k);
    ~ProfileScope();
};

// Library-owned device scratch, grown on demand from the mwg_malloc pool
// (contents are not kept when it grows).
struct Buffer {
    void *p = nullptr;
    size_t bytes = 0;
    mwg_status ensure(size_t need);
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Compute units of the current device.
int device_cus();

// A level's segments in device memory and the segment of every output node
// (valid until the next upload_level call).
struct LevelDevice {
    const mwg_segment *segs = nullptr;
    const int *node_seg = nullptr;
};
mwg_status check_level(const mwg_level *lvl);
mwg_status upload_level(const mwg_level *lvl, LevelDevice &out);
// First screen block of every segment of a level (nseg + 1 entries, the
// last the total); fails when a segment's screening table is not built or the
// level is too large for one screen launch.
mwg_status screen_block_table(const mwg_level *lvl, std::vector<int> &seg_blk);
// The level's screen block table staged by mwg_level_upload when it equals
// seg_blk (same segments), else null.
const int *staged_seg_blk(const mwg_level *lvl, const std::vector<int> &seg_blk);

// Exclusive prefix sum of in[0..n) into out[0..n], out[n] = total; in and
// out may alias.  Enqueued on the null stream.
// With n_dev, only the first min(n, *n_dev * mult) items (a length known on
// the device) are scanned, their total written after them.
mwg_status scan_exclusive(const int *in, int *out, int n, const int *n_dev = nullptr, int mult = 0);
mwg_status scan_exclusive(const long long *in, long long *out, int n);

#define MWG_HIP(call)                                                                                                   \
    do {                                                                                                                \
        hipError_t e_ = (call);                                                                                         \
        if (e_ != hipSuccess) return mwgdev::fail(MWG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

#define MWG_LAUNCHED()                                                                                                  \
    do {                                                                                                                \
        hipError_t e_ = hipGetLastError();                                                                              \
        if (e_ != hipSuccess) return mwgdev::fail(MWG_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

// Read-only view of device memory that no kernel writes while the reader
// runs: loads through it at wave-uniform addresses become scalar loads.
template <typename T> using const_as4 = const __attribute__((address_space(4))) T;
template <typename T> __device__ inline const_as4<T> *ro(const T *p) { return (const_as4<T> *)p; }
// Copy of a read-only struct into registers (dword by dword: scalar loads).
template <typename T> __device__ inline T ld(const_as4<T> &x) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized struct");
    T out;
    const auto src = (const_as4<unsigned> *)&x;
    unsigned *dst = reinterpret_cast<unsigned *>(&out);
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 4; i++) dst[i] = src[i];
    return out;
}

__device__ inline int f_lookup(const mwg_ftree &f, int depth, int x, int y, int z) {
    const long long span = 1ll << depth;
    const uint32_t ox = (uint32_t)((long long)x - f.corner[0] * span);
    const uint32_t oy = (uint32_t)((long long)y - f.corner[1] * span);
    const uint32_t oz = (uint32_t)((long long)z - f.corner[2] * span);
    const mwgkey::Key key = mwgkey::pack(depth, ox, oy, oz);
    const int log2cap = 31 - __clz(f.hash_cap);
    const uint32_t mask = (uint32_t)f.hash_cap - 1;
    uint32_t s = mwgkey::slot(key, log2cap);
    for (;;) {
        const unsigned long long hi = f.hash_keys[2 * (size_t)s];
        if (hi == 0ull) return -1;
        if (hi == key.hi && f.hash_keys[2 * (size_t)s + 1] == key.lo) return f.hash_vals[s];
        s = (s + 1) & mask;
    }
}

// Block-wide exclusive scan of one value per thread (NT threads, a multiple
// of 64): wave64 shuffles, then the wave totals through LDS (tmp holds NT/64
// entries).  Returns the block total.
template <typename T, int NT> __device__ inline T block_exclusive_scan(T v, T *tmp, T &excl) {
    static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    if (lane == 63) tmp[wave] = incl;
    __syncthreads();
    T before = 0, total = 0;
    for (int w = 0; w < NT / 64; w++) {
        const T s = tmp[w];
        if (w < wave) before += s;
        total += s;
    }
    excl = before + incl - v;
    __syncthreads();
    return total;
}

// Kernel dispatch over the supported orders k+1 in {4, 6, 8, 10}.
template <template <int> class Launcher, typename... Args> mwg_status dispatch(int kp1, Args... args) {
    switch (kp1) {
    case 4:
        return Launcher<4>::run(args...);
    case 6:
        return Launcher<6>::run(args...);
    case 8:
        return Launcher<8>::run(args...);
    case 10:
        return Launcher<10>::run(args...);
    default:
        return fail(MWG_ERR_UNSUPPORTED, "unsupported order: k+1 = " + std::to_string(kp1));
    }
}

// LDS layout of a node's 8 blocks in the transform kernels: one padding
// double after every row of N1 values, so the rows that the lanes of a pass
// read (N1 apart without padding: lanes on few banks) spread over the banks.
template <int N1> struct Lds {
    static constexpr int N3 = N1 * N1 * N1, size = 8 * N3 + 8 * N3 / N1;
};
template <int N1> __device__ __forceinline__ int pad(int o) { return o + o / N1; }

// The first `count` filter matrices (N1 x N1 each) of a filter bank into LDS:
// the passes read every filter element once per output, from LDS rather than
// through the vector memory path.
template <int N1> __device__ inline void stage_filters(double *dst, const double *filt, int count) {
    for (int i = threadIdx.x; i < count * N1 * N1; i += kThreads) dst[i] = filt[i];
}

// One pass of the separable 3D multiwavelet transform over 8 blocks in LDS
// (padded layout).
// dir_base 0: compression sub-filters, 4: reconstruction sub-filters (the
// device filter array holds H0t,H1t,G0t,G1t then H0,G0,H1,G1).  out[gt] =
// in[ft0]^T F(2 b + 0) + in[ft1]^T F(2 b + 1), ft0 < ft1 the two blocks that
// agree with gt outside direction dim, b = bit dim of gt, in viewed
// (N1 x N1^2) column-major (math_utils::apply_filter).
// Register tiles of a pass: a thread forms J = N1/2 output columns j of R = 4
// rows r (2 for N1 = 10, within the register budget), rows N2/R apart so that consecutive lanes read consecutive (padded)
// rows; every input row element and filter element is read once per tile
// instead of once per output.
template <int N1> struct PassTile {
    static constexpr int N2 = N1 * N1, J = N1 / 2, R = N1 >= 10 ? 2 : 4, RB = N2 / R, TILES = 8 * 2 * RB;
    static_assert(N1 % 2 == 0 && N2 % R == 0, "pass tile shape");
};

// s0[p][q] = sum_i in[ft0][r_q][i] F0[j_p][i], s1 likewise with ft1 / F1, the
// sum over i in increasing order (the order of MRCPP's apply_filter loops).
template <int N1>
__device__ inline void pass_tile(const double *in, const double *filt, int dir_base, int dim, int t, int &gt, int &j0, int &rb,
                                 double (&s0)[N1 / 2][PassTile<N1>::R], double (&s1)[N1 / 2][PassTile<N1>::R]) {
    using P = PassTile<N1>;
    constexpr int N2 = P::N2, N3 = N2 * N1, J = P::J, R = P::R, RB = P::RB;
    gt = t / (2 * RB);
    j0 = ((t / RB) % 2) * J;
    rb = t % RB;
    const int bit = 1 << dim;
    const int ft0 = gt & ~bit, ft1 = gt | bit;
    const double *F0 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 0) * N2 + N1 * j0;
    const double *F1 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 1) * N2 + N1 * j0;
#pragma unroll
    for (int p = 0; p < J; p++)
#pragma unroll
        for (int q = 0; q < R; q++) {
            s0[p][q] = 0.0;
            s1[p][q] = 0.0;
        }
#pragma unroll 2
    for (int i = 0; i < N1; i++) {
        double x0[R], x1[R], f0[J], f1[J];
#pragma unroll
        for (int q = 0; q < R; q++) {
            x0[q] = in[pad<N1>(ft0 * N3 + N1 * (rb + q * RB) + i)];
            x1[q] = in[pad<N1>(ft1 * N3 + N1 * (rb + q * RB) + i)];
        }
#pragma unroll
        for (int p = 0; p < J; p++) {
            f0[p] = F0[N1 * p + i];
            f1[p] = F1[N1 * p + i];
        }
#pragma unroll
        for (int p = 0; p < J; p++)
#pragma unroll
            for (int q = 0; q < R; q++) {
                s0[p][q] += x0[q] * f0[p];
                s1[p][q] += x1[q] * f1[p];
            }
    }
}

template <int N1> __device__ inline void mw_pass(double *out, const double *in, const double *filt, int dir_base, int dim) {
    using P = PassTile<N1>;
    constexpr int N2 = P::N2, N3 = N2 * N1;
    for (int t = threadIdx.x; t < P::TILES; t += kThreads) {
        double s0[P::J][P::R], s1[P::J][P::R];
        int gt, j0, rb;
        pass_tile<N1>(in, filt, dir_base, dim, t, gt, j0, rb, s0, s1);
#pragma unroll
        for (int p = 0; p < P::J; p++)
#pragma unroll
            for (int q = 0; q < P::R; q++) out[pad<N1>(gt * N3 + (j0 + p) * N2 + rb + q * P::RB)] = s0[p][q] + s1[p][q];
    }
}

// Last reconstruction pass accumulating into the children's scaling blocks,
// each of the two contributions added separately (in: padded LDS layout).
template <int N1>
__device__ inline void mw_pass_add_children(double *coef, long long child_off, long long stride, const double *in, const double *filt) {
    using P = PassTile<N1>;
    constexpr int N2 = P::N2;
    for (int t = threadIdx.x; t < P::TILES; t += kThreads) {
        double s0[P::J][P::R], s1[P::J][P::R];
        int gt, j0, rb;
        pass_tile<N1>(in, filt, 4, 2, t, gt, j0, rb, s0, s1);
#pragma unroll
        for (int p = 0; p < P::J; p++)
#pragma unroll
            for (int q = 0; q < P::R; q++) {
                double *dst = coef + child_off + (long long)gt * stride + (j0 + p) * N2 + rb + q * P::RB;
                double v = *dst;
                v += s0[p][q];
                v += s1[p][q];
                *dst = v;
            }
    }
}

// Component norms of 8 blocks, written to nrm[0..8] (8 norms + square norm);
// PAD: c in the padded LDS layout.
template <int N1, bool PAD = false> __device__ inline void block_norms(const double *c, double *nrm, double *red) {
    constexpr int N3 = N1 * N1 * N1;
    double ss[8];
#pragma unroll
    for (int t = 0; t < 8; t++) ss[t] = 0.0;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const double v = c[PAD ? pad<N1>(o) : o];
        ss[o / N3] += v * v;
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        double v = ss[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        ss[t] = v;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int t = 0; t < 8; t++) red[wave * 8 + t] = ss[t];
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int t = 0; t < 8; t++) {
            double v = 0.0;
            for (int w = 0; w < kThreads / 64; w++) v += red[w * 8 + t];
            const double nv = sqrt(v);
            nrm[t] = nv;
            tot += nv * nv;
        }
        nrm[8] = tot;
    }
    __syncthreads();
}

} // namespace mwgdev
