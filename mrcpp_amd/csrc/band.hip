// Band sums of one breadth-first level (ConvolutionCalculator::calcNode,
// applyOperComp, applyOperator, tensorApplyOperComp,
// src/treebuilders/ConvolutionCalculator.cpp:224-382).
//
// Pipeline, all on one stream:
//  1. k_group_screen: one wave per candidate pair; lane t tests term c0 + t
//     against every (ft, gt) with the exact norm test and the pair's groups
//     (ft, term, gt mask) are written ft-major to a fixed-stride list.
//  2. scan of the group counts over pairs, k_unit_count per output node,
//     scan of the unit counts, k_unit_write per pair: the groups of every
//     output node are cut into work units of at most G consecutive groups
//     (G grows with the level so the partial results fit a fixed budget).
//  3. k_band_units: persistent workgroups take units from an atomic counter;
//     a unit accumulates its groups' 3-sweep tensor products in registers
//     and writes one 8-block result, the node's first unit straight into the
//     output coefficients, later units into partial slots.
//  4. k_reduce_units: per output node and chunk of coefficients, the partial
//     slots are added in unit order; k_level_norms computes the component
//     norms (MWNode::calcNorms, src/trees/MWNode.cpp:609-655).
// Every summation order is fixed by the pair and group order, so results do
// not depend on scheduling.
#include "device_common.h"

namespace mwgdev {

namespace {

constexpr int kGroupMinPerUnit = 16;
constexpr int kMaxUnitsPerNode = 512;
constexpr double kPartialBudget = 1024.0 * 1024.0 * 1024.0; // bytes of partial slots

__device__ inline unsigned group_code(int term, int ft, unsigned mask) { return (unsigned)term | ((unsigned)ft << 16) | (mask << 24); }

// ---------------------------------------------------------------------------
// 1. exact screening, one wave per pair
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_group_screen(mwg_level lvl, mwg_ftree f, mwg_band_tables op, mwg_pairs pr, unsigned *codes, int stride,
                                                     int *gcount, unsigned long long *n_applied) {
    __shared__ unsigned char smask[8 * kMaxTerms];
    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    const int g = pr.pair_g[p];
    const int fid = pr.pair_f[p];
    const int d = lvl.o_depth;
    const int M = op.M;
    if (fid < 0) { // unresolved band node: no contribution (the host generates every node first)
        if (lane == 0) gcount[p] = 0;
        return;
    }
    int dl[3], maxdl = 0;
    for (int i = 0; i < 3; i++) {
        dl[i] = pr.pair_l[3 * p + i] - lvl.l[3 * g + i];
        maxdl = max(maxdl, abs(dl[i]));
    }
    double fn[8];
    for (int t = 0; t < 8; t++) fn[t] = f.norms[(size_t)fid * 9 + t];
    const auto wd = ro(op.widths + (size_t)d * M * 5);
    const auto bsz = ro(op.band_size + (size_t)d * M * 64);
    const auto ms = ro(op.mat_start + (size_t)d * M);
    const auto mn = ro(op.mat_norm);
    unsigned napp = 0;
    for (int c0 = 0; c0 < M; c0 += 64) {
        const int i = c0 + lane;
        unsigned m8[8];
#pragma unroll
        for (int ft = 0; ft < 8; ft++) m8[ft] = 0;
        if (i < M && maxdl <= wd[i * 5 + 4]) {
            const auto w = wd + i * 5;
            const int w4 = w[4];
            double nrm[3][4];
#pragma unroll
            for (int dim = 0; dim < 3; dim++)
#pragma unroll
                for (int c = 0; c < 4; c++) nrm[dim][c] = (abs(dl[dim]) <= w[c]) ? mn[ms[i] + (long long)(dl[dim] + w4) * 4 + c] : -1.0;
#pragma unroll
            for (int ft = 0; ft < 8; ft++) {
                if (fn[ft] < kMachineZero) continue;
                unsigned m = 0;
#pragma unroll
                for (int gt = 0; gt < 8; gt++) {
                    if (!(d == 0 || gt != 0 || ft != 0)) continue;
                    const double nx = nrm[0][2 * (gt & 1) + (ft & 1)];
                    const double ny = nrm[1][2 * ((gt >> 1) & 1) + ((ft >> 1) & 1)];
                    const double nz = nrm[2][2 * ((gt >> 2) & 1) + ((ft >> 2) & 1)];
                    if (nx < 0.0 || ny < 0.0 || nz < 0.0) continue;
                    double onorm = 1.0;
                    onorm *= nx;
                    onorm *= ny;
                    onorm *= nz;
                    const double fthr = bsz[i * 64 + gt * 8 + ft] * fn[ft];
                    if (onorm * fthr > lvl.g_thr) m |= 1u << gt;
                }
                m8[ft] = m;
                napp += __popc(m);
            }
        }
        if (i < M)
#pragma unroll
            for (int ft = 0; ft < 8; ft++) smask[ft * M + i] = (unsigned char)m8[ft];
    }
    __syncthreads();
    // emit ft-major, terms ascending
    int base = 0;
    unsigned *dst = codes + (size_t)p * stride;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int ft = 0; ft < 8; ft++)
        for (int c0 = 0; c0 < M; c0 += 64) {
            const int i = c0 + lane;
            const unsigned m = (i < M) ? smask[ft * M + i] : 0u;
            const unsigned long long bal = __ballot(m != 0);
            if (m) dst[base + __popcll(bal & below)] = group_code(i, ft, m);
            base += __popcll(bal);
        }
    for (int off = 32; off > 0; off >>= 1) napp += __shfl_xor(napp, off);
    if (lane == 0) {
        gcount[p] = base;
        if (napp) atomicAdd(n_applied, (unsigned long long)napp);
    }
}

// ---------------------------------------------------------------------------
// 2. work units
// ---------------------------------------------------------------------------
// Groups per unit of an output node with T groups, total groups in the level:
// at least kGroupMinPerUnit, few enough partial slots for the budget, at most
// kMaxUnitsPerNode units per node (bounds the node's reduction).
__device__ inline int units_group_size(int total, int slots, int T) {
    const int G = max(kGroupMinPerUnit, (int)(((long long)total + slots - 1) / slots));
    return max(G, (T + kMaxUnitsPerNode - 1) / kMaxUnitsPerNode);
}

// Units per output node, packed with a nonzero flag so one scan yields both
// the first unit and the first partial slot of every node.
__global__ __launch_bounds__(256) void k_unit_count(int L, mwg_pairs pr, int npairs, const int *gstart, int slots, long long *packed, int *counter) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g == 0) *counter = 0;
    if (g >= L) return;
    const int gf = pr.g_first[g], np = pr.g_npairs[g];
    const int T = gstart[gf + np] - gstart[gf];
    const int G = units_group_size(gstart[npairs], slots, T);
    const int nu = (T + G - 1) / G;
    packed[g] = (long long)nu + ((long long)(nu > 0) << 32);
}

__device__ inline int unit_first(long long v) { return (int)(v & 0xffffffffll); }
__device__ inline int nonzero_before(long long v) { return (int)(v >> 32); }

// Descriptors {pair, first group in the pair, groups, partial slot or -1} of
// the units that start inside each pair.
__global__ __launch_bounds__(256) void k_unit_write(int npairs, mwg_pairs pr, const int *gstart, const long long *uscan, int slots, int4 *units) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npairs) return;
    const int g = pr.pair_g[p];
    const int gf = pr.g_first[g], np = pr.g_npairs[g];
    const int B = gstart[gf];
    const int T = gstart[gf + np] - B;
    const int G = units_group_size(gstart[npairs], slots, T);
    const int a = gstart[p] - B, e = gstart[p + 1] - B;
    if (a == e) return;
    const long long v = uscan[g];
    const int Ug = unit_first(v);
    const int Sg = Ug - nonzero_before(v);
    for (int j = (a + G - 1) / G; j * G < e; j++)
        units[Ug + j] = make_int4(p, j * G - a, min(G, T - j * G), j == 0 ? -1 : Sg + j - 1);
}

// ---------------------------------------------------------------------------
// 3. persistent unit kernel
//
// A workgroup of NT threads (one wave for (k+1)^2 <= 64, two for k = 9) owns
// one line of every block per thread: in each 1D sweep a thread reads N1
// values, does N1^2 FMAs with wave-uniform operator entries (scalar loads)
// and writes N1 values.  LDS arrays are [contracted index][line] with a row
// pitch chosen so the strided writes of one sweep spread over the banks.
// ---------------------------------------------------------------------------
template <int N1> struct UnitShape {
    static constexpr int N2 = N1 * N1;
    static constexpr int NT = N2 <= 64 ? 64 : 128;
    static constexpr int P = N1 == 8 ? 68 : N2 + 2;
};

template <int N1>
__global__ __launch_bounds__(UnitShape<N1>::NT) void k_band_units(mwg_level lvl, mwg_ftree f, mwg_band_tables op, mwg_pairs pr, const unsigned *codes,
                                                                  int stride, const int *gstart, const int4 *units, const long long *uscan, int L,
                                                                  double *partial, int *counter) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    constexpr int P = UnitShape<N1>::P;
    __shared__ double sF[N1 * P];
    __shared__ double sT1[N1 * P];
    __shared__ double sT2[N1 * P];
    __shared__ int s_unit;

    const int tid = threadIdx.x;
    const bool owner = tid < N2;
    const int d = lvl.o_depth;
    const int M = op.M;
    // tables, pairs and units are not written by this kernel: scalar loads
    const auto wd = ro(op.widths + (size_t)d * M * 5);
    const auto ms = ro(op.mat_start + (size_t)d * M);
    const auto mats = ro(op.mats);
    const auto lv = ro(lvl.l);
    const auto pg = ro(pr.pair_g);
    const auto pf = ro(pr.pair_f);
    const auto pl = ro(pr.pair_l);
    const auto gs = ro(gstart);
    const auto cd = ro(codes);
    const auto coff = ro(f.coef_off);
    const int total = unit_first(ro(uscan)[L]);

    for (;;) {
        __syncthreads();
        if (tid == 0) s_unit = atomicAdd(counter, 1);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        if (u >= total) break;
        const auto ud = ro(reinterpret_cast<const int *>(units)) + 4 * (size_t)u;
        int p = ud[0], k = ud[1];
        const int n = ud[2], slot = ud[3];
        const int g = pg[p];
        const int gl0 = lv[3 * g], gl1 = lv[3 * g + 1], gl2 = lv[3 * g + 2];

        double acc[8][N1];
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int j = 0; j < N1; j++) acc[t][j] = 0.0;

        int pend = gs[p + 1] - gs[p];
        int loaded = -1, curft = -1;
        int dl0 = 0, dl1 = 0, dl2 = 0;
        const double *fsrc = nullptr;
        for (int it = 0; it < n; it++) {
            while (k >= pend) { // next pair of the same output node that has groups
                p++;
                k = 0;
                pend = gs[p + 1] - gs[p];
            }
            if (p != loaded) { // a pair with groups always has a band node id
                dl0 = pl[3 * p] - gl0;
                dl1 = pl[3 * p + 1] - gl1;
                dl2 = pl[3 * p + 2] - gl2;
                fsrc = f.coef + coff[pf[p]];
                loaded = p;
                curft = -1;
            }
            const unsigned code = cd[(size_t)p * stride + k];
            k++;
            const int ti = (int)(code & 0xffffu);
            const int ft = (int)((code >> 16) & 7u);
            const unsigned gm = code >> 24;
            if (ft != curft) {
                __syncthreads();
                for (int e = tid; e < N3; e += UnitShape<N1>::NT) sF[(e % N1) * P + e / N1] = fsrc[(size_t)ft * N3 + e];
                curft = ft;
                __syncthreads();
            }
            const int w4 = wd[ti * 5 + 4];
            const long long mbase = ms[ti];
            unsigned xneed = 0, yneed = 0;
#pragma unroll
            for (int gt = 0; gt < 8; gt++)
                if (gm >> gt & 1) {
                    xneed |= 1u << (gt & 1);
                    yneed |= 1u << (gt & 3);
                }
#pragma unroll
            for (int gx = 0; gx < 2; gx++) {
                if (!(xneed >> gx & 1)) continue;
                // x sweep: T1[i1][i2 + N1 j0] = sum_i0 F[i0][i1 + N1 i2] Ox(i0, j0)
                const auto Ox = mats + (size_t)(mbase + (long long)(dl0 + w4) * 4 + 2 * gx + (ft & 1)) * N2;
                if (owner) {
                    double x[N1];
#pragma unroll
                    for (int a = 0; a < N1; a++) x[a] = sF[a * P + tid];
                    const int i1 = tid % N1, i2 = tid / N1;
#pragma unroll
                    for (int j = 0; j < N1; j++) {
                        double s = 0.0;
#pragma unroll
                        for (int a = 0; a < N1; a++) s += x[a] * Ox[a + N1 * j];
                        sT1[i1 * P + i2 + N1 * j] = s;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int gy = 0; gy < 2; gy++) {
                    if (!(yneed >> (gx + 2 * gy) & 1)) continue;
                    // y sweep: T2[i2][j0 + N1 j1] = sum_i1 T1[i1][i2 + N1 j0] Oy(i1, j1)
                    const auto Oy = mats + (size_t)(mbase + (long long)(dl1 + w4) * 4 + 2 * gy + ((ft >> 1) & 1)) * N2;
                    if (owner) {
                        double y[N1];
#pragma unroll
                        for (int a = 0; a < N1; a++) y[a] = sT1[a * P + tid];
                        const int i2 = tid % N1, j0 = tid / N1;
#pragma unroll
                        for (int j = 0; j < N1; j++) {
                            double s = 0.0;
#pragma unroll
                            for (int a = 0; a < N1; a++) s += y[a] * Oy[a + N1 * j];
                            sT2[i2 * P + j0 + N1 * j] = s;
                        }
                    }
                    __syncthreads();
                    // z sweep into the owned line (j0, j1): G[gt][j0 + N1 j1 + N2 j2] += sum_i2 T2[i2][line] Oz(i2, j2)
                    if (owner) {
                        double z[N1];
#pragma unroll
                        for (int a = 0; a < N1; a++) z[a] = sT2[a * P + tid];
#pragma unroll
                        for (int gz = 0; gz < 2; gz++) {
                            const int gt = gx + 2 * gy + 4 * gz;
                            if (!(gm >> gt & 1)) continue;
                            const auto Oz = mats + (size_t)(mbase + (long long)(dl2 + w4) * 4 + 2 * gz + ((ft >> 2) & 1)) * N2;
#pragma unroll
                            for (int j = 0; j < N1; j++) {
                                double s = 0.0;
#pragma unroll
                                for (int a = 0; a < N1; a++) s += z[a] * Oz[a + N1 * j];
                                acc[gt][j] += s;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
        }
        double *dst = (slot < 0) ? lvl.g_coef + ro(lvl.g_off)[g] : partial + (size_t)slot * 8 * N3;
        if (owner) {
#pragma unroll
            for (int t = 0; t < 8; t++)
#pragma unroll
                for (int j = 0; j < N1; j++) dst[t * N3 + tid + N2 * j] = acc[t][j];
        }
    }
}

// ---------------------------------------------------------------------------
// 4. per output node and chunk of 256 coefficients: first unit + partial
//    slots in unit order; then the component norms per node
// ---------------------------------------------------------------------------
template <int N1> __global__ __launch_bounds__(256) void k_reduce_units(mwg_level lvl, const long long *uscan, const double *partial) {
    constexpr int N3 = N1 * N1 * N1;
    const int g = blockIdx.x;
    const int o = blockIdx.y * 256 + threadIdx.x;
    const long long v = uscan[g];
    const int Ug = unit_first(v);
    const int nu = unit_first(uscan[g + 1]) - Ug;
    if (nu == 1 || o >= 8 * N3) return; // a single unit wrote the node already
    double *dst = lvl.g_coef + lvl.g_off[g];
    if (nu == 0) {
        dst[o] = 0.0;
        return;
    }
    const int Sg = Ug - nonzero_before(v);
    double acc = dst[o];
    const double *src = partial + (size_t)Sg * 8 * N3 + o;
    for (int j = 0; j + 1 < nu; j++) acc += src[(size_t)j * 8 * N3];
    dst[o] = acc;
}

template <int N1> __global__ __launch_bounds__(kThreads) void k_level_norms(mwg_level lvl) {
    __shared__ double red[8 * (kThreads / 64)];
    const int g = blockIdx.x;
    block_norms<N1>(lvl.g_coef + lvl.g_off[g], lvl.g_norms + (size_t)g * 9, red);
}

// A level without candidate pairs: zero coefficients and norms.
template <int N1> __global__ __launch_bounds__(kThreads) void k_zero_level(mwg_level lvl) {
    constexpr int N3 = N1 * N1 * N1;
    const int g = blockIdx.x;
    double *dst = lvl.g_coef + lvl.g_off[g];
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[o] = 0.0;
    if (threadIdx.x < 9) lvl.g_norms[(size_t)g * 9 + threadIdx.x] = 0.0;
}

struct BandWork {
    Buffer codes, gstart, uscan, units, partial, counter;
};
BandWork &work() {
    static BandWork w;
    return w;
}

template <int N1> int unit_grid() {
    static int grid = 0;
    if (!grid) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_band_units<N1>, UnitShape<N1>::NT, 0) != hipSuccess || per_cu <= 0) {
            (void)hipGetLastError();
            per_cu = 4;
        }
        grid = per_cu * device_cus();
    }
    return grid;
}

template <int N1> struct BandL {
    static mwg_status run(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, const mwg_pairs *pr, int npairs,
                          unsigned long long *n_applied) {
        constexpr int N3 = N1 * N1 * N1;
        const int L = lvl->count;
        if (npairs == 0) {
            ProfileScope prof_(MWG_KERNEL_REDUCE);
            hipLaunchKernelGGL(k_zero_level<N1>, dim3(L), dim3(kThreads), 0, 0, *lvl);
            MWG_LAUNCHED();
            return MWG_OK;
        }
        BandWork &w = work();
        const int stride = 8 * op->M;
        const size_t slot_bytes = sizeof(double) * 8 * N3;
        const int slots = (int)std::max<double>(256.0, kPartialBudget / (double)slot_bytes);
        if ((double)npairs * stride >= 2147483647.0) return fail(MWG_ERR_UNSUPPORTED, "too many groups in one level batch");
        if (w.codes.ensure(sizeof(unsigned) * (size_t)npairs * stride) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.gstart.ensure(sizeof(int) * ((size_t)npairs + 1)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.uscan.ensure(sizeof(long long) * ((size_t)L + 1)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.units.ensure(sizeof(int4) * ((size_t)slots + L)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.partial.ensure(slot_bytes * (size_t)slots) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.counter.ensure(sizeof(int)) != MWG_OK) return MWG_ERR_NOMEM;
        int *gstart = w.gstart.as<int>();
        long long *uscan = w.uscan.as<long long>();
        {
            ProfileScope prof_(MWG_KERNEL_GROUPS);
            hipLaunchKernelGGL(k_group_screen, dim3(npairs), dim3(64), 0, 0, *lvl, *f, *op, *pr, w.codes.as<unsigned>(), stride, gstart, n_applied);
            MWG_LAUNCHED();
            mwg_status s = scan_exclusive(gstart, gstart, npairs);
            if (s != MWG_OK) return s;
            hipLaunchKernelGGL(k_unit_count, dim3((L + 255) / 256), dim3(256), 0, 0, L, *pr, npairs, gstart, slots, uscan, w.counter.as<int>());
            MWG_LAUNCHED();
            s = scan_exclusive(uscan, uscan, L);
            if (s != MWG_OK) return s;
            hipLaunchKernelGGL(k_unit_write, dim3((npairs + 255) / 256), dim3(256), 0, 0, npairs, *pr, gstart, uscan, slots, w.units.as<int4>());
            MWG_LAUNCHED();
        }
        {
            ProfileScope prof_(MWG_KERNEL_BAND);
            hipLaunchKernelGGL(k_band_units<N1>, dim3(unit_grid<N1>()), dim3(UnitShape<N1>::NT), 0, 0, *lvl, *f, *op, *pr, w.codes.as<unsigned>(),
                               stride, gstart, w.units.as<int4>(), uscan, L, w.partial.as<double>(), w.counter.as<int>());
            MWG_LAUNCHED();
        }
        {
            ProfileScope prof_(MWG_KERNEL_REDUCE);
            hipLaunchKernelGGL(k_reduce_units<N1>, dim3(L, (8 * N3 + 255) / 256), dim3(256), 0, 0, *lvl, uscan, w.partial.as<double>());
            MWG_LAUNCHED();
            hipLaunchKernelGGL(k_level_norms<N1>, dim3(L), dim3(kThreads), 0, 0, *lvl);
            MWG_LAUNCHED();
        }
        return MWG_OK;
    }
};

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_band_apply(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, const mwg_pairs *pairs, int npairs,
                                     unsigned long long *n_applied, int precision) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!lvl || !f || !op || !pairs || !n_applied) return fail(MWG_ERR_ARG, "null argument");
    if (npairs < 0) return fail(MWG_ERR_ARG, "negative pair count");
    if (lvl->count <= 0) return MWG_OK;
    if (precision != 0) return fail(MWG_ERR_UNSUPPORTED, "only FP64 sweeps are available");
    if (op->M > kMaxTerms) return fail(MWG_ERR_UNSUPPORTED, "separation rank above 512");
    if (op->kp1 != f->kp1 || lvl->kp1 != f->kp1) return fail(MWG_ERR_ARG, "operator, level and tree orders differ");
    if (npairs > pairs->capacity) return fail(MWG_ERR_ARG, "more pairs than the pair arrays hold");
    return dispatch<BandL>(f->kp1, lvl, f, op, pairs, npairs, n_applied);
}
