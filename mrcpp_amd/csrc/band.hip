// Band sums of one breadth-first level (ConvolutionCalculator::calcNode,
// applyOperComp, applyOperator, tensorApplyOperComp,
// src/treebuilders/ConvolutionCalculator.cpp:224-382).
//
// Pipeline, all on one stream:
//  1. k_group_screen: two waves per candidate pair; lane t tests term c0 + t
//     against every (ft, gt) with the exact norm test and the pair's groups
//     (operator block, ft, gt mask) are written ft-major to a fixed-stride
//     list; the pair's translation difference and band node go to a pair
//     record.
//  2. scan of the group counts over pairs, k_unit_count per output node,
//     scan of the unit counts, k_unit_write per pair: the groups of every
//     output node are cut into work units of at most G consecutive groups
//     (G grows with the level so the partial results fit a fixed budget), and
//     the group codes are copied into level order.
//  3. k_band_mfma (matrix cores) or k_band_units (vector ALUs): persistent
//     workgroups take units from an atomic counter; a unit accumulates its
//     groups' 3-sweep tensor products in registers and writes one 8-block
//     result, the node's first unit straight into the output coefficients,
//     later units into partial slots.
//  4. k_reduce_units: per output node and chunk of coefficients, the partial
//     slots are added in unit order; k_level_norms computes the component
//     norms (MWNode::calcNorms, src/trees/MWNode.cpp:609-655).
// Every summation order is fixed by the pair and group order, so results do
// not depend on scheduling.
#include "device_common.h"
#include "mfma.h"

namespace mwgdev {

namespace {

constexpr int kGroupMinPerUnit = 32; // default minimum groups per work unit
int g_unit_groups = kGroupMinPerUnit;  // mwg_set_band_launch
int g_band_wg_per_cu = 0;              // workgroups per CU of the persistent band kernels, 0: as many as fit
constexpr int kPrecValuBand = 3; // FP64 sweeps on the vector ALUs (comparison path)
constexpr int kMaxUnitsPerNode = 512;
constexpr double kPartialBudget = 1024.0 * 1024.0 * 1024.0; // bytes of partial slots

// Group code: bits 0-19 the operator matrix block of displacement 0 of the
// group's term at the node's depth (mat_start / 4 + width), 20-22 the f
// component, 23-30 the gt mask, 31 set when the group's band block differs
// from the previous group's of the pair (a new band node or f component).
__device__ inline unsigned group_code(long long b40, int ft, unsigned mask, bool fresh) {
    return (unsigned)b40 | ((unsigned)ft << 20) | (mask << 23) | ((unsigned)fresh << 31);
}
__device__ inline int code_block(unsigned c) { return (int)(c & 0xfffffu); }
__device__ inline int code_ft(unsigned c) { return (int)((c >> 20) & 7u); }
__device__ inline unsigned code_mask(unsigned c) { return (c >> 23) & 0xffu; }
__device__ inline bool code_fresh(unsigned c) { return (c >> 31) != 0u; }

// ---------------------------------------------------------------------------
// 1. exact screening, two waves per pair
// ---------------------------------------------------------------------------
constexpr int kScreenWaves = 2; // waves per pair: wave v tests terms 64 v, 64 v + 1, ... (stride 128)

// Screening margins (optional): ratios r_j of the node threshold below which
// an applied tuple's |O| |f_ft| is summed into margin bin j.
struct MarginSpec {
    double ratio[MWG_MARGINS];
    double *pair; // [pair][MWG_MARGINS], nullptr: margins not computed
};

__global__ __launch_bounds__(64 * kScreenWaves) void k_group_screen(const mwg_segment *segs, const int *node_seg, const int *l, const double *node_thr,
                                                                   mwg_pairs pr, unsigned *codes, int stride, int *gcount, int4 *pdl, long long *pfo,
                                                                   int *papp, MarginSpec mg) {
    __shared__ unsigned char smask[8 * kMaxTerms];
    __shared__ unsigned s_napp[kScreenWaves];
    __shared__ double s_mg[kScreenWaves][MWG_MARGINS];
    __shared__ int sb40[kMaxTerms]; // band block index of each term (first matrix / 4 + width)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x;
    const int g = pr.pair_g[p];
    const int fid = pr.pair_f[p];
    const_as4<mwg_segment> &S = ro(segs)[ro(node_seg)[g]];
    const int d = S.o_depth;
    const int M = S.op.M;
    const double g_thr = node_thr ? node_thr[g] : S.g_thr;
    if (fid < 0) { // unresolved band node: no contribution (the host generates every node first)
        if (threadIdx.x == 0) {
            gcount[p] = 0;
            papp[p] = 0;
        }
        if (mg.pair && threadIdx.x < MWG_MARGINS) mg.pair[MWG_MARGINS * (size_t)p + threadIdx.x] = 0.0;
        return;
    }
    double mg_thr[MWG_MARGINS], mg_sum[MWG_MARGINS];
#pragma unroll
    for (int j = 0; j < MWG_MARGINS; j++) {
        mg_thr[j] = mg.ratio[j] * g_thr;
        mg_sum[j] = 0.0;
    }
    int dl[3], maxdl = 0;
    for (int i = 0; i < 3; i++) {
        dl[i] = pr.pair_l[3 * p + i] - l[3 * g + i];
        maxdl = max(maxdl, abs(dl[i]));
    }
    if (threadIdx.x == 0) { // the pair's record for the sweeps
        pdl[p] = make_int4(dl[0], dl[1], dl[2], fid);
        pfo[p] = S.f.coef_off[fid];
    }
    double fn[8];
    const double *fnorms = S.f.norms;
    for (int t = 0; t < 8; t++) fn[t] = fnorms[(size_t)fid * 9 + t];
    const int *wd = S.op.widths + (size_t)d * M * 5;
    const int *bsm = S.op.band_size_max + (size_t)d * M;
    const long long *ms = S.op.mat_start + (size_t)d * M;
    const double *mn = S.op.mat_norm;
    unsigned napp = 0;
    for (int c0 = 64 * wv; c0 < M; c0 += 64 * kScreenWaves) {
        const int i = c0 + lane;
        unsigned m8[8];
#pragma unroll
        for (int ft = 0; ft < 8; ft++) m8[ft] = 0;
        int w[5] = {-1, -1, -1, -1, -1};
        long long m0 = 0;
        double bmax = 0.0;
        if (i < M) {
#pragma unroll
            for (int c = 0; c < 5; c++) w[c] = wd[i * 5 + c];
            m0 = ms[i];
            bmax = (double)bsm[i];
            sb40[i] = (int)(m0 / 4 + w[4]);
        }
        if (maxdl <= w[4]) { // term i reaches the pair (so i < M and its matrices exist)
            // the 12 component norms at the pair's translation, all loads
            // issued before any is used (|dl| <= width: in the term's block)
            double nrm[3][4];
#pragma unroll
            for (int dim = 0; dim < 3; dim++)
#pragma unroll
                for (int c = 0; c < 4; c++) nrm[dim][c] = mn[m0 + (long long)(dl[dim] + w[4]) * 4 + c];
#pragma unroll
            for (int dim = 0; dim < 3; dim++)
#pragma unroll
                for (int c = 0; c < 4; c++)
                    if (abs(dl[dim]) > w[c]) nrm[dim][c] = -1.0;
            // band-size factor of component c in one direction (a negative
            // width drops its factor): band_size[gt][ft] = 64 prod_dim bf(c_dim)
            int bf[4];
#pragma unroll
            for (int c = 0; c < 4; c++) bf[c] = w[c] < 0 ? 1 : 2 * w[c] + 1;
#pragma unroll
            for (int ft = 0; ft < 8; ft++) {
                if (fn[ft] < kMachineZero) continue;
                // bound over the g components of (term, ft): the row maximum
                // of the band size times the largest component norm per
                // direction (at depth > 0 the (0,0) tuple is skipped, so some
                // direction carries the g bit); the slack covers the
                // reassociated roundings of the exact test below
                double g1[3], gm[3];
#pragma unroll
                for (int dim = 0; dim < 3; dim++) {
                    const int fb = (ft >> dim) & 1;
                    g1[dim] = fmax(nrm[dim][2 + fb], 0.0);
                    gm[dim] = fmax(fmax(nrm[dim][fb], 0.0), g1[dim]);
                }
                const double prod = (d > 0 && ft == 0) ? fmax(g1[0] * gm[1] * gm[2], fmax(gm[0] * g1[1] * gm[2], gm[0] * gm[1] * g1[2]))
                                                       : gm[0] * gm[1] * gm[2];
                if (!(bmax * prod * fn[ft] * kBoundSlack > g_thr)) continue;
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
                    const int bs = 64 * bf[2 * (gt & 1) + (ft & 1)] * bf[2 * ((gt >> 1) & 1) + ((ft >> 1) & 1)] *
                                   bf[2 * ((gt >> 2) & 1) + ((ft >> 2) & 1)];
                    const double fthr = bs * fn[ft];
                    const double val = onorm * fthr;
                    if (val > g_thr) {
                        m |= 1u << gt;
                        if (mg.pair) {
                            const double c = onorm * fn[ft];
#pragma unroll
                            for (int j = 0; j < MWG_MARGINS; j++)
                                if (val <= mg_thr[j]) mg_sum[j] += c;
                        }
                    }
                }
                m8[ft] = m;
                napp += __popc(m);
            }
        }
        if (i < M)
#pragma unroll
            for (int ft = 0; ft < 8; ft++) smask[ft * M + i] = (unsigned char)m8[ft];
    }
    // applied tuples of the pair (added to the segment's counter per output
    // node by k_unit_count: one atomic per node instead of one per wave)
    for (int off = 32; off > 0; off >>= 1) napp += __shfl_xor(napp, off);
    if (lane == 0) s_napp[wv] = napp;
    if (mg.pair) { // fixed-order tree over lanes, then waves in order
#pragma unroll
        for (int j = 0; j < MWG_MARGINS; j++) {
            double x = mg_sum[j];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
            if (lane == 0) s_mg[wv][j] = x;
        }
    }
    __syncthreads();
    if (wv != 0) return;
    if (lane == 0) {
        unsigned tot = 0;
        for (int v = 0; v < kScreenWaves; v++) tot += s_napp[v];
        papp[p] = (int)tot;
    }
    if (mg.pair && lane < MWG_MARGINS) {
        double x = 0.0;
        for (int v = 0; v < kScreenWaves; v++) x += s_mg[v][lane];
        mg.pair[MWG_MARGINS * (size_t)p + lane] = x;
    }
    // emit ft-major, terms ascending
    int base = 0;
    unsigned *dst = codes + (size_t)p * stride;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int ft = 0; ft < 8; ft++) {
        bool first = true; // the first group of f component ft starts a new band block
        for (int c0 = 0; c0 < M; c0 += 64) {
            const int i = c0 + lane;
            const unsigned m = (i < M) ? smask[ft * M + i] : 0u;
            const unsigned long long bal = __ballot(m != 0);
            if (m) {
                const int rank = __popcll(bal & below);
                dst[base + rank] = group_code(sb40[i], ft, m, first && rank == 0);
            }
            if (bal) first = false;
            base += __popcll(bal);
        }
    }
    if (lane == 0) gcount[p] = base;
}

// ---------------------------------------------------------------------------
// 2. work units
// ---------------------------------------------------------------------------
// Groups per unit of an output node with T groups, total groups in the level:
// at least gmin, few enough partial slots for the budget, at most
// kMaxUnitsPerNode units per node (bounds the node's reduction).
__device__ inline int units_group_size(int gmin, int total, int slots, int T) {
    const int G = max(gmin, (int)(((long long)total + slots - 1) / slots));
    return max(G, (T + kMaxUnitsPerNode - 1) / kMaxUnitsPerNode);
}

// Units per output node, packed with a nonzero flag so one scan yields both
// the first unit and the first partial slot of every node.
__global__ __launch_bounds__(256) void k_unit_count(int L, mwg_pairs pr, int npairs, const int *gstart, int gmin, int slots, long long *packed, int *counter,
                                                    const mwg_segment *segs, const int *node_seg, const int *papp, const double *pmg, double *g_margin) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g == 0) *counter = 0;
    // applied tuples and candidate pairs of the node's pairs, added to its
    // segment's counters: one atomic per wave when the wave's nodes share a
    // segment (lane 0 holds the wave's first node)
    const bool live = g < L;
    const int gf = live ? pr.g_first[g] : 0, np = live ? pr.g_npairs[g] : 0;
    const int s = live ? node_seg[g] : -1;
    unsigned long long app = 0, npl = (unsigned long long)np;
    for (int q = gf; q < gf + np; q++) app += (unsigned)papp[q];
    const int s0 = __shfl(s, 0);
    if (__all(!live || s == s0)) {
        for (int off = 32; off > 0; off >>= 1) {
            app += __shfl_xor(app, off);
            npl += __shfl_xor(npl, off);
        }
        if ((threadIdx.x & 63) == 0 && s0 >= 0) {
            if (app) atomicAdd(segs[s0].counters, app);
            if (npl) atomicAdd(segs[s0].counters + 1, npl);
        }
    } else if (live) {
        if (app) atomicAdd(segs[s].counters, app);
        if (np) atomicAdd(segs[s].counters + 1, npl);
    }
    if (!live) return;
    if (g_margin) { // the node's screening margins, its pairs in order
        double m[MWG_MARGINS];
#pragma unroll
        for (int j = 0; j < MWG_MARGINS; j++) m[j] = 0.0;
        for (int q = gf; q < gf + np; q++)
#pragma unroll
            for (int j = 0; j < MWG_MARGINS; j++) m[j] += pmg[MWG_MARGINS * (size_t)q + j];
#pragma unroll
        for (int j = 0; j < MWG_MARGINS; j++) g_margin[MWG_MARGINS * (size_t)g + j] = m[j];
    }
    const int T = gstart[gf + np] - gstart[gf];
    const int G = units_group_size(gmin, gstart[npairs], slots, T);
    const int nu = (T + G - 1) / G;
    packed[g] = (long long)nu + ((long long)(nu > 0) << 32);
}

__device__ inline int unit_first(long long v) { return (int)(v & 0xffffffffll); }
__device__ inline int nonzero_before(long long v) { return (int)(v >> 32); }

// One wave per pair: the descriptors {pair, first group in the pair, groups,
// partial slot or -1} of the units that start inside the pair and, when grec
// is given, the pair's group codes in level order (gstart numbering) with the
// pair index.
__global__ __launch_bounds__(64) void k_unit_write(int npairs, mwg_pairs pr, const int *gstart, const long long *uscan, int gmin, int slots, int4 *units,
                                                   const unsigned *codes, int stride, int2 *grec) {
    const int p = blockIdx.x;
    const int base = gstart[p], cnt = gstart[p + 1] - base;
    if (cnt == 0) return;
    if (grec)
        for (int k = threadIdx.x; k < cnt; k += 64) grec[base + k] = make_int2((int)codes[(size_t)p * stride + k], p);
    if (threadIdx.x != 0) return;
    const int g = pr.pair_g[p];
    const int gf = pr.g_first[g], np = pr.g_npairs[g];
    const int B = gstart[gf];
    const int T = gstart[gf + np] - B;
    const int G = units_group_size(gmin, gstart[npairs], slots, T);
    const int a = base - B, e = a + cnt;
    const long long v = uscan[g];
    const int Ug = unit_first(v);
    const int Sg = Ug - nonzero_before(v);
    for (int j = (a + G - 1) / G; j * G < e; j++)
        units[Ug + j] = make_int4(p, j * G - a, min(G, T - j * G), j == 0 ? -1 : Sg + j - 1);
}

// ---------------------------------------------------------------------------
// 3. persistent unit kernels
//
// Workgroups take units from an atomic counter until every unit is taken; a
// unit walks its groups in order (UnitCursor), accumulates the tensor
// products of every group in registers and writes one 8-block result.
// ---------------------------------------------------------------------------

// Read-only views a unit kernel walks (nothing here is written by the band
// kernels, so loads at wave-uniform addresses are scalar loads).
struct UnitView {
    const_as4<mwg_segment> *segs;
    const_as4<int> *node_seg, *lv, *pg, *pf, *pl, *gs, *unit;
    const_as4<unsigned> *cd;
    const_as4<long long> *g_off;
    int stride, total;
    __device__ UnitView(const mwg_segment *segs_, const int *node_seg_, const int *l, const long long *g_off_, const mwg_pairs &pr,
                        const unsigned *codes, int stride_, const int *gstart, const int4 *units, const long long *uscan, int L)
            : segs(ro(segs_)), node_seg(ro(node_seg_)), lv(ro(l)), pg(ro(pr.pair_g)), pf(ro(pr.pair_f)), pl(ro(pr.pair_l)), gs(ro(gstart)),
              unit(ro(reinterpret_cast<const int *>(units))), cd(ro(codes)), g_off(ro(g_off_)), stride(stride_), total(unit_first(ro(uscan)[L])) {}
};

// Next unit for the workgroup (all threads get the same index), or -1.
__device__ inline int take_unit(int *counter, int *s_unit, int total) {
    __syncthreads();
    if (threadIdx.x == 0) *s_unit = atomicAdd(counter, 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(*s_unit);
    return u < total ? u : -1;
}

// One group of a unit: term, f component, gt mask, and the band node it acts on.
struct Group {
    int b40, ft;  // operator matrix block of displacement 0, f component
    unsigned mask;
    bool fresh_f; // band node or ft differs from the previous group's
    int dl[3];    // band node translation minus output translation
    const double *fsrc;
};

// Walks the groups of one unit; holds the unit's segment tables.
struct UnitCursor {
    int g, slot, left;
    int gl[3], dl[3];
    int p, k, pend, loaded, curft;
    const double *fsrc;
    // segment of the unit's output node
    const_as4<int> *wd;       // widths at the node's depth
    const_as4<long long> *ms; // matrix starts at the node's depth
    const double *coef;
    const_as4<long long> *coff;
    double *out;              // the node's output slot

    __device__ void start(const UnitView &v, int u) {
        const auto ud = v.unit + 4 * (size_t)u;
        p = ud[0];
        k = ud[1];
        left = ud[2];
        slot = ud[3];
        g = v.pg[p];
        for (int i = 0; i < 3; i++) gl[i] = v.lv[3 * g + i];
        const_as4<mwg_segment> &S = v.segs[v.node_seg[g]];
        const int d = S.o_depth, M = S.op.M;
        wd = ro(S.op.widths + (size_t)d * M * 5);
        ms = ro(S.op.mat_start + (size_t)d * M);
        coef = S.f.coef;
        coff = ro(S.f.coef_off);
        out = S.g_coef + v.g_off[g];
        pend = v.gs[p + 1] - v.gs[p];
        loaded = -1;
        curft = -1;
        fsrc = nullptr;
    }
    __device__ bool next(const UnitView &v, Group &grp) {
        if (left == 0) return false;
        left--;
        while (k >= pend) { // next pair of the same output node that has groups
            p++;
            k = 0;
            pend = v.gs[p + 1] - v.gs[p];
        }
        bool fresh = false;
        if (p != loaded) { // a pair with groups always has a band node id
            for (int i = 0; i < 3; i++) dl[i] = v.pl[3 * p + i] - gl[i];
            fsrc = coef + coff[v.pf[p]];
            loaded = p;
            fresh = true;
        }
        const unsigned code = v.cd[(size_t)p * v.stride + k];
        k++;
        grp.b40 = code_block(code);
        grp.ft = code_ft(code);
        grp.mask = code_mask(code);
        grp.fresh_f = fresh || grp.ft != curft;
        curft = grp.ft;
        for (int i = 0; i < 3; i++) grp.dl[i] = dl[i];
        grp.fsrc = fsrc;
        return true;
    }
    // the unit's result: the output node itself (first unit) or a partial slot
    __device__ double *result(double *partial, int ncoef) const { return slot < 0 ? out : partial + (size_t)slot * ncoef; }
};

// x / y needs of a gt mask: bit gx of xneed, bit gx + 2 gy of yneed.
__device__ inline void sweep_needs(unsigned mask, unsigned &xneed, unsigned &yneed) {
    xneed = 0;
    yneed = 0;
#pragma unroll
    for (int gt = 0; gt < 8; gt++)
        if (mask >> gt & 1) {
            xneed |= 1u << (gt & 1);
            yneed |= 1u << (gt & 3);
        }
}

// ---------------------------------------------------------------------------
// 3a. FP64 on the vector ALUs: a workgroup of NT threads (one wave for
// (k+1)^2 <= 64, two for k = 9) owns one line of every block per thread: in
// each 1D sweep a thread reads N1 values, does N1^2 FMAs with wave-uniform
// operator entries (scalar loads) and writes N1 values.  LDS arrays are
// [contracted index][line] with a row pitch chosen so the strided writes of
// one sweep spread over the banks.
// ---------------------------------------------------------------------------
template <int N1> struct UnitShape {
    static constexpr int N2 = N1 * N1;
    static constexpr int NT = N2 <= 64 ? 64 : 128;
    static constexpr int P = N1 == 8 ? 68 : N2 + 2;
};

template <int N1>
__global__ __launch_bounds__(UnitShape<N1>::NT) void k_band_units(const mwg_segment *segs, const int *node_seg, const int *l, const long long *g_off,
                                                                  mwg_pairs pr, const unsigned *codes, int stride, const int *gstart, const int4 *units,
                                                                  const long long *uscan, int L, double *partial, int *counter) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    constexpr int P = UnitShape<N1>::P;
    __shared__ double sF[N1 * P];
    __shared__ double sT1[N1 * P];
    __shared__ double sT2[N1 * P];
    __shared__ int s_unit;

    const int tid = threadIdx.x;
    const bool owner = tid < N2;
    const UnitView v(segs, node_seg, l, g_off, pr, codes, stride, gstart, units, uscan, L);

    for (int u; (u = take_unit(counter, &s_unit, v.total)) >= 0;) {
        UnitCursor cur;
        cur.start(v, u);
        const auto mats = ro(v.segs[v.node_seg[cur.g]].op.mats);
        double acc[8][N1];
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int j = 0; j < N1; j++) acc[t][j] = 0.0;

        Group grp;
        while (cur.next(v, grp)) {
            const int ft = grp.ft;
            const unsigned gm = grp.mask;
            if (grp.fresh_f) {
                __syncthreads();
                for (int e = tid; e < N3; e += UnitShape<N1>::NT) sF[(e % N1) * P + e / N1] = grp.fsrc[(size_t)ft * N3 + e];
                __syncthreads();
            }
            const long long b4 = grp.b40;
            unsigned xneed, yneed;
            sweep_needs(gm, xneed, yneed);
#pragma unroll
            for (int gx = 0; gx < 2; gx++) {
                if (!(xneed >> gx & 1)) continue;
                // x sweep: T1[i1][i2 + N1 j0] = sum_i0 F[i0][i1 + N1 i2] Ox(i0, j0)
                const auto Ox = mats + (size_t)((b4 + grp.dl[0]) * 4 + 2 * gx + (ft & 1)) * N2;
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
                    const auto Oy = mats + (size_t)((b4 + grp.dl[1]) * 4 + 2 * gy + ((ft >> 1) & 1)) * N2;
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
                            const auto Oz = mats + (size_t)((b4 + grp.dl[2]) * 4 + 2 * gz + ((ft >> 2) & 1)) * N2;
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
        double *dst = cur.result(partial, 8 * N3);
        if (owner) {
#pragma unroll
            for (int t = 0; t < 8; t++)
#pragma unroll
                for (int j = 0; j < N1; j++) dst[t * N3 + tid + N2 * j] = acc[t][j];
        }
    }
}

// ---------------------------------------------------------------------------
// 3b. the sweeps on the matrix cores
//
// Four waves per workgroup.  Per group (ft, term, gt mask): the x sweep
// F^T [Ox_0 | Ox_1] (tiles shared out over the waves), the y sweep for each
// needed gx, then each wave runs the z sweep of its own (gx, gy) (rotated
// from unit to unit) into its accumulators, the columns of a gz outside the
// mask zeroed in the operand.  LDS holds F = [i0][i1 + N1 i2], T1 = [gx][i1][j0 + N1 i2],
// T2 = [gx][gy][i2][j1 + N1 j0] (layouts and swizzles in BandGeom).
//
// PREC selects the operand types:
//   FP64: 16x16x4 f64, LDS [contracted index][row], accumulate in place;
//   FP32: 16x16x4 f32, same layout, each group's product added to FP64;
//   BF16X3: 16x16x32 bf16 with each operand split into three bf16 parts
//     x = hi + mid + lo (about 27 significant bits together) and the
//     contraction concatenated in two instructions per 8-wide block,
//     [A_hi | A_hi | A_mid | A_hi] x [B_hi ; B_mid ; B_hi ; B_lo] and
//     [A_lo | A_mid | 0 | 0] x [B_hi ; B_mid ; 0 ; 0], so the six products
//     hi*hi + hi*mid + mid*hi + hi*lo + lo*hi + mid*mid (everything down to
//     2^-16 of the leading one) accumulate exactly formed in FP32; LDS in FP32
//     [row][contracted index] with 4 floats of padding after every N1 rows,
//     split into parts as it is read; each group's product added to FP64.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int N1, int PREC> struct BandGeom {
    using S = SweepShape<N1>;
    static constexpr bool kBF = PREC == MWG_PREC_BF16X3;
    static constexpr bool k64 = PREC == MWG_PREC_FP64;
    static constexpr int KB = (N1 + 7) / 8;                  // 8-wide contraction blocks (BF16X3)
    static constexpr int KSTEPS = kBF ? 2 * KB : S::KS;     // instructions (operand fragments) per tile
    static constexpr int PK = 8 * KB;                        // row length (BF16X3 layout)
    static constexpr int ROWS = 16 * S::MT;
    // Matrix-core layouts (FP64, FP32): block [contracted index kk][row m],
    // row pitch PM = 16 mod 32 elements, the rows XOR-swizzled inside their
    // aligned 32-row group so that every 16-lane (and 32-lane) group of an
    // operand read or a result write touches distinct LDS banks: reads take
    // 16 consecutive rows of one kk; result writes put the 16 columns of one
    // output row side by side (the column index is the fastest row index of
    // the next block).  FP32 swizzles with 16 and needs rows up to the next 32.
    static constexpr int ROWS_SW = k64 ? ROWS : 32 * ((S::MT + 1) / 2);
    static constexpr int PM = ROWS_SW + ((16 - ROWS_SW % 32) + 32) % 32;
    static constexpr int CORE = ((S::KP * PM + 63) / 64) * 64;
    static constexpr int BF_BLK = ROWS * PK + (ROWS / N1 + 1) * 4;
    static constexpr int BLKF = kBF ? BF_BLK : S::KP * PM;     // F
    static constexpr int B1 = kBF ? BF_BLK : CORE + (k64 ? 0 : 32);  // stride of the two T1 blocks (gx)
    static constexpr int B2 = kBF ? BF_BLK : CORE + (k64 ? 16 : 0);  // stride of the four T2 blocks (2 gx + gy)
    // BF16X3: [row][contracted index] in FP32 with 4 floats of padding after every N1 rows
    __device__ static inline int at_bf(int kk, int m) { return m * PK + (m / N1) * 4 + kk; }
    // F: kk = i0, m = i1 + N1 i2
    __device__ static inline int atF(int kk, int m) { return kBF ? at_bf(kk, m) : kk * PM + (m ^ ((2 * kk) & 15)); }
    // T1 block b = gx: kk = i1, m = j0 + N1 i2
    __device__ static inline int atT1(int b, int kk, int m) {
        if constexpr (kBF) return b * B1 + at_bf(kk, m);
        else if constexpr (k64) return b * B1 + kk * PM + (m ^ (8 * (b & 1)));
        else return b * B1 + kk * PM + (m ^ (16 * ((kk >> 2) & 1)) ^ (8 * (b & 1)));
    }
    // T2 block b = 2 gx + gy: kk = i2, m = j1 + N1 j0
    __device__ static inline int atT2(int b, int kk, int m) {
        if constexpr (kBF) return b * B2 + at_bf(kk, m);
        else if constexpr (k64) return b * B2 + kk * PM + (m ^ (8 * (b & 1)));
        else return b * B2 + kk * PM + (m ^ (16 * ((m >> 5) & 1)) ^ (8 * (b & 1)));
    }
    // operand element (kk, m) of block `blk` of array `arr` (0: F, 1: T1, 2: T2)
    __device__ static inline int at(int arr, int blk, int kk, int m) { return arr == 0 ? atF(kk, m) : arr == 1 ? atT1(blk, kk, m) : atT2(blk, kk, m); }
};

template <int PREC> struct BandTypes;
template <> struct BandTypes<MWG_PREC_FP64> {
    using L = double; // LDS element
    using B = double; // operand table element
    using MM = Mfma<double>;
};
template <> struct BandTypes<MWG_PREC_FP32> {
    using L = float;
    using B = float;
    using MM = Mfma<float>;
};
template <> struct BandTypes<MWG_PREC_BF16X3> {
    using L = float;
    using B = uint4; // 8 bf16
    using MM = Mfma<float>;
};

// One 16x16 tile of a sweep: rows 16 mt.. of block blk of LDS array ARR
// (0: F, 1: T1, 2: T2, each starting at lds) times the operand fragments
// b[0..KSTEPS) of column tile nt.
template <int N1, int PREC, int ARR>
__device__ inline typename BandTypes<PREC>::MM::C sweep_tile(const typename BandTypes<PREC>::L *lds, int blk, const typename BandTypes<PREC>::B *b,
                                                             int mt, int lane) {
    using G = BandGeom<N1, PREC>;
    using MM = typename BandTypes<PREC>::MM;
    typename MM::C c = {};
    const int li = lane & 15, lq = lane >> 4;
    if constexpr (PREC == MWG_PREC_BF16X3) {
        const int m = 16 * mt + li;
#pragma unroll
        for (int kb = 0; kb < G::KB; kb++) {
            const float *row = lds + G::at(ARR, blk, 8 * kb, m);
            const float4 v0 = *reinterpret_cast<const float4 *>(row);
            const float4 v1 = *reinterpret_cast<const float4 *>(row + 4);
            const float x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
            bf16x8 a1, a2;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const __bf16 hi = (__bf16)x[e];
                const float r1 = x[e] - (float)hi; // exact
                const __bf16 mid = (__bf16)r1;
                const __bf16 lo = (__bf16)(r1 - (float)mid);
                a1[e] = lq == 2 ? mid : hi;
                a2[e] = lq == 0 ? lo : (lq == 1 ? mid : (__bf16)0.0f);
            }
            bf16x8 b1, b2;
            const uint4 bb1 = b[2 * kb], bb2 = b[2 * kb + 1];
            __builtin_memcpy(&b1, &bb1, sizeof(b1));
            __builtin_memcpy(&b2, &bb2, sizeof(b2));
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b2, c, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < G::KSTEPS; ks++) c = MM::mma(lds[G::at(ARR, blk, 4 * ks + lq, 16 * mt + li)], b[ks], c);
    }
    return c;
}

// An empty use of a loaded value: the compiler waits for its load here.
__device__ inline void settled(double x) { asm volatile("" ::"v"(x)); }
__device__ inline void settled(float x) { asm volatile("" ::"v"(x)); }
__device__ inline void settled(const uint4 &x) {
    const unsigned a = x.x, b = x.y, c = x.z, d = x.w;
    asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d));
}

// Order 9 (N1 = 10): at least 2 waves per SIMD (the registers allowed one;
// 3 spills too much).
template <int N1, int PREC>
__global__ __launch_bounds__(256, N1 == 10 ? 2 : 1) void k_band_mfma(const mwg_segment *segs, const int *node_seg, const long long *g_off, mwg_pairs pr, const int2 *grec,
                                                   const int4 *pdl, const long long *pfo, const int *gstart, const int4 *units,
                                                   const long long *uscan, int L, double *partial, int *counter) {
    using S = SweepShape<N1>;
    using G = BandGeom<N1, PREC>;
    using LT = typename BandTypes<PREC>::L;
    using BT = typename BandTypes<PREC>::B;
    using MM = typename BandTypes<PREC>::MM;
    constexpr int N2 = S::N2, N3 = S::N3, MT = S::MT, NT = S::NT, KSTEPS = G::KSTEPS;
    constexpr int FRAG = KSTEPS * NT * 64; // operand elements per (matrix block, f bit)
    constexpr int FPT = (N3 + 255) / 256;  // band block elements per thread
    __shared__ __align__(16) LT sF[G::BLKF];
    __shared__ __align__(16) LT sT1[2 * G::B1];
    __shared__ __align__(16) LT sT2[4 * G::B2];
    __shared__ int s_unit;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15;
    const auto useg = ro(segs);
    const auto unode = ro(node_seg);
    const auto ud4 = ro(reinterpret_cast<const int *>(units));
    const int total = unit_first(ro(uscan)[L]);

    for (int i = tid; i < G::BLKF; i += 256) sF[i] = LT(0);
    for (int i = tid; i < 2 * G::B1; i += 256) sT1[i] = LT(0);
    for (int i = tid; i < 4 * G::B2; i += 256) sT2[i] = LT(0);

    for (int u; (u = take_unit(counter, &s_unit, total)) >= 0;) {
        // unit: groups [gi0, gi0 + ng) of one output node, contiguous in level order
        const int p0 = ud4[4 * (size_t)u], ng = ud4[4 * (size_t)u + 2], slot = ud4[4 * (size_t)u + 3];
        // this wave's (gx, gy) for the unit, rotated from unit to unit so that
        // every SIMD of a CU serves all four z sweeps (how often each is needed
        // depends on the gt masks)
        const int role = (w + u) & 3;
        const int gxw = role & 1, gyw = role >> 1;
        const int gi0 = ro(gstart)[p0] + ud4[4 * (size_t)u + 1];
        const int g = ro(pr.pair_g)[p0];
        const_as4<mwg_segment> &SG = useg[unode[g]];
        const BT *frag = static_cast<const BT *>(SG.op.frag[PREC]);
        const double *coef = SG.f.coef;
        double *dst = slot < 0 ? SG.g_coef + ro(g_off)[g] : partial + (size_t)slot * 8 * N3;
        dvec4 acc[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++) acc[mt][nt] = dvec4{0.0, 0.0, 0.0, 0.0};

        // The unit's group records, 64 to a chunk, lane j holding group 64 c + j:
        // its code and its pair's translation difference and band node offset.
        // Groups are read from registers (no memory access per group); chunk
        // c + 1 is requested when group 64 c becomes current.
        struct Chunk {
            int code, d0, d1, d2, flo, fhi;
        };
        auto load_chunk = [&](int c, Chunk &ch) {
            const int i = 64 * c + lane;
            if (i < ng) {
                const int2 r = grec[gi0 + i];
                const int4 d = pdl[r.y];
                const long long fo = pfo[r.y];
                ch = Chunk{r.x, d.x, d.y, d.z, (int)(unsigned)(fo & 0xffffffffll), (int)(fo >> 32)};
            }
        };
        auto take = [&](const Chunk &ch, int i, Group &q) {
            const int j = i & 63;
            const unsigned code = (unsigned)__builtin_amdgcn_readlane(ch.code, j);
            q.b40 = code_block(code);
            q.ft = code_ft(code);
            q.mask = code_mask(code);
            q.fresh_f = code_fresh(code);
            q.dl[0] = __builtin_amdgcn_readlane(ch.d0, j);
            q.dl[1] = __builtin_amdgcn_readlane(ch.d1, j);
            q.dl[2] = __builtin_amdgcn_readlane(ch.d2, j);
            const long long fo = ((long long)__builtin_amdgcn_readlane(ch.fhi, j) << 32) | (unsigned)__builtin_amdgcn_readlane(ch.flo, j);
            q.fsrc = coef + fo + (long long)q.ft * N3;
        };
        // operand fragments of a group: matrix block b40 + dl of each direction
        // (nothing here may use the loaded values: the loads of the next group
        // must stay in flight during the current group's sweeps)
        auto fetch = [&](const Group &q, BT (&fx)[NT][KSTEPS], BT (&fy)[NT][KSTEPS], BT (&fz)[NT][KSTEPS]) {
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
#pragma unroll
                for (int ks = 0; ks < KSTEPS; ks++) {
                    const int o = (ks * NT + nt) * 64 + lane;
                    fx[nt][ks] = frag[(size_t)((q.b40 + q.dl[0]) * 2 + (q.ft & 1)) * FRAG + o];
                    fy[nt][ks] = frag[(size_t)((q.b40 + q.dl[1]) * 2 + ((q.ft >> 1) & 1)) * FRAG + o];
                    fz[nt][ks] = frag[(size_t)((q.b40 + q.dl[2]) * 2 + ((q.ft >> 2) & 1)) * FRAG + o];
                }
        };

        // software pipeline: the next group's fragments are requested and, when
        // it starts a new band block, that block loaded into registers before
        // the current group's sweeps; the block goes to LDS between the current
        // group's x and y sweeps (F is free then, and the barrier after the y
        // sweep publishes it to the next x sweep)
        Chunk ca{}, cb{};
        load_chunk(0, ca);
        if (ng > 64) load_chunk(1, cb);
        using Frags = BT[NT][KSTEPS];
        Group ga, gb;
        Frags ax, ay, az, bx, by, bz; // operand fragments of two groups, used in turn
        double fpre[FPT];
        take(ca, 0, ga);
        fetch(ga, ax, ay, az);
        // F[i0][i1 + N1 i2] for element e = i0 + N1 (i1 + N1 i2)
        for (int e = tid; e < N3; e += 256) sF[G::atF(e % N1, e / N1)] = (LT)ga.fsrc[e];
        __syncthreads();
        // group i with its fragments in (fx, fy, fz); group i + 1 is read into nq
        // and its fragments requested into (nfx, nfy, nfz).  Two buffer sets
        // used in turn keep every load in flight until its group runs.
        auto step = [&](int i, Group &cur, Frags &fx, Frags &fy, Frags &fz, Group &nq, Frags &nfx, Frags &nfy, Frags &nfz) -> bool {
            // this group's fragments (requested one group ago) complete before
            // anything new is requested, so no wait below queues behind the
            // loads for the next group
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
#pragma unroll
                for (int ks = 0; ks < KSTEPS; ks++) {
                    settled(fx[nt][ks]);
                    settled(fy[nt][ks]);
                    settled(fz[nt][ks]);
                }
            // Every group requests the same loads in the same order (the last
            // one its own data again, unused), so the waits below are exact.
            const bool have_next = i + 1 < ng;
            if (have_next && ((i + 1) & 63) == 0) take(cb, i + 1, nq);
            else take(ca, have_next ? i + 1 : i, nq);
            // the band block first: it is written to LDS within this group,
            // so its loads must not queue behind the operand loads
            const bool next_f = have_next && nq.fresh_f;
#pragma unroll
            for (int k = 0; k < FPT; k++) fpre[k] = nq.fsrc[min(tid + 256 * k, N3 - 1)];
            fetch(nq, nfx, nfy, nfz);
            const unsigned gm = cur.mask;
            unsigned xneed, yneed;
            sweep_needs(gm, xneed, yneed);
            // x sweep: T1[gx][i1][j0 + N1 i2] = sum_i0 F[i0][i1 + N1 i2] Ox_gx(i0, j0)
            for (int t = w; t < MT * NT; t += 4) {
                const int mt = t / NT, nt = t % NT;
                const auto c = sweep_tile<N1, PREC, 0>(sF, 0, fx[nt], mt, lane);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int line = 16 * mt + MM::row(lane, r), nn = 16 * nt + li;
                    if (line < N2 && nn < 2 * N1) {
                        const int gx = nn / N1, j0 = nn % N1, i1 = line % N1, i2 = line / N1;
                        sT1[G::atT1(gx, i1, j0 + N1 * i2)] = c[r];
                    }
                }
            }
            __syncthreads();
            if (next_f) {
#pragma unroll
                for (int k = 0; k < FPT; k++) {
                    const int e = tid + 256 * k;
                    if (e < N3) sF[G::atF(e % N1, e / N1)] = (LT)fpre[k];
                }
            }
            // y sweep: T2[gx][gy][i2][j1 + N1 j0] = sum_i1 T1[gx][i1][j0 + N1 i2] Oy_gy(i1, j1)
            for (int t = w; t < 2 * MT * NT; t += 4) {
                const int gx = t / (MT * NT), tt = t % (MT * NT);
                if (!(xneed >> gx & 1)) continue;
                const int mt = tt / NT, nt = tt % NT;
                const auto c = sweep_tile<N1, PREC, 1>(sT1, gx, fy[nt], mt, lane);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = 16 * mt + MM::row(lane, r), nn = 16 * nt + li;
                    if (m < N2 && nn < 2 * N1) {
                        const int j0 = m % N1, i2 = m / N1, gy = nn / N1, j1 = nn % N1;
                        sT2[G::atT2(gx * 2 + gy, i2, j1 + N1 * j0)] = c[r];
                    }
                }
            }
            __syncthreads();
            // z sweep of this wave's (gx, gy): G[gt][j0 + N1 j1 + N2 j2] += sum_i2 T2[gx][gy][i2][j1 + N1 j0] Oz_gz(i2, j2)
            if (yneed >> (gxw + 2 * gyw) & 1) {
                const int zb2 = gxw * 2 + gyw;
                // the columns of a gz outside the mask zeroed
                BT zb[NT][KSTEPS];
#pragma unroll
                for (int nt = 0; nt < NT; nt++) {
                    const int nn = 16 * nt + li;
                    const bool keep = nn < 2 * N1 && ((gm >> (gxw + 2 * gyw + 4 * (nn / N1))) & 1);
#pragma unroll
                    for (int ks = 0; ks < KSTEPS; ks++) zb[nt][ks] = keep ? fz[nt][ks] : BT{};
                }
#pragma unroll
                for (int mt = 0; mt < MT; mt++)
#pragma unroll
                    for (int nt = 0; nt < NT; nt++) {
                        if constexpr (PREC == MWG_PREC_FP64) {
#pragma unroll
                            for (int ks = 0; ks < KSTEPS; ks++)
                                acc[mt][nt] = MM::mma(sT2[G::atT2(zb2, 4 * ks + (lane >> 4), 16 * mt + li)], zb[nt][ks], acc[mt][nt]);
                        } else {
                            const auto c = sweep_tile<N1, PREC, 2>(sT2, zb2, zb[nt], mt, lane);
#pragma unroll
                            for (int r = 0; r < 4; r++) acc[mt][nt][r] += (double)c[r];
                        }
                    }
            }
            if (have_next && ((i + 1) & 63) == 0) { // group i + 1 opens chunk (i + 1) / 64
                ca = cb;
                if (i + 1 + 64 < ng) load_chunk((i + 1) / 64 + 1, cb);
            }
            return have_next;
        };
        for (int i = 0;; i += 2) {
            if (!step(i, ga, ax, ay, az, gb, bx, by, bz)) break;
            if (!step(i + 1, gb, bx, by, bz, ga, ax, ay, az)) break;
        }
        // this wave's blocks gt = gx + 2 gy + 4 gz of the unit result
#pragma unroll
        for (int mt = 0; mt < MT; mt++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = 16 * mt + MM::row(lane, r), nn = 16 * nt + li;
                    if (m < N2 && nn < 2 * N1) {
                        // row m = j1 + N1 j0 of T2: output element j0 + N1 j1 + N2 j2
                        const int gz = nn / N1, j2 = nn % N1, j1 = m % N1, j0 = m / N1;
                        dst[(gxw + 2 * gyw + 4 * gz) * N3 + j0 + N1 * j1 + N2 * j2] = acc[mt][nt][r];
                    }
                }
    }
}

// ---------------------------------------------------------------------------
// 3c. order 7 (N1 = 8) on the matrix cores with wave-owned rows (FP64, FP32).
//
// The same sweeps and instructions as k_band_mfma<8, PREC>, distributed so
// that only the x -> y step crosses waves.  Wave w computes the x sweep rows
// i1 + N1 i2 with i2 in {2w, 2w+1} (it writes and reads its own F rows), and
// the y and z sweeps for the output rows with j0 in {2w, 2w+1}: the y sweep
// reads T1 rows written by every wave (one barrier per group; T1 is kept
// twice, by group parity, so the next group's x sweep does not wait for the
// slowest wave's y sweep) and writes T2 rows only this wave reads; the z
// sweep runs every needed (gx, gy) block, so the four waves do equal work
// whatever the gt mask.  Every output element goes through the same products
// in the same order as in k_band_mfma: the results are bit-identical.
//
// LDS layouts (element offsets; conflict-free for these access patterns:
// each 32-lane half of an access touches distinct banks, and for FP64 each
// 16-lane quarter too):
//   F[kk = i0][m = i1 + N1 i2], T1[gx][kk = i1][j0 + N1 i2],
//   T2 (per wave) [b = 2 gx + gy][kk = i2][r = j1 + N1 (j0 - 2w)].
// ---------------------------------------------------------------------------
template <int PREC> struct RowGeom;
template <> struct RowGeom<MWG_PREC_FP64> {
    static constexpr int F = 512, T1B = 520, T1 = 2 * T1B, T2W = 512;
    __device__ static inline int atF(int kk, int m) { return kk * 64 + (m ^ (2 * kk) ^ (16 * (kk & 1))); }
    __device__ static inline int atT1(int gx, int kk, int j0, int i2) { return gx * T1B + kk * 64 + (((j0 + 8 * i2) ^ (2 * (i2 >> 1))) ^ (16 * (kk & 1))); }
    __device__ static inline int atT2(int b, int kk, int r) {
        const int gy = b & 1;
        return b * 128 + ((kk * 16 + (r ^ (8 * gy))) ^ (16 * gy));
    }
};
template <> struct RowGeom<MWG_PREC_FP32> {
    static constexpr int F = 512, T1B = 552, T1 = 2 * T1B, T2W = 576;
    __device__ static inline int atF(int kk, int m) { return kk * 64 + (m ^ (16 * (kk & 1)) ^ (4 * (kk >> 1))); }
    __device__ static inline int atT1(int gx, int kk, int j0, int i2) { return gx * T1B + kk * 68 + ((j0 + 8 * i2) ^ (2 * (i2 >> 2))); }
    __device__ static inline int atT2(int b, int kk, int r) { return b * 144 + kk * 16 + (r ^ (8 * ((kk >> 1) & 1))); }
};

// FP32: at least 5 waves per SIMD (the registers allow 4 otherwise; a few
// spilled bytes cost less than the latency the fifth wave hides).
template <int PREC>
__global__ __launch_bounds__(256, PREC == MWG_PREC_FP32 ? 5 : 1) void k_band_rows(const mwg_segment *segs, const int *node_seg, const long long *g_off, mwg_pairs pr, const int2 *grec,
                                                   const int4 *pdl, const long long *pfo, const int *gstart, const int4 *units,
                                                   const long long *uscan, int L, double *partial, int *counter) {
    constexpr int N1 = 8, N3 = 512, KSTEPS = 2;
    constexpr int FRAG = KSTEPS * 64; // operand elements per (matrix block, f bit)
    using R = RowGeom<PREC>;
    using LT = typename BandTypes<PREC>::L;
    using BT = typename BandTypes<PREC>::B;
    using MM = typename BandTypes<PREC>::MM;
    __shared__ __align__(16) LT sF[R::F];
    __shared__ __align__(16) LT sT1[2 * R::T1];
    __shared__ __align__(16) LT sT2[4 * R::T2W];
    __shared__ int s_unit;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = lane & 15, lq = lane >> 4;
    const auto useg = ro(segs);
    const auto unode = ro(node_seg);
    const auto ud4 = ro(reinterpret_cast<const int *>(units));
    const int total = unit_first(ro(uscan)[L]);
    LT *const myT2 = sT2 + w * R::T2W;

    for (int i = tid; i < R::F; i += 256) sF[i] = LT(0);
    for (int i = tid; i < 2 * R::T1; i += 256) sT1[i] = LT(0);
    for (int i = tid; i < 4 * R::T2W; i += 256) sT2[i] = LT(0);

    // this wave's F rows: band block elements 128 w + lane + 64 k
    auto write_f = [&](const double (&v)[2]) {
#pragma unroll
        for (int k = 0; k < 2; k++) sF[R::atF(lane & 7, 16 * w + 8 * k + (lane >> 3))] = (LT)v[k];
    };

    for (int u; (u = take_unit(counter, &s_unit, total)) >= 0;) {
        // unit: groups [gi0, gi0 + ng) of one output node, contiguous in level order
        const int p0 = ud4[4 * (size_t)u], ng = ud4[4 * (size_t)u + 2], slot = ud4[4 * (size_t)u + 3];
        const int gi0 = ro(gstart)[p0] + ud4[4 * (size_t)u + 1];
        const int g = ro(pr.pair_g)[p0];
        const_as4<mwg_segment> &SG = useg[unode[g]];
        const BT *frag = static_cast<const BT *>(SG.op.frag[PREC]);
        const double *coef = SG.f.coef;
        double *dst = slot < 0 ? SG.g_coef + ro(g_off)[g] : partial + (size_t)slot * 8 * N3;
        dvec4 acc[4]; // z sweep result of block b = 2 gx + gy, rows j1 + N1 (j0 - 2w), columns gz, j2
#pragma unroll
        for (int b = 0; b < 4; b++) acc[b] = dvec4{0.0, 0.0, 0.0, 0.0};

        // group records, 64 to a chunk, lane j holding group 64 c + j (as in k_band_mfma)
        struct Chunk {
            int code, d0, d1, d2, flo, fhi;
        };
        auto load_chunk = [&](int c, Chunk &ch) {
            const int i = 64 * c + lane;
            if (i < ng) {
                const int2 r = grec[gi0 + i];
                const int4 d = pdl[r.y];
                const long long fo = pfo[r.y];
                ch = Chunk{r.x, d.x, d.y, d.z, (int)(unsigned)(fo & 0xffffffffll), (int)(fo >> 32)};
            }
        };
        auto take = [&](const Chunk &ch, int i, Group &q) {
            const int j = i & 63;
            const unsigned code = (unsigned)__builtin_amdgcn_readlane(ch.code, j);
            q.b40 = code_block(code);
            q.ft = code_ft(code);
            q.mask = code_mask(code);
            q.fresh_f = code_fresh(code);
            q.dl[0] = __builtin_amdgcn_readlane(ch.d0, j);
            q.dl[1] = __builtin_amdgcn_readlane(ch.d1, j);
            q.dl[2] = __builtin_amdgcn_readlane(ch.d2, j);
            const long long fo = ((long long)__builtin_amdgcn_readlane(ch.fhi, j) << 32) | (unsigned)__builtin_amdgcn_readlane(ch.flo, j);
            q.fsrc = coef + fo + (long long)q.ft * N3;
        };
        using Frags = BT[KSTEPS];
        auto fetch = [&](const Group &q, Frags &fx, Frags &fy, Frags &fz) {
#pragma unroll
            for (int ks = 0; ks < KSTEPS; ks++) {
                const int o = ks * 64 + lane;
                fx[ks] = frag[(size_t)((q.b40 + q.dl[0]) * 2 + (q.ft & 1)) * FRAG + o];
                fy[ks] = frag[(size_t)((q.b40 + q.dl[1]) * 2 + ((q.ft >> 1) & 1)) * FRAG + o];
                fz[ks] = frag[(size_t)((q.b40 + q.dl[2]) * 2 + ((q.ft >> 2) & 1)) * FRAG + o];
            }
        };

        Chunk ca{}, cb{};
        load_chunk(0, ca);
        if (ng > 64) load_chunk(1, cb);
        Group ga, gb;
        Frags ax, ay, az, bx, by, bz;
        double fpre[2];
        take(ca, 0, ga);
        fetch(ga, ax, ay, az);
        {
            double v[2];
#pragma unroll
            for (int k = 0; k < 2; k++) v[k] = ga.fsrc[128 * w + lane + 64 * k];
            write_f(v);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        auto step = [&](int i, Group &cur, Frags &fx, Frags &fy, Frags &fz, Group &nq, Frags &nfx, Frags &nfy, Frags &nfz) -> bool {
#pragma unroll
            for (int ks = 0; ks < KSTEPS; ks++) {
                settled(fx[ks]);
                settled(fy[ks]);
                settled(fz[ks]);
            }
            const bool have_next = i + 1 < ng;
            if (have_next && ((i + 1) & 63) == 0) take(cb, i + 1, nq);
            else take(ca, have_next ? i + 1 : i, nq);
            const bool next_f = have_next && nq.fresh_f;
#pragma unroll
            for (int k = 0; k < 2; k++) fpre[k] = nq.fsrc[128 * w + lane + 64 * k];
            fetch(nq, nfx, nfy, nfz);
            const unsigned gm = cur.mask;
            unsigned xneed, yneed;
            sweep_needs(gm, xneed, yneed);
            LT *const t1 = sT1 + (i & 1) * R::T1;
            // x sweep, rows m = i1 + N1 i2 of this wave: T1[gx][i1][j0 + N1 i2] = sum_i0 F[i0][m] Ox_gx(i0, j0)
            {
                typename MM::C c = {};
#pragma unroll
                for (int ks = 0; ks < KSTEPS; ks++) c = MM::mma(sF[R::atF(4 * ks + lq, 16 * w + li)], fx[ks], c);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rho = MM::row(lane, r);
                    t1[R::atT1(li >> 3, rho & 7, li & 7, 2 * w + (rho >> 3))] = c[r];
                }
            }
            __syncthreads();
            // y sweep, rows (j0 = 2w + (row & 1), i2 = row >> 1):
            // T2[gx][gy][i2][j1 + N1 (j0 - 2w)] = sum_i1 T1[gx][i1][j0 + N1 i2] Oy_gy(i1, j1)
#pragma unroll
            for (int gx = 0; gx < 2; gx++) {
                if (!(xneed >> gx & 1)) continue;
                typename MM::C c = {};
#pragma unroll
                for (int ks = 0; ks < KSTEPS; ks++) c = MM::mma(t1[R::atT1(gx, 4 * ks + lq, 2 * w + (li & 1), li >> 1)], fy[ks], c);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rho = MM::row(lane, r);
                    myT2[R::atT2(2 * gx + (li >> 3), rho >> 1, (li & 7) + 8 * (rho & 1))] = c[r];
                }
            }
            // T2 and F rows pass between lanes of this wave only: the wave's LDS
            // accesses complete in issue order, so a compiler fence suffices
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            // z sweep of every needed (gx, gy), rows j1 + N1 (j0 - 2w):
            // G[gt][j0 + N1 j1 + N2 j2] += sum_i2 T2[gx][gy][i2][.] Oz_gz(i2, j2), columns of a gz outside the mask zeroed
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int gx = b >> 1, gy = b & 1;
                if (!(yneed >> (gx + 2 * gy) & 1)) continue;
                const bool keep = (gm >> (gx + 2 * gy + 4 * (li >> 3))) & 1;
                BT zb[KSTEPS];
#pragma unroll
                for (int ks = 0; ks < KSTEPS; ks++) zb[ks] = keep ? fz[ks] : BT{};
                if constexpr (PREC == MWG_PREC_FP64) {
#pragma unroll
                    for (int ks = 0; ks < KSTEPS; ks++) acc[b] = MM::mma(myT2[R::atT2(b, 4 * ks + lq, li)], zb[ks], acc[b]);
                } else {
                    typename MM::C c = {};
#pragma unroll
                    for (int ks = 0; ks < KSTEPS; ks++) c = MM::mma(myT2[R::atT2(b, 4 * ks + lq, li)], zb[ks], c);
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[b][r] += (double)c[r];
                }
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            // the next band block into this wave's F rows (its x sweep above has read them)
            if (next_f) write_f(fpre);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (have_next && ((i + 1) & 63) == 0) {
                ca = cb;
                if (i + 1 + 64 < ng) load_chunk((i + 1) / 64 + 1, cb);
            }
            return have_next;
        };
        for (int i = 0;; i += 2) {
            if (!step(i, ga, ax, ay, az, gb, bx, by, bz)) break;
            if (!step(i + 1, gb, bx, by, bz, ga, ax, ay, az)) break;
        }
        // blocks gt = gx + 2 gy + 4 gz, output rows j0 in {2w, 2w + 1}
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int gx = b >> 1, gy = b & 1;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int rho = MM::row(lane, r);
                const int j1 = rho & 7, j0 = 2 * w + (rho >> 3), gz = li >> 3, j2 = li & 7;
                dst[(gx + 2 * gy + 4 * gz) * N3 + j0 + N1 * j1 + 64 * j2] = acc[b][r];
            }
        }
    }
}

// Operand table of one precision from the FP64 operator matrices.  Element
// (matrix block b4, f bit, column tile nt, step ks, lane) of P[k][n] =
// O_(2 (n / N1) + f)[k][n % N1], zero in the padding; FP64/FP32 hold
// k = 4 ks + lane/16, n = 16 nt + lane%16; BF16X3 holds 8 bf16 for
// k = 8 (ks / 2) + 0..7 in the slot lane/16 of [hi ; mid ; hi ; lo] (even ks)
// or [hi ; mid ; 0 ; 0] (odd ks).
template <int N1, int PREC> __global__ __launch_bounds__(256) void k_build_frag(const double *mats, long long nb4, typename BandTypes<PREC>::B *out) {
    using S = SweepShape<N1>;
    using G = BandGeom<N1, PREC>;
    constexpr int FRAG = G::KSTEPS * S::NT * 64;
    const size_t total = (size_t)nb4 * 2 * FRAG;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int lane = (int)(idx % 64);
    size_t rest = idx / 64;
    const int nt = (int)(rest % S::NT);
    rest /= S::NT;
    const int ks = (int)(rest % G::KSTEPS);
    rest /= G::KSTEPS;
    const int fb = (int)(rest % 2);
    const long long b4 = (long long)(rest / 2);
    const int nn = 16 * nt + lane % 16;
    auto P = [&](int kk) {
        if (kk >= N1 || nn >= 2 * N1) return 0.0;
        return mats[(size_t)(b4 * 4 + 2 * (nn / N1) + fb) * S::N2 + kk + N1 * (nn % N1)];
    };
    if constexpr (PREC == MWG_PREC_BF16X3) {
        // three-part split of the FP64 entry (residuals exact in FP64); fragment
        // 2 kb + 0 holds [hi ; mid ; hi ; lo], 2 kb + 1 holds [hi ; mid ; 0 ; 0]
        // for the 8-wide block kb, slot lane / 16
        const int slot = lane / 16, kb = ks / 2, part = ks % 2;
        bf16x8 val;
        for (int e = 0; e < 8; e++) {
            const double x = P(8 * kb + e);
            const __bf16 hi = (__bf16)(float)x;
            const double r1 = x - (double)(float)hi;
            const __bf16 mid = (__bf16)(float)r1;
            const __bf16 lo = (__bf16)(float)(r1 - (double)(float)mid);
            const __bf16 parts[2][4] = {{hi, mid, hi, lo}, {hi, mid, (__bf16)0.0f, (__bf16)0.0f}};
            val[e] = parts[part][slot];
        }
        uint4 o;
        __builtin_memcpy(&o, &val, sizeof(o));
        out[idx] = o;
    } else {
        out[idx] = (typename BandTypes<PREC>::B)P(4 * ks + lane / 16);
    }
}

// ---------------------------------------------------------------------------
// 4. per output node and chunk of 256 coefficients: first unit + partial
//    slots in unit order; then the component norms per node
// ---------------------------------------------------------------------------
// Output slot of node g (its segment's pool + g_off).
__device__ inline double *node_out(const mwg_segment *segs, const int *node_seg, const long long *g_off, int g) {
    return ro(segs)[ro(node_seg)[g]].g_coef + g_off[g];
}

template <int N1>
__global__ __launch_bounds__(256) void k_reduce_units(const mwg_segment *segs, const int *node_seg, const long long *g_off, const long long *uscan,
                                                      const double *partial) {
    constexpr int N3 = N1 * N1 * N1;
    const int g = blockIdx.x;
    const int o = blockIdx.y * 256 + threadIdx.x;
    const long long v = uscan[g];
    const int Ug = unit_first(v);
    const int nu = unit_first(uscan[g + 1]) - Ug;
    if (nu == 1 || o >= 8 * N3) return; // a single unit wrote the node already
    double *dst = node_out(segs, node_seg, g_off, g);
    if (nu == 0) {
        dst[o] = 0.0;
        return;
    }
    const int Sg = Ug - nonzero_before(v);
    double acc = dst[o];
    const double *src = partial + (size_t)Sg * 8 * N3 + o;
    // slots added in unit order; their loads issued 8 at a time
    int j = 0;
    for (; j + 8 < nu; j += 8) {
        double t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = src[(size_t)(j + q) * 8 * N3];
#pragma unroll
        for (int q = 0; q < 8; q++) acc += t[q];
    }
    for (; j + 1 < nu; j++) acc += src[(size_t)j * 8 * N3];
    dst[o] = acc;
}

template <int N1>
__global__ __launch_bounds__(kThreads) void k_level_norms(const mwg_segment *segs, const int *node_seg, const long long *g_off, double *g_norms) {
    __shared__ double red[8 * (kThreads / 64)];
    const int g = blockIdx.x;
    block_norms<N1>(node_out(segs, node_seg, g_off, g), g_norms + (size_t)g * 9, red);
}

// A level without candidate pairs: zero coefficients, norms and margins.
template <int N1>
__global__ __launch_bounds__(kThreads) void k_zero_level(const mwg_segment *segs, const int *node_seg, const long long *g_off, double *g_norms,
                                                         double *g_margin) {
    constexpr int N3 = N1 * N1 * N1;
    const int g = blockIdx.x;
    double *dst = node_out(segs, node_seg, g_off, g);
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[o] = 0.0;
    if (threadIdx.x < 9) g_norms[(size_t)g * 9 + threadIdx.x] = 0.0;
    if (g_margin && threadIdx.x < MWG_MARGINS) g_margin[(size_t)g * MWG_MARGINS + threadIdx.x] = 0.0;
}

struct BandWork {
    Buffer codes, gstart, uscan, units, partial, counter, grec, pdl, pfo, papp, pmg;
};
BandWork &work() {
    static BandWork w;
    return w;
}

// Persistent grid: as many workgroups as fit on the device at once.
template <typename K> int resident_grid(K kernel, int threads) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 2;
    }
    return per_cu * device_cus();
}
// the resident grid, or g_band_wg_per_cu workgroups per CU when that is smaller
int capped_grid(int resident) {
    if (g_band_wg_per_cu <= 0) return resident;
    return std::min(resident, g_band_wg_per_cu * device_cus());
}
template <int N1> int unit_grid() {
    static const int grid = resident_grid(k_band_units<N1>, UnitShape<N1>::NT);
    return capped_grid(grid);
}
template <int N1, int PREC> int mfma_grid() {
    static const int grid = resident_grid(k_band_mfma<N1, PREC>, 256);
    return capped_grid(grid);
}
template <int PREC> int rows_grid() {
    static const int grid = resident_grid(k_band_rows<PREC>, 256);
    return capped_grid(grid);
}
int g_band_layout = MWG_BAND_AUTO;
template <int N1, int PREC>
void launch_mfma(const mwg_level *lvl, const LevelDevice &dev, const mwg_pairs *pr, BandWork &w, int stride) {
    if constexpr (N1 == 8 && (PREC == MWG_PREC_FP64 || PREC == MWG_PREC_FP32)) {
        if (g_band_layout != MWG_BAND_TILES) {
            hipLaunchKernelGGL((k_band_rows<PREC>), dim3(rows_grid<PREC>()), dim3(256), 0, 0, dev.segs, dev.node_seg, lvl->g_off, *pr,
                               w.grec.as<int2>(), w.pdl.as<int4>(), w.pfo.as<long long>(), w.gstart.as<int>(), w.units.as<int4>(),
                               w.uscan.as<long long>(), lvl->count, w.partial.as<double>(), w.counter.as<int>());
            return;
        }
    }
    hipLaunchKernelGGL((k_band_mfma<N1, PREC>), dim3(mfma_grid<N1, PREC>()), dim3(256), 0, 0, dev.segs, dev.node_seg, lvl->g_off, *pr,
                       w.grec.as<int2>(), w.pdl.as<int4>(), w.pfo.as<long long>(), w.gstart.as<int>(), w.units.as<int4>(),
                       w.uscan.as<long long>(), lvl->count, w.partial.as<double>(), w.counter.as<int>());
}

template <int N1> struct BandL {
    static mwg_status run(const mwg_level *lvl, const mwg_pairs *pr, int npairs, int precision) {
        constexpr int N3 = N1 * N1 * N1;
        const int L = lvl->count;
        LevelDevice dev;
        mwg_status st = upload_level(lvl, dev);
        if (st != MWG_OK) return st;
        if (npairs == 0) {
            ProfileScope prof_(MWG_KERNEL_REDUCE);
            hipLaunchKernelGGL(k_zero_level<N1>, dim3(L), dim3(kThreads), 0, 0, dev.segs, dev.node_seg, lvl->g_off, lvl->g_norms, lvl->g_margin);
            MWG_LAUNCHED();
            return MWG_OK;
        }
        BandWork &w = work();
        int maxM = 1;
        for (int s = 0; s < lvl->nseg; s++) maxM = std::max(maxM, lvl->segs[s].op.M);
        const int stride = 8 * maxM; // group-list stride per pair
        const size_t slot_bytes = sizeof(double) * 8 * N3;
        const int slots = (int)std::max<double>(256.0, kPartialBudget / (double)slot_bytes);
        if ((double)npairs * stride >= 2147483647.0) return fail(MWG_ERR_UNSUPPORTED, "too many groups in one level batch");
        if (w.codes.ensure(sizeof(unsigned) * (size_t)npairs * stride) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.gstart.ensure(sizeof(int) * ((size_t)npairs + 1)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.uscan.ensure(sizeof(long long) * ((size_t)L + 1)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.units.ensure(sizeof(int4) * ((size_t)slots + L)) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.partial.ensure(slot_bytes * (size_t)slots) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.counter.ensure(sizeof(int)) != MWG_OK) return MWG_ERR_NOMEM;
        const bool mfma = precision != kPrecValuBand;
        if (mfma && w.grec.ensure(sizeof(int2) * (size_t)npairs * stride) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.pdl.ensure(sizeof(int4) * (size_t)npairs) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.pfo.ensure(sizeof(long long) * (size_t)npairs) != MWG_OK) return MWG_ERR_NOMEM;
        if (w.papp.ensure(sizeof(int) * (size_t)npairs) != MWG_OK) return MWG_ERR_NOMEM;
        MarginSpec mg{};
        if (lvl->g_margin) {
            if (w.pmg.ensure(sizeof(double) * MWG_MARGINS * (size_t)npairs) != MWG_OK) return MWG_ERR_NOMEM;
            for (int j = 0; j < MWG_MARGINS; j++) mg.ratio[j] = lvl->margin_ratio[j];
            mg.pair = w.pmg.as<double>();
        }
        int *gstart = w.gstart.as<int>();
        long long *uscan = w.uscan.as<long long>();
        {
            ProfileScope prof_(MWG_KERNEL_GROUPS);
            hipLaunchKernelGGL(k_group_screen, dim3(npairs), dim3(64 * kScreenWaves), 0, 0, dev.segs, dev.node_seg, lvl->l, lvl->node_thr, *pr,
                               w.codes.as<unsigned>(), stride, gstart, w.pdl.as<int4>(), w.pfo.as<long long>(), w.papp.as<int>(), mg);
            MWG_LAUNCHED();
            mwg_status s = scan_exclusive(gstart, gstart, npairs);
            if (s != MWG_OK) return s;
            hipLaunchKernelGGL(k_unit_count, dim3((L + 255) / 256), dim3(256), 0, 0, L, *pr, npairs, gstart, g_unit_groups, slots, uscan, w.counter.as<int>(),
                               dev.segs, dev.node_seg, w.papp.as<int>(), mg.pair, lvl->g_margin);
            MWG_LAUNCHED();
            s = scan_exclusive(uscan, uscan, L);
            if (s != MWG_OK) return s;
            hipLaunchKernelGGL(k_unit_write, dim3(npairs), dim3(64), 0, 0, npairs, *pr, gstart, uscan, g_unit_groups, slots, w.units.as<int4>(), w.codes.as<unsigned>(),
                               stride, mfma ? w.grec.as<int2>() : nullptr);
            MWG_LAUNCHED();
        }
        {
            ProfileScope prof_(MWG_KERNEL_BAND);
            if (precision == MWG_PREC_FP64)
                launch_mfma<N1, MWG_PREC_FP64>(lvl, dev, pr, w, stride);
            else if (precision == MWG_PREC_FP32)
                launch_mfma<N1, MWG_PREC_FP32>(lvl, dev, pr, w, stride);
            else if (precision == MWG_PREC_BF16X3)
                launch_mfma<N1, MWG_PREC_BF16X3>(lvl, dev, pr, w, stride);
            else
                hipLaunchKernelGGL(k_band_units<N1>, dim3(unit_grid<N1>()), dim3(UnitShape<N1>::NT), 0, 0, dev.segs, dev.node_seg, lvl->l,
                                   lvl->g_off, *pr, w.codes.as<unsigned>(), stride, gstart, w.units.as<int4>(), uscan, L, w.partial.as<double>(),
                                   w.counter.as<int>());
            MWG_LAUNCHED();
        }
        {
            ProfileScope prof_(MWG_KERNEL_REDUCE);
            hipLaunchKernelGGL(k_reduce_units<N1>, dim3(L, (8 * N3 + 255) / 256), dim3(256), 0, 0, dev.segs, dev.node_seg, lvl->g_off, uscan,
                               w.partial.as<double>());
            MWG_LAUNCHED();
            hipLaunchKernelGGL(k_level_norms<N1>, dim3(L), dim3(kThreads), 0, 0, dev.segs, dev.node_seg, lvl->g_off, lvl->g_norms);
            MWG_LAUNCHED();
        }
        return MWG_OK;
    }
};

template <int N1, int PREC> size_t frag_count(long long n_mat) {
    return (size_t)(n_mat / 4) * 2 * BandGeom<N1, PREC>::KSTEPS * SweepShape<N1>::NT * 64;
}
template <int N1, int PREC> mwg_status build_frag(const mwg_band_tables *op, void *dst) {
    const size_t total = frag_count<N1, PREC>(op->n_mat);
    if (total == 0) return MWG_OK;
    hipLaunchKernelGGL((k_build_frag<N1, PREC>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, op->mats, op->n_mat / 4,
                       static_cast<typename BandTypes<PREC>::B *>(dst));
    MWG_LAUNCHED();
    return MWG_OK;
}
template <int N1> struct FragL {
    static mwg_status run(const mwg_band_tables *op, int precision, void *dst) {
        if (precision == MWG_PREC_FP64) return build_frag<N1, MWG_PREC_FP64>(op, dst);
        if (precision == MWG_PREC_FP32) return build_frag<N1, MWG_PREC_FP32>(op, dst);
        return build_frag<N1, MWG_PREC_BF16X3>(op, dst);
    }
};
template <int N1> size_t frag_bytes(long long n_mat, int precision) {
    if (precision == MWG_PREC_FP64) return frag_count<N1, MWG_PREC_FP64>(n_mat) * sizeof(double);
    if (precision == MWG_PREC_FP32) return frag_count<N1, MWG_PREC_FP32>(n_mat) * sizeof(float);
    if (precision == MWG_PREC_BF16X3) return frag_count<N1, MWG_PREC_BF16X3>(n_mat) * sizeof(uint4);
    return 0;
}

} // namespace

} // namespace mwgdev

using namespace mwgdev;

// Precision 3 selects the FP64 sweeps on the vector ALUs (comparison path).
constexpr int kPrecValu = kPrecValuBand;

extern "C" size_t mwg_fragment_bytes(int kp1, long long n_mat, int precision) {
    if (n_mat < 0) return 0;
    switch (kp1) {
    case 4:
        return frag_bytes<4>(n_mat, precision);
    case 6:
        return frag_bytes<6>(n_mat, precision);
    case 8:
        return frag_bytes<8>(n_mat, precision);
    case 10:
        return frag_bytes<10>(n_mat, precision);
    default:
        return 0;
    }
}

extern "C" mwg_status mwg_build_fragments(const mwg_band_tables *op, int precision, void *dst) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!op || !dst) return fail(MWG_ERR_ARG, "null argument");
    if (op->n_mat % 4) return fail(MWG_ERR_ARG, "matrix count is not a multiple of 4");
    if (precision != MWG_PREC_FP64 && precision != MWG_PREC_FP32 && precision != MWG_PREC_BF16X3)
        return fail(MWG_ERR_UNSUPPORTED, "no operand table for this precision");
    return dispatch<FragL>(op->kp1, op, precision, dst);
}

extern "C" mwg_status mwg_set_band_layout(int layout) {
    if (layout != MWG_BAND_AUTO && layout != MWG_BAND_TILES && layout != MWG_BAND_ROWS) return fail(MWG_ERR_ARG, "unknown band layout");
    g_band_layout = layout;
    return MWG_OK;
}

extern "C" int mwg_band_layout(void) { return g_band_layout; }

extern "C" mwg_status mwg_set_band_launch(int unit_groups, int workgroups_per_cu) {
    if (unit_groups < 8 || unit_groups > 1024) return fail(MWG_ERR_ARG, "unit groups must be in 8..1024");
    if (workgroups_per_cu < 0 || workgroups_per_cu > 64) return fail(MWG_ERR_ARG, "workgroups per CU must be in 0..64");
    g_unit_groups = unit_groups;
    g_band_wg_per_cu = workgroups_per_cu;
    return MWG_OK;
}

extern "C" void mwg_band_launch(int *unit_groups, int *workgroups_per_cu) {
    if (unit_groups) *unit_groups = g_unit_groups;
    if (workgroups_per_cu) *workgroups_per_cu = g_band_wg_per_cu;
}

extern "C" mwg_status mwg_band_apply(const mwg_level *lvl, const mwg_pairs *pairs, int npairs, int precision) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!pairs) return fail(MWG_ERR_ARG, "null argument");
    if (npairs < 0) return fail(MWG_ERR_ARG, "negative pair count");
    const mwg_status st = check_level(lvl);
    if (st != MWG_OK) return st;
    if (lvl->count == 0) return MWG_OK;
    if (precision < MWG_PREC_FP64 || precision > kPrecValu) return fail(MWG_ERR_UNSUPPORTED, "unsupported sweep precision");
    if (npairs > 0 && precision != kPrecValu)
        for (int s = 0; s < lvl->nseg; s++)
            if (lvl->segs[s].count > 0 && lvl->segs[s].band_w >= 0 && !lvl->segs[s].op.frag[precision])
                return fail(MWG_ERR_ARG, "operand table of this precision not built");
    if (npairs > pairs->capacity) return fail(MWG_ERR_ARG, "more pairs than the pair arrays hold");
    return dispatch<BandL>(lvl->kp1, lvl, pairs, npairs, precision);
}
