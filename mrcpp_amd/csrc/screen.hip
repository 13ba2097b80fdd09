// Candidate (output node, band node) pairs of one breadth-first level
// (ConvolutionCalculator::makeOperBand / fillOperBand,
// src/treebuilders/ConvolutionCalculator.cpp:142-222, with the norm test of
// applyOperator, :297-329, applied as upper bounds).
//
// The band box of output node g is split into blocks of 256 positions, one
// thread per position, so a level launches (#nodes x #blocks per box)
// workgroups.  Each block compacts its survivors in position order; an
// exclusive scan over the block counts (g-major) then places every pair, so
// the pairs of one output node are contiguous and in position order (x
// fastest), independent of scheduling.
#include "device_common.h"

namespace mwgdev {

namespace {

constexpr int kScreenThreads = 256;

struct Box {
    int s[3], n[3];
};

// Band box of output translation gl at depth d, clamped to the world box.
__device__ inline Box band_box(const mwg_ftree &f, int d, int W, const int *gl) {
    Box b;
    for (int i = 0; i < 3; i++) {
        const int c = f.corner[i] * (1 << d);
        const int nb = f.nbox[i] * (1 << d);
        b.s[i] = max(gl[i] - W, c);
        const int e = min(gl[i] + W, c + nb - 1);
        b.n[i] = max(e - b.s[i] + 1, 0);
    }
    return b;
}

// Whether band position (x, y, z) can contribute to output node gl; *code is
// the band node id, or -(ancestor id + 1) when the node is missing.
__device__ inline bool screen_position(const mwg_level &lvl, const mwg_ftree &f, const mwg_band_tables &op, const int *gl, int x, int y, int z,
                                       int *code) {
    const int d = lvl.o_depth;
    const double *pb = op.pair_bound + (size_t)d * (2 * op.wmax + 1) + op.wmax;
    double a = pb[x - gl[0]] * pb[y - gl[1]] * pb[z - gl[2]];
    if (d > 0) {
        // below the root scale every tuple has a wavelet component (C, B or A)
        // in at least one direction
        const double *pw = op.pair_bound_w + (size_t)d * (2 * op.wmax + 1) + op.wmax;
        const double mx = pb[x - gl[0]], my = pb[y - gl[1]], mz = pb[z - gl[2]];
        a = fmax(pw[x - gl[0]] * my * mz, fmax(mx * pw[y - gl[1]] * mz, mx * my * pw[z - gl[2]]));
    }
    // a zero bound cannot pass a non-negative threshold: skip the lookup
    if (a == 0.0 && lvl.g_thr >= 0.0) return false;
    const int id = f_lookup(f, d, x, y, z);
    double maxf;
    double fnv[8]; // per-component norms of the band node (or their bound)
    if (id >= 0) {
        const double *nr = f.norms + (size_t)id * 9;
        maxf = 0.0;
        for (int t = 0; t < 8; t++) {
            fnv[t] = nr[t];
            maxf = fmax(maxf, nr[t]);
        }
        *code = id;
    } else {
        int dd = d, xx = x, yy = y, zz = z, anc = -1;
        while (anc < 0 && dd > 0) {
            dd--;
            xx >>= 1;
            yy >>= 1;
            zz >>= 1;
            anc = f_lookup(f, dd, xx, yy, zz);
        }
        if (anc < 0) return false; // outside the world (not reachable for clamped boxes)
        const double *nr = f.norms + (size_t)anc * 9;
        // a generated node holds a scaling block only, whose norm is bounded
        // by its nearest existing ancestor's (orthogonal reconstruction)
        maxf = f.is_gen[anc] ? nr[0] : sqrt(nr[8]);
        fnv[0] = maxf;
        for (int t = 1; t < 8; t++) fnv[t] = 0.0;
        *code = -(anc + 1);
    }
    if (!(a * maxf * kBoundSlack > lvl.g_thr)) return false;
    // Per-term bound, resolved per f component ft: the direction-d factor of
    // a tuple is the norm of component 2*gbit_d + ftbit_d, so over all gt it is
    // at most max(n_d(ftbit_d), n_d(2+ftbit_d)); at depth > 0 the (ft,gt) =
    // (0,0) tuple is skipped, so some direction carries the gt bit.
    const int M = op.M;
    const int dl3[3] = {x - gl[0], y - gl[1], z - gl[2]};
    const int maxdl = max(abs(dl3[0]), max(abs(dl3[1]), abs(dl3[2])));
    const int *wd = op.widths + (size_t)d * M * 5;
    const long long *ms = op.mat_start + (size_t)d * M;
    const int *bsm = op.band_size_max + (size_t)d * M;
    for (int i = 0; i < M; i++) {
        const int *wi = wd + i * 5;
        const int w = wi[4];
        if (maxdl > w) continue;
        // cheap per-term bound first: largest component norm per direction
        const double *tn = op.term_norm + ms[i] / 4;
        const double bs = (double)bsm[i];
        if (!(bs * tn[dl3[0] + w] * tn[dl3[1] + w] * tn[dl3[2] + w] * maxf * kBoundSlack > lvl.g_thr)) continue;
        const double *mn = op.mat_norm + ms[i];
        double n[3][4];
        for (int dim = 0; dim < 3; dim++)
            for (int c = 0; c < 4; c++) n[dim][c] = (abs(dl3[dim]) <= wi[c]) ? mn[(dl3[dim] + w) * 4 + c] : 0.0;
        for (int ft = 0; ft < 8; ft++) {
            if (fnv[ft] < kMachineZero) continue;
            double g0[3], g1[3], gm[3];
            for (int dim = 0; dim < 3; dim++) {
                const int fb = (ft >> dim) & 1;
                g0[dim] = n[dim][fb];
                g1[dim] = n[dim][2 + fb];
                gm[dim] = fmax(g0[dim], g1[dim]);
            }
            double prod;
            if (d > 0 && ft == 0)
                prod = fmax(g1[0] * gm[1] * gm[2], fmax(gm[0] * g1[1] * gm[2], gm[0] * gm[1] * g1[2]));
            else
                prod = gm[0] * gm[1] * gm[2];
            if (bs * prod * fnv[ft] * kBoundSlack > lvl.g_thr) return true;
        }
    }
    return false;
}

// One thread per band position; block-local compaction in position order.
__global__ __launch_bounds__(kScreenThreads) void k_screen_positions(mwg_level lvl, mwg_ftree f, mwg_band_tables op, int nb, int *cand_pos,
                                                                    int *cand_code, int *blk_count) {
    __shared__ int tmp[kScreenThreads / 64];
    const int blk = blockIdx.x;
    const int g = blk / nb, b = blk - g * nb;
    const int d = lvl.o_depth;
    const int gl[3] = {lvl.l[3 * g], lvl.l[3 * g + 1], lvl.l[3 * g + 2]};
    const Box bx = band_box(f, d, op.band_max[d], gl);
    const int P = bx.n[0] * bx.n[1] * bx.n[2];
    const int p = b * kScreenThreads + threadIdx.x;
    int code = 0;
    bool ok = false;
    if (p < P) {
        const int x = bx.s[0] + p % bx.n[0];
        const int y = bx.s[1] + (p / bx.n[0]) % bx.n[1];
        const int z = bx.s[2] + p / (bx.n[0] * bx.n[1]);
        ok = screen_position(lvl, f, op, gl, x, y, z, &code);
    }
    int excl;
    const int total = block_exclusive_scan<int, kScreenThreads>(ok ? 1 : 0, tmp, excl);
    if (ok) {
        cand_pos[(size_t)blk * kScreenThreads + excl] = p;
        cand_code[(size_t)blk * kScreenThreads + excl] = code;
    }
    if (threadIdx.x == 0) blk_count[blk] = total;
}

__global__ void k_screen_reset(int *count) { count[1] = 0; }

// Place the survivors of every block at their scanned offset.
__global__ __launch_bounds__(kScreenThreads) void k_screen_scatter(mwg_level lvl, mwg_ftree f, mwg_band_tables op, int nb, const int *cand_pos,
                                                                  const int *cand_code, const int *blk_off, mwg_pairs out) {
    const int blk = blockIdx.x;
    const int g = blk / nb, b = blk - g * nb;
    const int d = lvl.o_depth;
    const int base = blk_off[blk];
    const int cnt = blk_off[blk + 1] - base;
    if (threadIdx.x < cnt) {
        const int gl[3] = {lvl.l[3 * g], lvl.l[3 * g + 1], lvl.l[3 * g + 2]};
        const Box bx = band_box(f, d, op.band_max[d], gl);
        const int p = cand_pos[(size_t)blk * kScreenThreads + threadIdx.x];
        const int q = base + threadIdx.x;
        const int code = cand_code[(size_t)blk * kScreenThreads + threadIdx.x];
        if (code < 0) atomicAdd(out.count + 1, 1);
        if (q < out.capacity) {
            out.pair_g[q] = g;
            out.pair_f[q] = code;
            out.pair_l[3 * q + 0] = bx.s[0] + p % bx.n[0];
            out.pair_l[3 * q + 1] = bx.s[1] + (p / bx.n[0]) % bx.n[1];
            out.pair_l[3 * q + 2] = bx.s[2] + p / (bx.n[0] * bx.n[1]);
        }
    }
    if (b == 0 && threadIdx.x == 0) {
        out.g_first[g] = base;
        out.g_npairs[g] = blk_off[(g + 1) * nb] - base;
    }
    if (blk == 0 && threadIdx.x == 0) {
        // the grid covers the box only if the host's band width is the table's
        const int W = op.band_max[d];
        const long long need = (long long)(2 * W + 1) * (2 * W + 1) * (2 * W + 1);
        *out.count = (need <= (long long)nb * kScreenThreads) ? blk_off[gridDim.x] : -1;
    }
}

Buffer g_cand_pos, g_cand_code, g_blk;

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_pair_screen(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, mwg_pairs *out) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!lvl || !f || !op || !out || !out->count) return fail(MWG_ERR_ARG, "null argument");
    if (lvl->o_depth < 0 || lvl->o_depth >= op->n_depth) return fail(MWG_ERR_ARG, "level depth outside the band tables");
    if (lvl->o_depth > mwgkey::kMaxDepth) return fail(MWG_ERR_UNSUPPORTED, "level deeper than the device key range");
    if (lvl->band_w < 0) return fail(MWG_ERR_ARG, "negative band width");
    if (lvl->count <= 0) {
        MWG_HIP(hipMemset(out->count, 0, 2 * sizeof(int)));
        return MWG_OK;
    }
    const long long W = lvl->band_w;
    const long long box = (2 * W + 1) * (2 * W + 1) * (2 * W + 1);
    const long long nb = (box + kScreenThreads - 1) / kScreenThreads;
    const long long nblk = nb * lvl->count;
    if (nblk >= (1ll << 31) / kScreenThreads) return fail(MWG_ERR_UNSUPPORTED, "level too large for one screen launch");
    if (g_cand_pos.ensure(sizeof(int) * (size_t)nblk * kScreenThreads) != MWG_OK) return MWG_ERR_NOMEM;
    if (g_cand_code.ensure(sizeof(int) * (size_t)nblk * kScreenThreads) != MWG_OK) return MWG_ERR_NOMEM;
    if (g_blk.ensure(sizeof(int) * ((size_t)nblk + 1)) != MWG_OK) return MWG_ERR_NOMEM;
    ProfileScope prof_(MWG_KERNEL_SCREEN);
    hipLaunchKernelGGL(k_screen_reset, dim3(1), dim3(1), 0, 0, out->count);
    MWG_LAUNCHED();
    hipLaunchKernelGGL(k_screen_positions, dim3((unsigned)nblk), dim3(kScreenThreads), 0, 0, *lvl, *f, *op, (int)nb, g_cand_pos.as<int>(),
                       g_cand_code.as<int>(), g_blk.as<int>());
    MWG_LAUNCHED();
    const mwg_status s = scan_exclusive(g_blk.as<int>(), g_blk.as<int>(), (int)nblk);
    if (s != MWG_OK) return s;
    hipLaunchKernelGGL(k_screen_scatter, dim3((unsigned)nblk), dim3(kScreenThreads), 0, 0, *lvl, *f, *op, (int)nb, g_cand_pos.as<int>(),
                       g_cand_code.as<int>(), g_blk.as<int>(), *out);
    MWG_LAUNCHED();
    return MWG_OK;
}
