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

// Whether band position (x, y, z) can contribute to output node gl of segment
// S; *code is the band node id, or -(ancestor id + 1) when the node is missing.
__device__ inline bool screen_position(const_as4<mwg_segment> &S, double g_thr, const int *gl, int x, int y, int z, int *code) {
    const int d = S.o_depth;
    const int wmax = S.op.wmax;
    const double *pb = S.op.pair_bound + (size_t)d * (2 * wmax + 1) + wmax;
    double a = pb[x - gl[0]] * pb[y - gl[1]] * pb[z - gl[2]];
    if (d > 0) {
        // below the root scale every tuple has a wavelet component (C, B or A)
        // in at least one direction
        const double *pw = S.op.pair_bound_w + (size_t)d * (2 * wmax + 1) + wmax;
        const double mx = pb[x - gl[0]], my = pb[y - gl[1]], mz = pb[z - gl[2]];
        a = fmax(pw[x - gl[0]] * my * mz, fmax(mx * pw[y - gl[1]] * mz, mx * my * pw[z - gl[2]]));
    }
    // a zero bound cannot pass a non-negative threshold: skip the lookup
    if (a == 0.0 && g_thr >= 0.0) return false;
    const mwg_ftree f = ld(S.f);
    // nor can a bound that fails with the largest norm of any node of the
    // tree (the test below with maxf <= norm_max; products of non-negative
    // numbers round monotonically)
    if (g_thr >= 0.0 && !(a * f.norm_max * kBoundSlack > g_thr)) return false;
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
    if (!(a * maxf * kBoundSlack > g_thr)) return false;
    // Per-term bound per f component, largest over the terms (tuple_bound,
    // tabulated by k_screen_table): rounded products are monotone, so the
    // table entry passes exactly when some term's bound passes.
    const int bm = S.op.band_max[d], nb = 2 * bm + 1;
    const double *tb = S.op.screen_bound + 8 * (S.op.screen_off[d] + (x - gl[0] + bm) + nb * ((y - gl[1] + bm) + (long long)nb * (z - gl[2] + bm)));
    for (int ft = 0; ft < 8; ft++) {
        if (fnv[ft] < kMachineZero) continue;
        if (tb[ft] * fnv[ft] * kBoundSlack > g_thr) return true;
    }
    return false;
}

// Per-term bound of the tuples (f component ft, any g component) of term i at
// translation difference dl3, without the band node's norm: the direction-d
// factor of a tuple is the norm of component 2*gbit_d + ftbit_d, so over all
// gt it is at most max(n_d(ftbit_d), n_d(2+ftbit_d)); at depth > 0 the
// (ft,gt) = (0,0) tuple is skipped, so some direction carries the gt bit.
// The band-size factor bs is the row maximum over gt.  Negative when term i
// does not reach dl3.
__device__ inline void tuple_bound(const mwg_band_tables &op, int d, int i, const int *dl3, double *out) {
    const int M = op.M;
    const int *wi = op.widths + ((size_t)d * M + i) * 5;
    const int w = wi[4];
    const int maxdl = max(abs(dl3[0]), max(abs(dl3[1]), abs(dl3[2])));
    if (maxdl > w) {
        for (int ft = 0; ft < 8; ft++) out[ft] = -1.0;
        return;
    }
    const double bs = (double)op.band_size_max[(size_t)d * M + i];
    const double *mn = op.mat_norm + op.mat_start[(size_t)d * M + i];
    double n[3][4];
    for (int dim = 0; dim < 3; dim++)
        for (int c = 0; c < 4; c++) n[dim][c] = (abs(dl3[dim]) <= wi[c]) ? mn[(dl3[dim] + w) * 4 + c] : 0.0;
    for (int ft = 0; ft < 8; ft++) {
        double g1[3], gm[3];
        for (int dim = 0; dim < 3; dim++) {
            const int fb = (ft >> dim) & 1;
            g1[dim] = n[dim][2 + fb];
            gm[dim] = fmax(n[dim][fb], g1[dim]);
        }
        double prod;
        if (d > 0 && ft == 0)
            prod = fmax(g1[0] * gm[1] * gm[2], fmax(gm[0] * g1[1] * gm[2], gm[0] * gm[1] * g1[2]));
        else
            prod = gm[0] * gm[1] * gm[2];
        out[ft] = bs * prod;
    }
}

// Screen blocks per output node of a segment (one for an empty band, so its
// nodes still get their zero pair counts).
__host__ __device__ inline int screen_blocks(int band_w) {
    if (band_w < 0) return 1;
    const long long box = (2ll * band_w + 1) * (2ll * band_w + 1) * (2ll * band_w + 1);
    return (int)((box + kScreenThreads - 1) / kScreenThreads);
}

// Segment, output node and block of a screen workgroup (seg_blk: first block
// of every segment, nseg + 1 entries).
struct ScreenBlock {
    int s, g, b;
};
__device__ inline ScreenBlock locate(const int *seg_blk, int nseg, const_as4<mwg_segment> *segs, int blk) {
    int lo = 0, hi = nseg - 1; // last segment whose first block <= blk
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (seg_blk[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const int nb = screen_blocks(segs[lo].band_w);
    const int rel = blk - seg_blk[lo];
    return {lo, segs[lo].first + rel / nb, rel % nb};
}

__device__ inline Box segment_box(const_as4<mwg_segment> &S, const int *gl) {
    if (S.band_w < 0) {
        Box b;
        for (int i = 0; i < 3; i++) {
            b.s[i] = gl[i];
            b.n[i] = 0;
        }
        return b;
    }
    const mwg_ftree f = ld(S.f);
    return band_box(f, S.o_depth, S.op.band_max[S.o_depth], gl);
}

// Screen blocks of the nodes below a device node count (one segment: its
// nodes' blocks are contiguous and of one size), else nblk.
__device__ inline int live_blocks(int nblk, const int *count_dev, const_as4<mwg_segment> *segs) {
    return count_dev ? min(nblk, *count_dev * screen_blocks(segs[0].band_w)) : nblk;
}

// One thread per band position of screen block blk; block-local compaction
// in position order.
__device__ __forceinline__ void screen_block(const_as4<mwg_segment> *segs, int nseg, const int *seg_blk, int blk, const int *l, const double *node_thr,
                                             int *cand_pos, int *cand_code, int *blk_count, int *tmp) {
    const ScreenBlock sb = locate(seg_blk, nseg, segs, blk);
    const_as4<mwg_segment> &S = segs[sb.s];
    const int gl[3] = {l[3 * sb.g], l[3 * sb.g + 1], l[3 * sb.g + 2]};
    const Box bx = segment_box(S, gl);
    const int P = bx.n[0] * bx.n[1] * bx.n[2];
    const int p = sb.b * kScreenThreads + threadIdx.x;
    int code = 0;
    bool ok = false;
    if (p < P) {
        const int x = bx.s[0] + p % bx.n[0];
        const int y = bx.s[1] + (p / bx.n[0]) % bx.n[1];
        const int z = bx.s[2] + p / (bx.n[0] * bx.n[1]);
        ok = screen_position(S, node_thr ? node_thr[sb.g] : S.g_thr, gl, x, y, z, &code);
    }
    int excl;
    const int total = block_exclusive_scan<int, kScreenThreads>(ok ? 1 : 0, tmp, excl);
    if (ok) {
        cand_pos[(size_t)blk * kScreenThreads + excl] = p;
        cand_code[(size_t)blk * kScreenThreads + excl] = code;
    }
    if (threadIdx.x == 0) blk_count[blk] = total;
}

// One workgroup per screen block; with a device node count (DEV, one
// segment) the workgroups stride over the blocks of the nodes below it.
template <bool DEV>
__global__ __launch_bounds__(kScreenThreads) void k_screen_positions(const mwg_segment *segs_, int nseg, const int *seg_blk, int nblk, const int *count_dev,
                                                                    const int *l, const double *node_thr, int *cand_pos, int *cand_code, int *blk_count,
                                                                    int *count) {
    __shared__ int tmp[kScreenThreads / 64];
    const auto segs = ro(segs_);
    if (blockIdx.x == 0 && threadIdx.x == 0) { // counters the scatter adds to
        count[1] = 0;
        count[2] = 0;
        count[3] = 0;
    }
    if constexpr (!DEV) {
        screen_block(segs, nseg, seg_blk, blockIdx.x, l, node_thr, cand_pos, cand_code, blk_count, tmp);
    } else {
        const int live = live_blocks(nblk, count_dev, segs);
        for (int blk = blockIdx.x; blk < live; blk += gridDim.x) screen_block(segs, nseg, seg_blk, blk, l, node_thr, cand_pos, cand_code, blk_count, tmp);
    }
}

// Key of missing band node (x, y, z) of segment s for the deduplication
// table: the segment and the node's offsets from the world corner, or 0 when
// they do not fit (17 bits per offset, 12 for the segment).
__device__ inline unsigned long long dedup_key(const_as4<mwg_segment> &S, int s, int x, int y, int z) {
    const int d = S.o_depth;
    const long long ox = (long long)x - (long long)S.f.corner[0] * (1ll << d);
    const long long oy = (long long)y - (long long)S.f.corner[1] * (1ll << d);
    const long long oz = (long long)z - (long long)S.f.corner[2] * (1ll << d);
    const long long lim = 1ll << 17;
    if (s >= 4096 || ox < 0 || oy < 0 || oz < 0 || ox >= lim || oy >= lim || oz >= lim) return 0ull;
    return (1ull << 63) | ((unsigned long long)s << 51) | ((unsigned long long)ox << 34) | ((unsigned long long)oy << 17) | (unsigned long long)oz;
}
__device__ inline unsigned dedup_slot(unsigned long long k) {
    k ^= k >> 31;
    k *= 0x7fb5d329728ea185ull;
    k ^= k >> 27;
    k *= 0x81dadef4bc2dd44dull;
    k ^= k >> 33;
    return (unsigned)k;
}

// Place the survivors of screen block blk at their scanned offset.
__device__ __forceinline__ void scatter_block(const_as4<mwg_segment> *segs, int nseg, const int *seg_blk, int blk, const int *l, const int *cand_pos,
                                              const int *cand_code, const int *blk_off, const mwg_pairs &out) {
    const ScreenBlock sb = locate(seg_blk, nseg, segs, blk);
    const int base = blk_off[blk];
    const int cnt = blk_off[blk + 1] - base;
    if (threadIdx.x < cnt) {
        const_as4<mwg_segment> &S = segs[sb.s];
        const int gl[3] = {l[3 * sb.g], l[3 * sb.g + 1], l[3 * sb.g + 2]};
        const Box bx = segment_box(S, gl);
        const int p = cand_pos[(size_t)blk * kScreenThreads + threadIdx.x];
        const int q = base + threadIdx.x;
        const int code = cand_code[(size_t)blk * kScreenThreads + threadIdx.x];
        const int x = bx.s[0] + p % bx.n[0], y = bx.s[1] + (p / bx.n[0]) % bx.n[1], z = bx.s[2] + p / (bx.n[0] * bx.n[1]);
        if (code < 0) {
            // one record per distinct missing band node of the segment: the
            // first pair to claim its key in the table writes it; every
            // missing pair lists its slot (or, without a table entry, its
            // own record as -(record + 1)).  A probe visits each slot at
            // most once: a table filled by a level with more pairs than
            // the capacity (screened again with larger arrays) leaves the
            // remaining pairs without an entry.
            const int k = atomicAdd(out.count + 3, 1);
            int e = 0;
            bool own = true, slot = false;
            const unsigned long long key = dedup_key(S, sb.s, x, y, z);
            if (key && out.dedup_cap > 0) {
                const unsigned mask = (unsigned)out.dedup_cap - 1;
                unsigned h = dedup_slot(key) & mask;
                for (int probe = 0; probe < out.dedup_cap; probe++, h = (h + 1) & mask) {
                    const unsigned long long prev = atomicCAS(out.dedup_keys + h, 0ull, key);
                    if (prev == 0ull || prev == key) {
                        own = prev == 0ull;
                        e = (int)h;
                        slot = true;
                        break;
                    }
                }
            }
            if (own) {
                const int m = atomicAdd(out.count + 1, 1);
                if (m < out.missing_cap) {
                    int *rec = out.missing + 6 * (size_t)m;
                    rec[0] = q;
                    rec[1] = code;
                    rec[2] = sb.g;
                    rec[3] = x;
                    rec[4] = y;
                    rec[5] = z;
                }
                if (slot) out.dedup_rec[e] = m;
                else e = -(m + 1);
            }
            if (k < out.capacity) {
                out.missing_pairs[2 * (size_t)k] = q;
                out.missing_pairs[2 * (size_t)k + 1] = e;
            }
        }
        if (q < out.capacity) {
            out.pair_g[q] = sb.g;
            out.pair_f[q] = code;
            out.pair_l[3 * q + 0] = x;
            out.pair_l[3 * q + 1] = y;
            out.pair_l[3 * q + 2] = z;
        }
    }
    if (sb.b == 0 && threadIdx.x == 0) {
        const_as4<mwg_segment> &S = segs[sb.s];
        const int nb = screen_blocks(S.band_w);
        const int np = blk_off[blk + nb] - base;
        out.g_first[sb.g] = base;
        out.g_npairs[sb.g] = np;
        // the block count covers the box only if band_w is the table's band width
        if (sb.g == S.first && S.band_w >= 0 && S.op.band_max[S.o_depth] != S.band_w) out.count[2] = 1;
    }

}

template <bool DEV>
__global__ __launch_bounds__(kScreenThreads) void k_screen_scatter(const mwg_segment *segs_, int nseg, const int *seg_blk, int nblk, const int *count_dev,
                                                                  const int *l, const int *cand_pos, const int *cand_code, const int *blk_off, mwg_pairs out) {
    const auto segs = ro(segs_);
    if constexpr (!DEV) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *out.count = blk_off[nblk];
        scatter_block(segs, nseg, seg_blk, blockIdx.x, l, cand_pos, cand_code, blk_off, out);
    } else {
        const int live = live_blocks(nblk, count_dev, segs);
        if (blockIdx.x == 0 && threadIdx.x == 0) *out.count = blk_off[live];
        for (int blk = blockIdx.x; blk < live; blk += gridDim.x) scatter_block(segs, nseg, seg_blk, blk, l, cand_pos, cand_code, blk_off, out);
    }
}

Buffer g_cand_pos, g_cand_code, g_blk, g_seg_blk;

} // namespace

} // namespace mwgdev

using namespace mwgdev;

// One thread per (depth, translation difference): the largest tuple_bound
// over the terms for each f component, NaN (which passes no threshold) when
// no term reaches the difference.
__global__ __launch_bounds__(256) void k_screen_table(mwg_band_tables op, long long entries, double *out) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= entries) return;
    int d = 0;
    while (op.screen_off[d + 1] <= e) d++;
    const int bm = op.band_max[d], nb = 2 * bm + 1;
    const long long p = e - op.screen_off[d];
    const int dl3[3] = {(int)(p % nb) - bm, (int)((p / nb) % nb) - bm, (int)(p / ((long long)nb * nb)) - bm};
    double best[8];
    bool any = false;
    for (int ft = 0; ft < 8; ft++) best[ft] = 0.0;
    for (int i = 0; i < op.M; i++) {
        double b[8];
        tuple_bound(op, d, i, dl3, b);
        if (b[0] < 0.0) continue;
        any = true;
        for (int ft = 0; ft < 8; ft++) best[ft] = fmax(best[ft], b[ft]);
    }
    for (int ft = 0; ft < 8; ft++) out[8 * e + ft] = any ? best[ft] : __builtin_nan("");
}

extern "C" mwg_status mwg_build_screen_table(const mwg_band_tables *op, long long entries, double *dst) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!op || !dst || !op->screen_off) return fail(MWG_ERR_ARG, "null argument");
    if (entries <= 0) return MWG_OK;
    hipLaunchKernelGGL(k_screen_table, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, 0, *op, entries, dst);
    MWG_LAUNCHED();
    return MWG_OK;
}

// Band node of every pair whose node was missing, from its record's node;
// the deduplication slots used go back to zero.
__global__ __launch_bounds__(256) void k_patch_pairs(mwg_pairs pr, const int *node_of_record, int count) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const int q = pr.missing_pairs[2 * (size_t)k], e = pr.missing_pairs[2 * (size_t)k + 1];
    const int m = e >= 0 ? pr.dedup_rec[e] : -e - 1;
    pr.pair_f[q] = node_of_record[m];
    if (e >= 0) pr.dedup_keys[e] = 0ull;
}

extern "C" mwg_status mwg_patch_pairs(mwg_pairs *pairs, const int *node_of_record, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!pairs || !node_of_record || !pairs->missing_pairs) return fail(MWG_ERR_ARG, "null argument");
    if (count > pairs->capacity) return fail(MWG_ERR_ARG, "more missing pairs than the pair arrays hold");
    hipLaunchKernelGGL(k_patch_pairs, dim3((count + 255) / 256), dim3(256), 0, 0, *pairs, node_of_record, count);
    MWG_LAUNCHED();
    return MWG_OK;
}

namespace mwgdev {
mwg_status screen_block_table(const mwg_level *lvl, std::vector<int> &seg_blk) {
    seg_blk.assign(lvl->nseg + 1, 0);
    long long nblk = 0;
    for (int s = 0; s < lvl->nseg; s++) {
        if (lvl->segs[s].band_w >= 0 && (!lvl->segs[s].op.screen_off || !lvl->segs[s].op.screen_bound))
            return fail(MWG_ERR_ARG, "operator screening table not built");
        seg_blk[s] = (int)nblk;
        nblk += (long long)screen_blocks(lvl->segs[s].band_w) * lvl->segs[s].count;
        if (nblk >= (1ll << 31) / kScreenThreads) return fail(MWG_ERR_UNSUPPORTED, "level too large for one screen launch");
    }
    seg_blk[lvl->nseg] = (int)nblk;
    return MWG_OK;
}
} // namespace mwgdev

extern "C" mwg_status mwg_pair_screen(const mwg_level *lvl, mwg_pairs *out) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!out || !out->count) return fail(MWG_ERR_ARG, "null argument");
    mwg_status st = check_level(lvl);
    if (st != MWG_OK) return st;
    if (!out->missing_pairs) return fail(MWG_ERR_ARG, "no missing-pair list");
    if (out->dedup_cap < 0 || (out->dedup_cap > 0 && ((out->dedup_cap & (out->dedup_cap - 1)) != 0 || (long long)out->dedup_cap < 2ll * out->capacity ||
                                                       !out->dedup_keys || !out->dedup_rec)))
        return fail(MWG_ERR_ARG, "deduplication table must be a power of two of at least twice the pair capacity");
    if (lvl->count == 0) {
        MWG_HIP(hipMemset(out->count, 0, 4 * sizeof(int)));
        return MWG_OK;
    }
    if (lvl->count_dev && lvl->nseg != 1) return fail(MWG_ERR_ARG, "a device node count needs a level of one segment");
    std::vector<int> seg_blk;
    if ((st = screen_block_table(lvl, seg_blk)) != MWG_OK) return st;
    const long long nblk = seg_blk[lvl->nseg];
    LevelDevice dev;
    if ((st = upload_level(lvl, dev)) != MWG_OK) return st;
    if (g_cand_pos.ensure(sizeof(int) * (size_t)nblk * kScreenThreads) != MWG_OK) return MWG_ERR_NOMEM;
    if (g_cand_code.ensure(sizeof(int) * (size_t)nblk * kScreenThreads) != MWG_OK) return MWG_ERR_NOMEM;
    if (g_blk.ensure(sizeof(int) * ((size_t)nblk + 1)) != MWG_OK) return MWG_ERR_NOMEM;
    const int *seg_blk_dev = staged_seg_blk(lvl, seg_blk);
    if (!seg_blk_dev) {
        if (g_seg_blk.ensure(sizeof(int) * seg_blk.size()) != MWG_OK) return MWG_ERR_NOMEM;
        if ((st = mwg_upload_async(g_seg_blk.p, seg_blk.data(), sizeof(int) * seg_blk.size())) != MWG_OK) return st;
        seg_blk_dev = g_seg_blk.as<int>();
    }
    ProfileScope prof_(MWG_KERNEL_SCREEN);
    // one workgroup per screen block; with a device node count (an upper
    // bound on the host) a grid of at most 8 workgroups per compute unit
    // strides over the blocks
    const unsigned grid = lvl->count_dev ? (unsigned)std::min<long long>(nblk, 8ll * device_cus()) : (unsigned)nblk;
    if (lvl->count_dev)
        hipLaunchKernelGGL(k_screen_positions<true>, dim3(grid), dim3(kScreenThreads), 0, 0, dev.segs, lvl->nseg, seg_blk_dev, (int)nblk, lvl->count_dev,
                           lvl->l, lvl->node_thr, g_cand_pos.as<int>(), g_cand_code.as<int>(), g_blk.as<int>(), out->count);
    else
        hipLaunchKernelGGL(k_screen_positions<false>, dim3(grid), dim3(kScreenThreads), 0, 0, dev.segs, lvl->nseg, seg_blk_dev, (int)nblk, nullptr, lvl->l,
                           lvl->node_thr, g_cand_pos.as<int>(), g_cand_code.as<int>(), g_blk.as<int>(), out->count);
    MWG_LAUNCHED();
    if ((st = scan_exclusive(g_blk.as<int>(), g_blk.as<int>(), (int)nblk, lvl->count_dev, screen_blocks(lvl->segs[0].band_w))) != MWG_OK) return st;
    if (lvl->count_dev)
        hipLaunchKernelGGL(k_screen_scatter<true>, dim3(grid), dim3(kScreenThreads), 0, 0, dev.segs, lvl->nseg, seg_blk_dev, (int)nblk, lvl->count_dev,
                           lvl->l, g_cand_pos.as<int>(), g_cand_code.as<int>(), g_blk.as<int>(), *out);
    else
        hipLaunchKernelGGL(k_screen_scatter<false>, dim3(grid), dim3(kScreenThreads), 0, 0, dev.segs, lvl->nseg, seg_blk_dev, (int)nblk, nullptr, lvl->l,
                           g_cand_pos.as<int>(), g_cand_code.as<int>(), g_blk.as<int>(), *out);
    MWG_LAUNCHED();
    return MWG_OK;
}
