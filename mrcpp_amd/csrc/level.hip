// Device copy of a level's segments and the segment of every output node;
// the next breadth-first level formed on the device.
#include "device_common.h"

#include <cstring>

// split decisions equal to the host's: every product and sum rounded on its own
#pragma clang fp contract(off)

namespace mwgdev {

namespace {

__global__ __launch_bounds__(256) void k_node_seg(const mwg_segment *segs, int nseg, int count, int *node_seg) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= count) return;
    int lo = 0, hi = nseg - 1; // last segment with first <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (segs[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    node_seg[g] = lo;
}

constexpr int kNextLevelMax = 4096; // nodes of a finished level (their wavelet norms held in LDS)

// Wavelet square norm of a node from its component norms
// (NodeStore::waveletNorm: sequential, -1 once a component norm is unset).
__device__ inline double node_wavelet_sq(const double *n) {
    double w = 0.0;
    for (int t = 1; t < 8; t++) {
        const double v = n[t];
        if (v >= 0.0) w += v * v;
        else w = -1.0;
    }
    return w;
}

// The next breadth-first level from a finished one (mwg_next_level).  One
// workgroup: the nodes' wavelet norms are formed in parallel into LDS and
// summed by one thread in node order; then the split flags, their ranks
// (wave ballots) and the children, a chunk of kNextThreads nodes at a time.
constexpr int kNextThreads = 256;
__global__ __launch_bounds__(kNextThreads) void k_next_level(mwg_next_level_args a) {
    __shared__ double s_w[kNextLevelMax];
    __shared__ double s_thr[2];
    __shared__ int s_wave[kNextThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < a.count; i += kNextThreads) s_w[i] = node_wavelet_sq(a.g_norms + 9 * (size_t)i);
    __syncthreads();
    if (tid == 0) {
        double step = 0.0; // in node order; eight LDS reads issued ahead of their additions
        int k = 0;
        for (; k + 8 <= a.count; k += 8) {
            double v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = s_w[k + j];
#pragma unroll
            for (int j = 0; j < 8; j++) step += v[j];
        }
        for (; k < a.count; k++) step += s_w[k];
        const double wn = a.w_norm + step;
        const double tsn = (a.s_norm < 0.0 || wn < 0.0) ? -1.0 : a.s_norm + wn;
        const double tn = (tsn > 0.0 && !a.abs_prec) ? sqrt(tsn) : 1.0;
        const double x = a.prec * tn * a.scale_fac;
        const double gthr = tsn > 0.0 ? a.prec * 1.0 * sqrt(tsn / a.n_terms) : tsn;
        s_thr[0] = a.thr_floor < x ? x : a.thr_floor; // std::max(floor, x)
        s_thr[1] = gthr;
        a.head[0] = tsn;
        a.head[1] = gthr;
        a.head[2] = wn;
    }
    __syncthreads();
    const double thr = s_thr[0], gthr = s_thr[1];
    const bool may = a.may_split && a.prec > 0.0;
    int before = 0;
    for (int c0 = 0; c0 < a.count; c0 += kNextThreads) {
        const int i = c0 + tid;
        bool sp = false;
        if (i < a.count) {
            sp = may && sqrt(s_w[i]) > thr;
            a.split[i] = sp ? 1 : 0;
        }
        const unsigned long long m = __ballot(sp);
        if (lane == 0) s_wave[wave] = __popcll(m);
        __syncthreads();
        int off = before, total = 0;
        for (int w = 0; w < kNextThreads / 64; w++) {
            if (w < wave) off += s_wave[w];
            total += s_wave[w];
        }
        if (sp) {
            const int k = off + __popcll(m & ((1ull << lane) - 1ull));
            for (int c = 0; c < 8; c++) {
                const int j = 8 * k + c;
                a.next_off[j] = (a.base_id + j) * a.ncoef;
                for (int d = 0; d < 3; d++) a.next_l[3 * (size_t)j + d] = 2 * a.l[3 * (size_t)i + d] + ((c >> d) & 1);
                a.next_thr[j] = gthr;
            }
        }
        before += total;
        __syncthreads();
    }
    if (tid == 0) {
        *a.next_count = 8 * before;
        a.head[3] = (double)(8 * before);
    }
}

Buffer g_segs, g_node_seg, g_arena;
// What the device holds for the last level: its segments and node count,
// and where its segment table, node segments and (when staged by
// mwg_level_upload) screen block table are.  Reset whenever one of these
// buffers is reallocated.
struct LevelCache {
    std::vector<char> segs;
    int count = -1;
    const mwg_segment *segs_dev = nullptr;
    const int *node_seg_dev = nullptr;
    std::vector<int> seg_blk;
    const int *seg_blk_dev = nullptr;
} g_cache;

bool cache_hit(const mwg_level *lvl) {
    const size_t sb = sizeof(mwg_segment) * (size_t)lvl->nseg;
    return g_cache.segs_dev && g_cache.count == lvl->count && g_cache.segs.size() == sb && std::memcmp(g_cache.segs.data(), lvl->segs, sb) == 0;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

} // namespace

mwg_status check_level(const mwg_level *lvl) {
    if (!lvl) return fail(MWG_ERR_ARG, "null level");
    if (lvl->count < 0 || lvl->nseg < 0) return fail(MWG_ERR_ARG, "negative level size");
    if (lvl->count == 0) return MWG_OK;
    if (!lvl->segs || lvl->nseg == 0 || !lvl->l || !lvl->g_off || !lvl->g_norms) return fail(MWG_ERR_ARG, "null level array");
    int next = 0;
    for (int s = 0; s < lvl->nseg; s++) {
        const mwg_segment &S = lvl->segs[s];
        if (S.first != next || S.count < 0) return fail(MWG_ERR_ARG, "segments do not partition the level in order");
        next += S.count;
        if (S.f.kp1 != lvl->kp1 || S.op.kp1 != lvl->kp1) return fail(MWG_ERR_ARG, "operator, level and tree orders differ");
        if (S.op.M > kMaxTerms) return fail(MWG_ERR_UNSUPPORTED, "separation rank above 1000");
        if (S.op.n_mat / 4 >= (1ll << 20)) return fail(MWG_ERR_UNSUPPORTED, "operator tables above 2^20 matrix blocks");
        if (!S.g_coef || !S.counters) return fail(MWG_ERR_ARG, "segment without output pool or counters");
        if (S.band_w >= 0) {
            if (S.o_depth < 0 || S.o_depth >= S.op.n_depth) return fail(MWG_ERR_ARG, "segment depth outside its band tables");
            if (S.o_depth > mwgkey::kMaxDepth) return fail(MWG_ERR_UNSUPPORTED, "level deeper than the device key range");
        }
    }
    if (next != lvl->count) return fail(MWG_ERR_ARG, "segments do not cover the level");
    return MWG_OK;
}

mwg_status upload_level(const mwg_level *lvl, LevelDevice &out) {
    if (!cache_hit(lvl)) {
        const size_t sb = sizeof(mwg_segment) * (size_t)lvl->nseg;
        g_cache = LevelCache{};
        if (g_segs.ensure(sb) != MWG_OK || g_node_seg.ensure(sizeof(int) * ((size_t)lvl->count + 1)) != MWG_OK) return MWG_ERR_NOMEM;
        const mwg_status s = mwg_upload_async(g_segs.p, lvl->segs, sb);
        if (s != MWG_OK) return s;
        hipLaunchKernelGGL(k_node_seg, dim3((lvl->count + 255) / 256), dim3(256), 0, 0, g_segs.as<mwg_segment>(), lvl->nseg, lvl->count,
                           g_node_seg.as<int>());
        MWG_LAUNCHED();
        g_cache.segs.assign(reinterpret_cast<const char *>(lvl->segs), reinterpret_cast<const char *>(lvl->segs) + sb);
        g_cache.count = lvl->count;
        g_cache.segs_dev = g_segs.as<mwg_segment>();
        g_cache.node_seg_dev = g_node_seg.as<int>();
    }
    out.segs = g_cache.segs_dev;
    out.node_seg = g_cache.node_seg_dev;
    return MWG_OK;
}

const int *staged_seg_blk(const mwg_level *lvl, const std::vector<int> &seg_blk) {
    return cache_hit(lvl) && g_cache.seg_blk_dev && g_cache.seg_blk == seg_blk ? g_cache.seg_blk_dev : nullptr;
}

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_next_level(const mwg_next_level_args *a) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!a || a->count < 0) return fail(MWG_ERR_ARG, "bad next-level arguments");
    if (a->count == 0) return MWG_OK;
    if (!a->g_norms || !a->l || !a->head || !a->split || !a->next_off || !a->next_l || !a->next_thr || !a->next_count)
        return fail(MWG_ERR_ARG, "null next-level array");
    if (!(a->n_terms > 0.0) || a->ncoef <= 0 || a->base_id < 0) return fail(MWG_ERR_ARG, "bad next-level sizes");
    if (a->count > kNextLevelMax) return fail(MWG_ERR_UNSUPPORTED, "more than 4096 nodes for the next-level kernel");
    hipLaunchKernelGGL(k_next_level, dim3(1), dim3(kNextThreads), 0, 0, *a);
    MWG_LAUNCHED();
    return MWG_OK;
}

extern "C" mwg_status mwg_level_upload(const mwg_level *lvl, const void *data, size_t bytes, void **data_dev) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!lvl || !data_dev || (bytes && !data)) return fail(MWG_ERR_ARG, "null argument");
    if (lvl->count < 0 || lvl->nseg < 0 || (lvl->nseg && !lvl->segs)) return fail(MWG_ERR_ARG, "bad level segments");
    int next = 0;
    for (int s = 0; s < lvl->nseg; s++) {
        if (lvl->segs[s].first != next || lvl->segs[s].count < 0) return fail(MWG_ERR_ARG, "segments do not partition the level in order");
        next += lvl->segs[s].count;
    }
    if (next != lvl->count) return fail(MWG_ERR_ARG, "segments do not cover the level");
    // [data | segments | node segments | screen block table], 256-byte aligned parts
    std::vector<int> seg_blk;
    const bool blk = lvl->count > 0 && screen_block_table(lvl, seg_blk) == MWG_OK;
    const size_t sb = sizeof(mwg_segment) * (size_t)lvl->nseg;
    const size_t o_segs = align256(bytes), o_node = align256(o_segs + sb), o_blk = align256(o_node + sizeof(int) * ((size_t)lvl->count + 1));
    const size_t total = o_blk + (blk ? sizeof(int) * seg_blk.size() : 0);
    std::vector<unsigned char> buf(total, 0);
    if (bytes) std::memcpy(buf.data(), data, bytes);
    if (sb) std::memcpy(buf.data() + o_segs, lvl->segs, sb);
    int *ns = reinterpret_cast<int *>(buf.data() + o_node);
    for (int s = 0; s < lvl->nseg; s++)
        for (int i = 0; i < lvl->segs[s].count; i++) ns[lvl->segs[s].first + i] = s;
    if (blk) std::memcpy(buf.data() + o_blk, seg_blk.data(), sizeof(int) * seg_blk.size());
    g_cache = LevelCache{};
    if (g_arena.ensure(total) != MWG_OK) return MWG_ERR_NOMEM;
    const mwg_status st = mwg_upload_async(g_arena.p, buf.data(), total);
    if (st != MWG_OK) return st;
    unsigned char *base = g_arena.as<unsigned char>();
    g_cache.segs.assign(reinterpret_cast<const char *>(lvl->segs), reinterpret_cast<const char *>(lvl->segs) + sb);
    g_cache.count = lvl->count;
    g_cache.segs_dev = reinterpret_cast<const mwg_segment *>(base + o_segs);
    g_cache.node_seg_dev = reinterpret_cast<const int *>(base + o_node);
    if (blk) {
        g_cache.seg_blk = seg_blk;
        g_cache.seg_blk_dev = reinterpret_cast<const int *>(base + o_blk);
    }
    *data_dev = base;
    return MWG_OK;
}
