// Device copy of a level's segments and the segment of every output node.
#include "device_common.h"

#include <cstring>

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
