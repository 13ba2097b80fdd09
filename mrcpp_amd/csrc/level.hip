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

Buffer g_segs, g_node_seg;
// the segments and node count of the last upload (a level whose segments did
// not change since is not uploaded again)
std::vector<char> g_last_segs;
int g_last_count = -1;
void *g_last_segs_dev = nullptr;

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
    const size_t sb = sizeof(mwg_segment) * (size_t)lvl->nseg;
    const size_t b0 = g_segs.bytes, b1 = g_node_seg.bytes;
    if (g_segs.ensure(sb) != MWG_OK || g_node_seg.ensure(sizeof(int) * ((size_t)lvl->count + 1)) != MWG_OK) return MWG_ERR_NOMEM;
    const bool grown = g_segs.bytes != b0 || g_node_seg.bytes != b1; // reallocated: contents lost
    const bool same = !grown && g_last_count == lvl->count && g_last_segs_dev == g_segs.p && g_last_segs.size() == sb &&
                      std::memcmp(g_last_segs.data(), lvl->segs, sb) == 0;
    if (!same) {
        const mwg_status s = mwg_upload_async(g_segs.p, lvl->segs, sb);
        if (s != MWG_OK) return s;
        hipLaunchKernelGGL(k_node_seg, dim3((lvl->count + 255) / 256), dim3(256), 0, 0, g_segs.as<mwg_segment>(), lvl->nseg, lvl->count,
                           g_node_seg.as<int>());
        MWG_LAUNCHED();
        g_last_segs.assign(reinterpret_cast<const char *>(lvl->segs), reinterpret_cast<const char *>(lvl->segs) + sb);
        g_last_count = lvl->count;
        g_last_segs_dev = g_segs.p;
    }
    out.segs = g_segs.as<mwg_segment>();
    out.node_seg = g_node_seg.as<int>();
    return MWG_OK;
}

} // namespace mwgdev
