// apply(): MRCPP's breadth-first operator application with every level's band
// sum, generated band nodes and final multiwavelet transforms on the MI355X.
// The host keeps the tree topology and the split decisions
// (src/treebuilders/apply.cpp:68-93, src/treebuilders/TreeBuilder.cpp:38-87);
// optionally (set_level_ahead) the device forms each next level and its pair
// screen first, and the host keeps it where its own split agrees bit for bit.
//
// apply_batch() runs many applications together: each breadth-first step
// gathers the work nodes of every unfinished tree into one level of
// (tree, scale) segments, so one screen and one band launch serve all trees.
// The other trees of a batch change a tree's coefficients only in the last
// bits: a level batch of more groups than the minimum unit size times the
// partial slots (32 x 32768 at order 7) uses larger work units,
// which adds the band sum's partial results in another order (the screening,
// and so the applied tuples, is per tree).
#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <omp.h>
#include <mutex>
#include <map>
#include <memory>
#include <set>

#include "function_tree.h"
#include "keys.h"
#include "max_norms.h"
#include "predictor.h"

namespace mrcpp {

namespace {

thread_local ApplyStats g_stats;
thread_local std::vector<ApplyStats> g_batch_stats;
thread_local SweepPrecision g_sweep = SweepPrecision::FP64;
thread_local bool g_ahead = false;

using clk = std::chrono::steady_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

// Device scratch reused across levels and applies.
struct Scratch {
    device::DeviceArray<int> pairG, pairF, pairL, gFirst, gCount, missing, missingPairs, dedupRec;
    // levels formed on the device (mwg_next_level): two sets, the level being
    // computed and the next one
    device::DeviceArray<long long> nextOff[2];
    device::DeviceArray<int> nextL[2];
    device::DeviceArray<double> nextThr[2];
    device::DeviceArray<int> split, nextCount;
    device::DeviceArray<double> nextHead;
    device::DeviceArray<unsigned long long> dedupKeys;
    bool dedupClean = false; // every slot of dedupKeys is zero
    device::DeviceArray<long long> gen; // one upload per level: generated nodes and pair patches
    device::DeviceArray<long long> pOff, cOff;
    device::DeviceArray<double> gNorms, gMargin;
    device::DeviceArray<unsigned long long> counters; // [job][2]: applied tuples, candidate pairs
};

// kept until the process ends (never destroyed: no device call runs from a
// static destructor)
Scratch &scratch() {
    static Scratch *s = new Scratch();
    return *s;
}

// Device bytes per level batch for the screen candidates and the per-pair
// group lists.
constexpr double kBatchBudget = 16.0e9;

// Local precision factors of apply with precision trees
// (src/treebuilders/apply.cpp:214-251): precFunc(idx) = 1 / max over the trees
// of sqrt(getMaxSquareNorm) of the tree's node at idx, 1 without trees.
// getMaxSquareNorm (src/trees/MWNode.h:84) is the largest scaled square norm
// 2^(3n) |c|^2 of the node and its real descendants (MWTree::makeMaxSquareNorms,
// src/trees/MWTree.cpp:536-543; MWNode::setMaxSquareNorm,
// src/trees/MWNode.cpp:1257-1268).  A node below a tree's leaves is generated
// from its nearest real ancestor (MWTree::getNode) and contributes its own
// scaled square norm; generated nodes live in a private cache, so the trees
// themselves are not changed.
class PrecFactors {
public:
    explicit PrecFactors(const std::vector<FunctionTree<3> *> &trees) {
        for (FunctionTree<3> *t : trees) src_.emplace_back(t->nodes());
    }

    double operator()(int n, const std::array<int, 3> &l) {
        double maxNorm = src_.empty() ? 1.0 : 0.0;
        for (detail::MaxNorms &m : src_) maxNorm = std::max(maxNorm, std::sqrt(m.at(n, l).maxSq));
        return 1.0 / maxNorm;
    }

private:
    std::vector<detail::MaxNorms> src_;
};

// One tree of a batch and its state across breadth-first steps.
enum class Mode { Plain, Replay };

struct Job {
    ApplyTask task;
    const MultiResolutionAnalysis<3> *mra = nullptr;
    const detail::BandTables *ht = nullptr;
    const mwg_band_tables *bt = nullptr;
    device::TreeMirror *fm = nullptr;
    detail::NodeStore<3> *fstore = nullptr;
    detail::NodeStore<3> *g = nullptr;
    const double *filters = nullptr;
    std::vector<int> work;
    double sNorm = 0.0, wNorm = 0.0, gthr = -1.0;
    int iter = 0;
    ApplyStats stats;
    // refinement predictor
    Mode mode = Mode::Plain;
    PredictorContext ctx;
    std::vector<int> replayWork; // nodes of the breadth-first step being judged (computed)
    std::vector<char> speculative; // computed ahead of its step (with an earlier step's threshold)
    bool edgeChecked = false;      // uncertain nodes of the current replay step recomputed
    bool marginPass = false;       // the next pass computes speculative nodes: record their screening margins
    std::vector<std::array<double, MWG_MARGINS>> margin; // per node: screening margins of the pass that computed it
    std::vector<double> thrUsed;               // per node: the screening threshold it was computed with
    double normErr = 0.0; // bound on |approximate tree norm - the plain method's| from kept speculative nodes
    // precision trees: precFunc of every output node (by id, -1 until needed)
    std::unique_ptr<PrecFactors> precFunc;
    std::vector<double> precFac;
};

// Precision of output node u (WaveletAdaptor::splitNode, src/treebuilders/WaveletAdaptor.h:51-54).
double node_prec(const Job &job, int u) { return job.precFunc ? job.task.prec * job.precFac[u] : job.task.prec; }

// Screening threshold of output node u (ConvolutionCalculator::calcNode,
// src/treebuilders/ConvolutionCalculator.cpp:241-246).
double node_threshold(const Job &job, int u) {
    double t = job.g->treeSquareNorm;
    // precFunc is 1 without precision trees (precFac is then empty)
    const double fac = job.precFunc ? job.precFac[u] : 1.0;
    if (t > 0.0) t = job.task.prec * fac * std::sqrt(t / (double)job.task.oper->size());
    return t;
}

// precFunc of the work nodes not evaluated yet.
void eval_prec_factors(Job &job) {
    if (!job.precFunc) return;
    const detail::NodeStore<3> &g = *job.g;
    if (job.precFac.size() < (size_t)g.size()) job.precFac.resize(g.size(), -1.0);
    for (int u : job.work)
        if (job.precFac[u] < 0.0) job.precFac[u] = (*job.precFunc)(g.scale(u), g.transl(u));
}

// Threshold ratios of the screening margins recorded for speculative nodes
// (mwg_level::margin_ratio): a node computed with threshold t and judged at a
// step whose own threshold t' satisfies t' <= kMarginRatio[j] t differs from
// its recomputation at t' by at most its margin j in 2-norm.  The screening
// threshold grows with the square root of the approximate tree norm; after
// the first two steps it typically grows by 0.1-0.3 % more for Poisson trees
// and by up to 30 % for Helmholtz trees, hence the fine bins near 1.
constexpr double kMarginRatio[MWG_MARGINS] = {1.0003, 1.001, 1.003, 1.01, 1.03, 1.1, 1.3, 2.0};

struct Shape {
    int kp1, n3, ncoef;
};

// Plan the generated nodes some pairs of a tree need into its input mirror's
// host table and set those pairs' band nodes (pf[i] < 0: ancestor code; band
// translation pl, target scale sc[i]); returns the nodes to generate, for
// MirrorBatch::add.  Touches only the tree's input mirror: trees with
// different inputs may be planned concurrently.
std::vector<device::GenNode> plan_band_nodes(Job &job, std::vector<int> &pf, const std::vector<int> &pl, const std::vector<int> &sc) {
    device::TreeMirror &fm = *job.fm;
    struct NewNode {
        int id, parent, child, scale;
    };
    std::vector<NewNode> made;
    for (size_t p = 0; p < pf.size(); p++) {
        if (pf[p] >= 0) continue;
        const int anc = -pf[p] - 1;
        const int scale = sc[p];
        const std::array<int, 3> l = {pl[3 * p], pl[3 * p + 1], pl[3 * p + 2]};
        int prev = anc;
        const int prevScale = (anc < fm.nReal) ? job.fstore->scale(anc) : fm.genScale[anc - fm.nReal];
        for (int s = prevScale + 1; s <= scale; s++) {
            std::array<int, 3> ls;
            for (int d = 0; d < 3; d++) ls[d] = l[d] >> (scale - s);
            const mwgkey::Key key = device::key_of(*job.mra, s, ls);
            int id = fm.table.find(key);
            if (id < 0) {
                id = fm.nReal + fm.nGen;
                fm.table.insert(key, id);
                fm.nGen++;
                fm.genParent.push_back(prev);
                fm.genScale.push_back(s);
                fm.genL.push_back(ls);
                int child = 0;
                for (int d = 0; d < 3; d++) child |= (ls[d] & 1) << d;
                made.push_back({id, prev, child, s});
            }
            prev = id;
        }
        pf[p] = prev;
    }
    std::vector<device::GenNode> plan;
    plan.reserve(made.size());
    for (const auto &m : made) plan.push_back({m.id, m.parent, m.child, m.scale});
    return plan;
}

// Runs body(i) for i in [0, n) on the host threads when n >= 4 (two cores
// left to the runtime's own threads, which feed the GPU); the first exception
// is rethrown once all have finished.
template <typename F> void host_parallel(int n, F &&body) {
    if (n < 4) { // no thread team for a few items
        for (int i = 0; i < n; i++) body(i);
        return;
    }
    std::exception_ptr err;
    std::mutex m;
    const int threads = std::max(1, std::min(n, omp_get_max_threads() - 2));
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int i = 0; i < n; i++) {
        try {
            body(i);
        } catch (...) {
            std::lock_guard<std::mutex> lk(m);
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
}

// A run of one level batch: nodes of one tree at one scale.
struct SegmentSpec {
    int job, scale;
    std::vector<int> nodes;
};

// The next level of one plain tree formed and screened on the device before
// the host has split the current one.  Right after a level's band sums,
// mwg_next_level forms its split decisions and the next level's node data,
// and the pair screen of the next level runs over an upper bound of its node
// count (eight per node) with the device's count, so that one download brings
// back the level's norms, the device's decisions and the next level's screen
// counters.  The host then splits with its own arithmetic and takes the
// device's level only when both agree bit for bit; otherwise that level is
// formed and screened again from the host's split.
constexpr int kAheadMaxNodes = 2048; // the device sums a level's norms sequentially
struct Ahead {
    bool want = false;  // in: form and screen the next level on the device
    bool ready = false; // in: this level was formed and screened ahead (device set `cur`)
    int cur = 0;
    bool formed = false; // out: the next level was formed (device set `set`)
    int set = 0;
    long long base = 0;      // first child id the device used
    std::vector<int> nodes;  // the level's nodes, in the order of the device's decisions
    std::vector<int> split;  // the device's split decisions
    double k1[4] = {};       // tree square norm, next screening threshold, wavelet norm sum, children
    std::vector<int> head;   // the next level's screen counters and first missing-node records
    mwg_pairs pairs{};       // the arrays that screen wrote
};

// Whether two bindings name the same pair arrays and sizes.
bool same_arrays(const mwg_pairs &a, const mwg_pairs &b) {
    return a.pair_g == b.pair_g && a.pair_f == b.pair_f && a.pair_l == b.pair_l && a.g_first == b.g_first && a.g_npairs == b.g_npairs &&
           a.missing == b.missing && a.missing_pairs == b.missing_pairs && a.dedup_keys == b.dedup_keys && a.dedup_rec == b.dedup_rec &&
           a.count == b.count && a.capacity == b.capacity && a.missing_cap == b.missing_cap && a.dedup_cap == b.dedup_cap;
}

mwg_segment make_segment(Job &job, int scale, int first, int count, unsigned long long *counters) {
    mwg_segment s{};
    s.first = first;
    s.count = count;
    s.scale = scale;
    s.o_depth = scale - job.task.oper->getOperatorRoot();
    const detail::BandTables &ht = *job.ht;
    const bool empty_band = s.o_depth < 0 || s.o_depth >= ht.nDepth || ht.bandMax[s.o_depth] < 0;
    s.band_w = empty_band ? -1 : ht.bandMax[s.o_depth];
    s.g_thr = job.gthr;
    s.f = job.fm->view(*job.mra);
    s.op = *job.bt;
    s.g_coef = job.task.out->deviceCoefs().get();
    s.counters = counters;
    return s;
}

// S.missing: the screen's counters (4 ints) then the missing-node records
// (6 ints each), so one download returns the counts and the first records.
constexpr size_t kHead = 4, kEagerRecords = 512;

// After the band sums of a level of one plain tree (its norms in S.gNorms):
// form the next level on the device, screen it over an upper bound of its
// node count, and bring back the level's norms (into nrm), the device's split
// decisions and the next level's screen head with one wait.
void form_ahead(Job &job, const Shape &sh, const SegmentSpec &sp, const mwg_level &lvl, mwg_pairs &pr, unsigned long long *counters, Ahead &ah,
                std::vector<double> &nrm) {
    Scratch &S = scratch();
    const int L = lvl.count, Lmax = 8 * L;
    const int nxt = ah.ready ? 1 - ah.cur : 0;
    S.nextOff[nxt].reserve(Lmax);
    S.nextL[nxt].reserve(3 * (size_t)Lmax);
    S.nextThr[nxt].reserve(Lmax);
    S.split.reserve(L);
    S.nextCount.reserve(1);
    S.nextHead.reserve(4);
    mwg_next_level_args a{};
    a.count = L;
    a.may_split = 1;
    a.abs_prec = job.task.absPrec ? 1 : 0;
    a.prec = job.task.prec;
    a.s_norm = job.sNorm;
    a.w_norm = job.wNorm;
    a.scale_fac = detail::wavelet_scale_factor(sp.scale);
    a.n_terms = (double)job.task.oper->size();
    a.thr_floor = 2.0 * MachinePrec;
    a.base_id = job.g->size();
    a.ncoef = sh.ncoef;
    a.g_norms = lvl.g_norms;
    a.l = lvl.l;
    a.head = S.nextHead.get();
    a.split = S.split.get();
    a.next_off = S.nextOff[nxt].get();
    a.next_l = S.nextL[nxt].get();
    a.next_thr = S.nextThr[nxt].get();
    a.next_count = S.nextCount.get();
    device::check(mwg_next_level(&a), "next level");
    // the next level: one segment of at most Lmax nodes, per-node thresholds
    std::vector<mwg_segment> seg = {make_segment(job, sp.scale + 1, 0, Lmax, counters)};
    mwg_level nl{};
    nl.kp1 = sh.kp1;
    nl.count = Lmax;
    nl.nseg = 1;
    nl.segs = seg.data();
    void *unused = nullptr;
    device::check(mwg_level_upload(&nl, nullptr, 0, &unused), "level upload");
    nl.g_off = a.next_off;
    nl.l = a.next_l;
    nl.node_thr = a.next_thr;
    nl.g_norms = lvl.g_norms;
    nl.count_dev = a.next_count;
    // per-node pair lists for the bound (the pair arrays keep their size;
    // the work still queued reads the old blocks)
    S.gFirst.reserve(Lmax);
    S.gCount.reserve(Lmax);
    pr.g_first = S.gFirst.get();
    pr.g_npairs = S.gCount.get();
    S.dedupClean = false;
    device::check(mwg_pair_screen(&nl, &pr), "pair screen");
    ah.split.resize(L);
    ah.head.resize(kHead + 6 * std::min<size_t>(kEagerRecords, (size_t)pr.missing_cap));
    void *dst[4] = {nrm.data(), ah.k1, ah.split.data(), ah.head.data()};
    const void *src[4] = {lvl.g_norms, a.head, a.split, S.missing.get()};
    const size_t bytes[4] = {sizeof(double) * nrm.size(), sizeof(ah.k1), sizeof(int) * ah.split.size(), sizeof(int) * ah.head.size()};
    device::check(mwg_download_many(4, dst, src, bytes), "download");
    ah.formed = true;
    ah.pairs = pr;
    ah.set = nxt;
    ah.base = a.base_id;
    ah.nodes = sp.nodes;
}

// Screen, generate, band-sum and reduce one level batch of segments; writes
// the computed nodes' norms back into their trees.  With `ahead` (one plain
// tree, one segment) the level may come formed and screened from the device,
// and the next level is formed and screened before the norms come back.
void run_level_batch(std::vector<Job> &jobs, const Shape &sh, const std::vector<SegmentSpec> &specs, unsigned long long *counters,
                     Ahead *ahead = nullptr) {
    Scratch &S = scratch();
    const bool ready = ahead && ahead->ready;
    int L = 0;
    for (const auto &sp : specs) L += (int)sp.nodes.size();
    std::vector<int> lv(3 * (size_t)L), nodeJob(L);
    std::vector<long long> off(L);
    std::vector<mwg_segment> segs;
    int at = 0;
    for (const auto &sp : specs) {
        Job &job = jobs[sp.job];
        segs.push_back(make_segment(job, sp.scale, at, (int)sp.nodes.size(), counters + 2 * sp.job));
        for (int u : sp.nodes) {
            for (int d = 0; d < 3; d++) lv[3 * at + d] = job.g->transl(u)[d];
            off[at] = (long long)u * sh.ncoef;
            nodeJob[at] = sp.job;
            at++;
        }
    }
    bool perNode = false, margins = false;
    for (const auto &sp : specs) {
        perNode = perNode || jobs[sp.job].precFunc != nullptr;
        margins = margins || jobs[sp.job].marginPass;
    }
    // one upload: g_off [L] | node thresholds [L] (precision trees only) | translations [3L ints]
    const size_t nThr = perNode ? (size_t)L : 0, nL8 = (3 * (size_t)L + 1) / 2;
    std::vector<long long> packed((size_t)L + nThr + nL8, 0);
    std::memcpy(packed.data(), off.data(), sizeof(long long) * L);
    if (perNode) {
        double *thr = reinterpret_cast<double *>(packed.data() + L);
        size_t i = 0;
        for (const auto &sp : specs)
            for (int u : sp.nodes) thr[i++] = jobs[sp.job].precFunc ? node_threshold(jobs[sp.job], u) : jobs[sp.job].gthr;
    }
    std::memcpy(packed.data() + L + nThr, lv.data(), sizeof(int) * lv.size());
    S.gNorms.reserve(9 * (size_t)L);
    S.gFirst.reserve(L);
    S.gCount.reserve(L);

    mwg_level lvl{};
    lvl.kp1 = sh.kp1;
    lvl.count = L;
    lvl.nseg = (int)segs.size();
    lvl.segs = segs.data();
    void *lvlDev = nullptr;
    if (ready) {
        // offsets, translations and screening thresholds formed on the device
        device::check(mwg_level_upload(&lvl, nullptr, 0, &lvlDev), "level upload");
        lvl.g_off = S.nextOff[ahead->cur].get();
        lvl.node_thr = S.nextThr[ahead->cur].get();
        lvl.l = S.nextL[ahead->cur].get();
    } else {
        // the level data goes with the segment tables in one upload
        device::check(mwg_level_upload(&lvl, packed.data(), sizeof(long long) * packed.size(), &lvlDev), "level upload");
        const long long *lvlData = static_cast<const long long *>(lvlDev);
        lvl.g_off = lvlData;
        lvl.node_thr = perNode ? reinterpret_cast<const double *>(lvlData + L) : nullptr;
        lvl.l = reinterpret_cast<const int *>(lvlData + L + nThr);
    }
    lvl.g_norms = S.gNorms.get();
    if (margins) {
        S.gMargin.reserve(MWG_MARGINS * (size_t)L);
        lvl.g_margin = S.gMargin.get();
        for (int j = 0; j < MWG_MARGINS; j++) lvl.margin_ratio[j] = kMarginRatio[j];
    }

    mwg_pairs pr{};
    int npairs = 0, nmissing = 0, nMissingPairs = 0;
    // the pair capacity the arrays hold now (binding it grows nothing)
    size_t cap = std::max<size_t>(std::min(S.pairF.capacity(), S.pairL.capacity() / 3), 4096);
    size_t mcap = std::max<size_t>(S.missing.capacity() > kHead ? (S.missing.capacity() - kHead) / 6 : 0, 1024);
    auto bind_pairs = [&](bool keepTable) {
        S.pairG.reserve(cap);
        S.pairF.reserve(cap);
        S.pairL.reserve(3 * cap);
        S.missing.reserve(kHead + 6 * mcap);
        pr.capacity = (int)std::min<size_t>(S.pairF.capacity(), S.pairL.capacity() / 3);
        pr.count = S.missing.get();
        pr.pair_g = S.pairG.get();
        pr.pair_f = S.pairF.get();
        pr.pair_l = S.pairL.get();
        pr.g_first = S.gFirst.get();
        pr.g_npairs = S.gCount.get();
        pr.missing = S.missing.get() + kHead;
        pr.missing_cap = (int)((S.missing.capacity() - kHead) / 6);
        S.missingPairs.reserve(2 * (size_t)pr.capacity);
        pr.missing_pairs = S.missingPairs.get();
        // deduplication table of missing band nodes: at least twice the pair
        // capacity (so every probe ends), zero before each screen
        size_t dcap = 1024;
        while (dcap < 2 * (size_t)pr.capacity) dcap *= 2;
        if (S.dedupKeys.capacity() < dcap) {
            S.dedupKeys.reserve(dcap);
            S.dedupRec.reserve(dcap);
            S.dedupClean = false;
        }
        pr.dedup_keys = S.dedupKeys.get();
        pr.dedup_rec = S.dedupRec.get();
        pr.dedup_cap = (int)dcap;
        if (keepTable) return; // the slots claimed by the level's own screen
        if (!S.dedupClean) S.dedupKeys.zero(S.dedupKeys.capacity());
        S.dedupClean = true;
    };
    // a level screened ahead left its pairs, records and table slots in
    // these same arrays (none grows here; if one had moved, the level is
    // screened again on a cleared table)
    bind_pairs(ready);
    bool screened = ready;
    if (ready && !same_arrays(pr, ahead->pairs)) {
        screened = false;
        bind_pairs(false);
    }
    auto t0 = clk::now();
    std::vector<int> head;
    for (;;) {
        if (screened) {
            head = ahead->head; // came back with the previous level's norms
            screened = false;
        } else {
            // a screen claims table slots for missing band nodes; the patch of
            // their pairs returns them to zero
            S.dedupClean = false;
            device::check(mwg_pair_screen(&lvl, &pr), "pair screen");
            head.resize(kHead + 6 * std::min<size_t>(kEagerRecords, (size_t)pr.missing_cap));
            S.missing.download(head.data(), head.size());
        }
        if (head[2]) throw Error("pair screen: band width does not match the operator tables");
        npairs = head[0];
        nmissing = head[1];
        nMissingPairs = head[3];
        if (npairs <= pr.capacity && nmissing <= pr.missing_cap) {
            if (nMissingPairs == 0) S.dedupClean = true;
            break;
        }
        cap = std::max(cap, (size_t)npairs + npairs / 4);
        mcap = std::max(mcap, (size_t)nmissing + nmissing / 4);
        bind_pairs(false);
    }
    const double tScreen = since(t0);
    double tGenerate = 0.0;
    if (npairs > 0 && nmissing > 0) {
        t0 = clk::now();
        // the pairs whose band node is missing: (pair, ancestor code, node, translation)
        std::vector<int> rec(6 * (size_t)nmissing);
        if ((size_t)nmissing <= kEagerRecords) std::copy(head.begin() + kHead, head.begin() + kHead + rec.size(), rec.begin());
        else S.missing.download(rec.data(), rec.size(), kHead);
        std::vector<int> segOf(L);
        for (size_t s = 0; s < segs.size(); s++)
            for (int i = 0; i < segs[s].count; i++) segOf[segs[s].first + i] = (int)s;
        // The screen gives one record per distinct missing band node of a
        // segment (all pairs of that node share it).  Records are planned in
        // this deterministic order; records naming the same node (segments
        // of one tree and scale, or nodes beyond the screen's key range) are
        // planned once.
        struct Miss {
            int job, anc, scale, l[3], rec;
        };
        // records grouped by tree (in record order), then each tree's sorted
        const int J = (int)jobs.size();
        std::vector<size_t> jb(J + 1, 0);
        for (int i = 0; i < nmissing; i++) jb[nodeJob[rec[6 * (size_t)i + 2]] + 1]++;
        for (int j = 0; j < J; j++) jb[j + 1] += jb[j];
        std::vector<Miss> ms(nmissing);
        {
            std::vector<size_t> put(jb.begin(), jb.end() - 1);
            for (int i = 0; i < nmissing; i++) {
                const int *r = rec.data() + 6 * (size_t)i;
                const int j = nodeJob[r[2]];
                ms[put[j]++] = Miss{j, r[1], segs[segOf[r[2]]].scale, {r[3], r[4], r[5]}, i};
            }
        }
        auto same = [](const Miss &a, const Miss &b) {
            return a.job == b.job && a.anc == b.anc && a.scale == b.scale && a.l[0] == b.l[0] && a.l[1] == b.l[1] && a.l[2] == b.l[2];
        };
        // trees sharing an input mirror are planned in tree order by one thread
        std::vector<int> withRecords;
        for (int j = 0; j < J; j++)
            if (jb[j + 1] > jb[j]) withRecords.push_back(j);
        std::vector<std::vector<int>> byMirror;
        {
            std::map<const device::TreeMirror *, size_t> slot;
            for (int j : withRecords) {
                auto it = slot.emplace(jobs[j].fm, byMirror.size()).first;
                if (it->second == byMirror.size()) byMirror.emplace_back();
                byMirror[it->second].push_back(j);
            }
        }
        std::vector<int> nodeOfRecord(nmissing);
        std::vector<std::vector<device::GenNode>> made(J);
        host_parallel((int)byMirror.size(), [&](int m) {
            for (int job : byMirror[m]) {
                const auto lo = ms.begin() + jb[job], hi = ms.begin() + jb[job + 1];
                std::sort(lo, hi, [](const Miss &a, const Miss &b) {
                    if (a.scale != b.scale) return a.scale < b.scale;
                    if (a.anc != b.anc) return a.anc < b.anc;
                    for (int d = 0; d < 3; d++)
                        if (a.l[d] != b.l[d]) return a.l[d] < b.l[d];
                    return a.rec < b.rec;
                });
                // the distinct band nodes of the tree
                std::vector<int> f, l3, sc;
                std::vector<size_t> first;
                const size_t a = jb[job], b = jb[job + 1];
                for (size_t i = a; i < b; i++)
                    if (i == a || !same(ms[i], ms[i - 1])) {
                        first.push_back(i);
                        f.push_back(ms[i].anc);
                        l3.insert(l3.end(), ms[i].l, ms[i].l + 3);
                        sc.push_back(ms[i].scale);
                    }
                made[job] = plan_band_nodes(jobs[job], f, l3, sc);
                for (size_t k = 0; k < first.size(); k++) {
                    const size_t e = k + 1 < first.size() ? first[k + 1] : b;
                    for (size_t i = first[k]; i < e; i++) nodeOfRecord[ms[i].rec] = f[k];
                }
            }
        });
        device::MirrorBatch batch;
        for (int j : withRecords) {
            if (made[j].empty()) continue;
            jobs[j].stats.nGenerated += (long long)made[j].size();
            batch.add(*jobs[j].fm, *jobs[j].mra, jobs[j].filters, made[j]);
        }
        // one upload for every tree's new nodes and the node of every record
        const size_t at = batch.addInts(nodeOfRecord.data(), nodeOfRecord.size());
        const int *nodes = batch.run(S.gen) + at;
        device::check(mwg_patch_pairs(&pr, nodes, nMissingPairs), "patch pairs");
        S.dedupClean = true;
        // mirrors may have moved: refresh every segment's input view
        for (size_t s = 0; s < segs.size(); s++) segs[s].f = jobs[nodeJob[segs[s].first]].fm->view(*jobs[nodeJob[segs[s].first]].mra);
        tGenerate = since(t0);
    }
    t0 = clk::now();
    // band sums, output coefficients and norms (zero when no pair survives)
    device::check(mwg_band_apply(&lvl, &pr, npairs, (int)g_sweep), "band apply");
    std::vector<double> nrm(9 * (size_t)L), mgn(margins ? MWG_MARGINS * (size_t)L : 0);
    if (ahead && ahead->want) {
        form_ahead(jobs[specs[0].job], sh, specs[0], lvl, pr, counters + 2 * specs[0].job, *ahead, nrm);
    } else {
        S.gNorms.download(nrm.data(), nrm.size());
        if (margins) S.gMargin.download(mgn.data(), mgn.size());
    }
    const double tBand = since(t0);
    std::set<int> touched;
    for (const auto &sp : specs) touched.insert(sp.job);
    for (int j : touched) {
        jobs[j].stats.tScreen += tScreen;
        jobs[j].stats.tGenerate += tGenerate;
        jobs[j].stats.tBand += tBand;
    }
    // the norms back into the trees, segments on the host threads (each
    // writes its own nodes; per-tree arrays sized first)
    std::vector<int> segAt(specs.size());
    at = 0;
    for (size_t s = 0; s < specs.size(); s++) {
        segAt[s] = at;
        at += (int)specs[s].nodes.size();
        Job &job = jobs[specs[s].job];
        if (job.marginPass) {
            job.margin.resize(job.g->size());
            job.thrUsed.resize(job.g->size(), -1.0);
        }
    }
    host_parallel((int)specs.size(), [&](int s) {
        const SegmentSpec &sp = specs[s];
        Job &job = jobs[sp.job];
        detail::NodeStore<3> &g = *job.g;
        int k = segAt[s];
        for (int u : sp.nodes) {
            double *cn = g.norms(u);
            for (int t = 0; t < 8; t++) cn[t] = nrm[9 * (size_t)k + t];
            g.setSquareNorm(u, nrm[9 * (size_t)k + 8]);
            g.setHasCoefs(u, true);
            if (job.marginPass) {
                for (int j = 0; j < MWG_MARGINS; j++) job.margin[u][j] = mgn[MWG_MARGINS * (size_t)k + j];
                job.thrUsed[u] = job.precFunc ? node_threshold(job, u) : job.gthr;
            }
            k++;
        }
    });
}

// Device bytes of one output node in a level batch (screen candidates and group lists).
double node_budget(const Job &job, int scale) {
    const int o_depth = scale - job.task.oper->getOperatorRoot();
    const detail::BandTables &ht = *job.ht;
    const int W = (o_depth >= 0 && o_depth < ht.nDepth) ? std::max(ht.bandMax[o_depth], 0) : 0;
    const double box = std::pow(2.0 * W + 1.0, 3.0);
    return 8.0 * (box + 256.0) + 128.0 * (96.0 * ht.M + 48.0); // screen candidates; per pair: group codes, level-order records, pair record
}

// Screening threshold of a step from the approximate tree norm
// (ConvolutionCalculator::calcNode, src/treebuilders/ConvolutionCalculator.cpp:241-246).
double screening_threshold(const Job &job) {
    double t = job.g->treeSquareNorm;
    if (t > 0.0) t = job.task.prec * 1.0 * std::sqrt(t / (double)job.task.oper->size());
    return t;
}

// Whether the next level of this step's one segment can be formed and
// screened on the device: a plain tree past its first step (the scaling norm
// is known), a step that may split, at most kAheadMaxNodes nodes, none with
// children yet.
bool ahead_eligible(const Job &job, const SegmentSpec &sp) {
    const detail::NodeStore<3> &g = *job.g;
    if (job.mode != Mode::Plain || job.precFunc || job.task.predictor || job.task.samples || job.marginPass) return false;
    if (job.iter < 1 || sp.nodes.empty() || (int)sp.nodes.size() > kAheadMaxNodes) return false;
    if (job.iter >= job.task.maxIter && job.task.maxIter >= 0) return false;
    if (sp.scale + 2 > job.mra->getMaxScale() || sp.scale + 1 - job.task.oper->getOperatorRoot() > mwgkey::kMaxDepth) return false;
    for (int u : sp.nodes)
        if (g.hasChildren(u) || g.depth(u) < 0) return false;
    return true;
}

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// After the host's split of a level formed ahead: whether the device formed
// the same next level (children, ids, approximate norm and threshold).
bool ahead_agrees(const Job &job, const Ahead &ah) {
    const detail::NodeStore<3> &g = *job.g;
    if ((double)job.work.size() != ah.k1[3] || ah.split.size() != ah.nodes.size()) return false;
    for (size_t i = 0; i < job.work.size(); i++)
        if (job.work[i] != ah.base + (long long)i) return false;
    for (size_t i = 0; i < ah.nodes.size(); i++)
        if ((g.hasChildren(ah.nodes[i]) ? 1 : 0) != ah.split[i]) return false;
    return same_bits(g.treeSquareNorm, ah.k1[0]) && same_bits(screening_threshold(job), ah.k1[1]) && same_bits(job.wNorm, ah.k1[2]);
}

// The final transforms of every tree of a batch together: TopDown
// (MWTree::mwTransformDown, overwrite=false) from the root depth by depth,
// then BottomUp (MWTree::mwTransformUp), one launch per depth for the parents
// of all trees, then the component norms of every node in one launch.  Node
// offsets are taken from the lowest output pool of the batch (each tree keeps
// its own pool); trees are independent, so merging them changes no result.
void transform_and_norms(std::vector<Job> &jobs, const Shape &sh) {
    Scratch &S = scratch();
    double *base = nullptr;
    for (auto &job : jobs) {
        double *c = job.task.out->deviceCoefs().get();
        if (c && (!base || c < base)) base = c;
    }
    if (!base) return;
    std::vector<std::vector<std::vector<int>>> br(jobs.size()); // [job][depth] parents with children
    std::vector<long long> poolOff(jobs.size());
    size_t depths = 0;
    std::vector<std::vector<int>> all(jobs.size()); // every real node of each tree
    host_parallel((int)jobs.size(), [&](int j) {
        const detail::NodeStore<3> &g = *jobs[j].g;
        const auto tab = detail::real_nodes_by_depth(g);
        br[j].resize(tab.size());
        for (size_t d = 0; d < tab.size(); d++)
            for (int u : tab[d])
                if (g.hasRealChildren(u)) br[j][d].push_back(u);
        all[j] = detail::real_nodes(g);
    });
    for (size_t j = 0; j < jobs.size(); j++) {
        poolOff[j] = (long long)(((uintptr_t)jobs[j].task.out->deviceCoefs().get() - (uintptr_t)base) / sizeof(double));
        depths = std::max(depths, br[j].size());
    }
    struct Pass {
        size_t at, n;
        bool up;
    };
    std::vector<Pass> passes;
    std::vector<long long> po, co;
    auto add = [&](size_t d, bool up) {
        const size_t at = po.size();
        for (size_t j = 0; j < jobs.size(); j++) {
            if (d >= br[j].size()) continue;
            const detail::NodeStore<3> &g = *jobs[j].g;
            for (int u : br[j][d]) {
                po.push_back(poolOff[j] + (long long)u * sh.ncoef);
                co.push_back(poolOff[j] + (long long)g.child0(u) * sh.ncoef);
            }
        }
        if (po.size() > at) passes.push_back({at, po.size() - at, up});
    };
    for (size_t d = 0; d < depths; d++) add(d, false);
    for (int d = (int)depths - 2; d >= 0; d--) add((size_t)d, true);
    // then the component norms of every node: offsets appended to po
    const size_t normAt = po.size();
    for (size_t j = 0; j < jobs.size(); j++)
        for (int u : all[j]) po.push_back(poolOff[j] + (long long)u * sh.ncoef);
    const size_t nAll = po.size() - normAt;
    S.pOff.reserve(po.size());
    S.cOff.reserve(std::max<size_t>(co.size(), 1));
    S.pOff.upload(po.data(), po.size());
    S.cOff.upload(co.data(), co.size());
    const double *filters = jobs[0].filters;
    for (const Pass &p : passes) {
        if (p.up)
            device::check(mwg_compress(sh.kp1, filters, base, S.pOff.get() + p.at, S.cOff.get() + p.at, (int)p.n, sh.ncoef), "compress");
        else
            device::check(mwg_reconstruct_add(sh.kp1, filters, base, S.pOff.get() + p.at, S.cOff.get() + p.at, (int)p.n, sh.ncoef),
                          "reconstruct");
    }
    S.gNorms.reserve(9 * nAll);
    device::check(mwg_node_norms(sh.kp1, base, S.pOff.get() + normAt, S.gNorms.get(), (int)nAll), "norms");
    std::vector<double> nrm(9 * nAll);
    S.gNorms.download(nrm.data(), nrm.size());
    size_t k = 0;
    for (size_t j = 0; j < jobs.size(); j++) {
        detail::NodeStore<3> &g = *jobs[j].g;
        for (int u : all[j]) {
            double *cn = g.norms(u);
            for (int t = 0; t < 8; t++) cn[t] = nrm[9 * k + t];
            g.setSquareNorm(u, nrm[9 * k + 8]);
            k++;
        }
    }
}


// TreeBuilder::calcWaveletNorm (src/treebuilders/TreeBuilder.cpp:156-163):
// the wavelet square norms of one step's nodes at depth >= 0, in work order.
double step_wavelet_norm(const detail::NodeStore<3> &g, const std::vector<int> &work) {
    double w = 0.0;
    for (int u : work)
        if (g.depth(u) >= 0) w += g.waveletNorm(u);
    return w;
}

// TreeBuilder::build bookkeeping after a step: approximate tree norm and
// WaveletAdaptor splits; returns the next work vector.
void finish_step(Job &job) {
    detail::NodeStore<3> &g = *job.g;
    const MultiResolutionAnalysis<3> &mra = *job.mra;
    if (job.iter == 0) {
        job.sNorm = 0.0;
        for (int u : job.work)
            if (g.depth(u) >= 0) job.sNorm += g.scalingNorm(u);
    }
    // the step's wavelet norms summed first, then added once
    // (wNorm += calcWaveletNorm(*workVec), src/treebuilders/TreeBuilder.cpp:58)
    job.wNorm += step_wavelet_norm(g, job.work);
    g.treeSquareNorm = (job.sNorm < 0.0 || job.wNorm < 0.0) ? -1.0 : job.sNorm + job.wNorm;
    if (job.iter >= job.task.maxIter && job.task.maxIter >= 0) job.work.clear();
    std::vector<int> next;
    // without precision trees the split threshold depends on the scale only
    // (the same value wavelet_split forms, computed once per scale)
    int thrScale = INT_MIN;
    double thr = 0.0;
    for (int u : job.work) {
        if (g.hasChildren(u)) continue;
        const int sc = g.scale(u);
        if (sc + 2 > mra.getMaxScale()) continue;
        bool split;
        if (job.precFunc) {
            split = detail::wavelet_split(node_prec(job, u), g.treeSquareNorm, job.task.absPrec, sc, g.waveletNorm(u));
        } else {
            if (sc != thrScale) {
                thrScale = sc;
                thr = detail::wavelet_threshold(job.task.prec, g.treeSquareNorm, job.task.absPrec, sc);
            }
            split = job.task.prec > 0.0 && std::sqrt(g.waveletNorm(u)) > thr;
        }
        if (split) {
            const int c0 = g.createChildren(u, false);
            for (int k = 0; k < 8; k++) next.push_back(c0 + k);
        }
    }
    job.work.swap(next);
    job.iter++;
}

// Breadth-first steps computed exactly before the predictor grows the grid
// (steps 0 .. kAnchorStep): every predicted node then has a computed parent.
constexpr int kAnchorStep = 1;

// Predictor inputs: the wavelet norms of the nodes computed in the exact
// steps (called after each of them, before the split).
void record_exact_step(Job &job) {
    PredictorContext &c = job.ctx;
    if (job.iter == 0) {
        c.inp = job.fstore;
        c.out = job.g;
        c.bands = job.ht;
        c.mra = job.mra;
        c.prec = job.task.prec;
        c.mu = job.task.oper->kernelMu();
        c.inpNorm = job.task.inp->getSquareNorm();
        c.centers = job.task.centers;
    }
    c.firstStep.resize(job.g->size(), 0);
    c.firstWavelet.resize(job.g->size(), 0.0);
    for (int u : job.work) {
        c.firstStep[u] = 1;
        c.firstWavelet[u] = job.g->waveletNorm(u);
    }
}

// Grow the output grid below the (certain) nodes of the step after the anchor
// wherever the predictor expects a refinement; the next pass computes all of them.
void pregrow(Job &job) {
    detail::NodeStore<3> &g = *job.g;
    const int maxScale = job.mra->getMaxScale();
    // predicted nodes stay within the input tree's depth (and the device key
    // range) and a budget of four times the input's nodes; anything deeper the
    // wavelet test asks for is computed in fallback passes
    const int maxDepth = std::min(job.fstore->getDepth() - 1, mwgkey::kMaxDepth - 1);
    const long long budget = 4ll * job.fm->nReal;
    job.replayWork = job.work;
    std::vector<int> all = job.work, cand = job.work;
    std::vector<float> feat;
    while (!cand.empty() && job.stats.nPredicted < budget) {
        feat.resize(cand.size() * kPredictorFeatures);
        for (size_t i = 0; i < cand.size(); i++) predictor_features(job.ctx, cand[i], feat.data() + i * kPredictorFeatures);
        const std::vector<float> prob = job.task.predictor->predict(feat.data(), (int)cand.size());
        std::vector<int> next;
        for (size_t i = 0; i < cand.size(); i++) {
            const int u = cand[i];
            if (prob[i] <= job.task.predictor->threshold || g.scale(u) + 2 > maxScale || g.depth(u) + 1 > maxDepth) continue;
            const int c0 = g.createChildren(u, false);
            for (int k = 0; k < 8; k++) next.push_back(c0 + k);
        }
        job.speculative.resize(g.size(), 0);
        for (int v : next) job.speculative[v] = 1;
        job.stats.nPredicted += (long long)next.size();
        all.insert(all.end(), next.begin(), next.end());
        cand.swap(next);
    }
    job.work.swap(all);
    job.mode = Mode::Replay;
    job.marginPass = true;
}

long long subtree_size(const detail::NodeStore<3> &g, int u) {
    long long n = 0;
    std::vector<int> st = {u};
    while (!st.empty()) {
        const int v = st.back();
        st.pop_back();
        if (!g.hasChildren(v)) continue;
        for (int c = 0; c < 8; c++) {
            n++;
            st.push_back(g.child0(v) + c);
        }
    }
    return n;
}

// Bound on how far speculative node u (computed with threshold thrUsed[u])
// can be from its computation with its own step's screening threshold: the
// tuples the higher threshold drops each add at most |Ox||Oy||Oz| |f_ft| in
// 2-norm (operator norms are Frobenius, which bound the spectral norm), and the
// margin sums exactly those over every tuple below ratio * thrUsed.  Infinite
// when no recorded ratio covers the step's threshold.
double speculative_bound(const Job &job, int u) {
    const double used = job.thrUsed[u], own = node_threshold(job, u);
    if (!(used > 0.0) || !(own >= used)) return HUGE_VAL;
    const double ratio = own / used;
    if (ratio == 1.0) return 0.0; // the same threshold: the same tuples
    for (int j = 0; j < MWG_MARGINS; j++)
        if (kMarginRatio[j] >= ratio * (1.0 + 1e-12)) return job.margin[u][j];
    return HUGE_VAL;
}

bool is_speculative(const Job &job, int u) { return u < (int)job.speculative.size() && job.speculative[u]; }

// The breadth-first steps of TreeBuilder::build replayed on computed nodes:
// the wavelet test decides every split with the running approximate norm;
// predicted children it confirms are kept, the others removed; children it
// wants that were not predicted go to job.work for the next pass.
void replay(Job &job) {
    detail::NodeStore<3> &g = *job.g;
    const int maxScale = job.mra->getMaxScale();
    job.marginPass = false;
    job.work.clear();
    while (!job.replayWork.empty()) {
        // speculative nodes of this step and their bounds (own threshold:
        // the approximate norm after the previous step, still in treeSquareNorm)
        std::vector<std::pair<int, double>> spec;
        for (int u : job.replayWork)
            if (is_speculative(job, u)) spec.push_back({u, speculative_bound(job, u)});
        auto err_of = [&](int u, double B) { return g.depth(u) >= 0 ? (2.0 * std::sqrt(g.waveletNorm(u)) + B) * B : 0.0; };
        if (!job.edgeChecked) {
            // A speculative node whose split could differ from the one its
            // recomputation would get is recomputed with this step's screening
            // threshold first: its wavelet norm is within B of the recomputed
            // one, and the split threshold within dthr of the plain method's
            // (the approximate norm holds every kept speculative node's error).
            const double w = job.wNorm + step_wavelet_norm(g, job.replayWork);
            const double norm = (job.sNorm < 0.0 || w < 0.0) ? -1.0 : job.sNorm + w;
            double err = job.normErr;
            for (const auto &sb : spec) err += err_of(sb.first, sb.second);
            std::vector<int> edge;
            for (const auto &sb : spec) {
                const int u = sb.first;
                const double B = sb.second;
                const double thr = detail::wavelet_threshold(node_prec(job, u), norm, job.task.absPrec, g.scale(u));
                const double dthr = (!job.task.absPrec && norm > 0.0) ? thr * (std::sqrt(1.0 + err / norm) - 1.0) : 0.0;
                if (!std::isfinite(B) || !std::isfinite(dthr) || std::fabs(std::sqrt(g.waveletNorm(u)) - thr) <= B * (1.0 + 1e-9) + dthr + 1e-12 * thr)
                    edge.push_back(u);
            }
            job.edgeChecked = true;
            if (!edge.empty()) {
                for (int u : edge) job.speculative[u] = 0;
                job.stats.nFallback += (long long)edge.size();
                job.work.swap(edge);
                return;
            }
        }
        job.edgeChecked = false;
        for (const auto &sb : spec)
            if (is_speculative(job, sb.first)) job.normErr += err_of(sb.first, sb.second);
        job.wNorm += step_wavelet_norm(g, job.replayWork);
        g.treeSquareNorm = (job.sNorm < 0.0 || job.wNorm < 0.0) ? -1.0 : job.sNorm + job.wNorm;
        const bool stop = job.iter >= job.task.maxIter && job.task.maxIter >= 0;
        std::vector<int> next, pending;
        for (int u : job.replayWork) {
            const bool split = !stop && g.scale(u) + 2 <= maxScale &&
                               detail::wavelet_split(node_prec(job, u), g.treeSquareNorm, job.task.absPrec, g.scale(u), g.waveletNorm(u));
            if (split) {
                if (!g.hasChildren(u)) {
                    const int c0 = g.createChildren(u, false);
                    for (int k = 0; k < 8; k++) pending.push_back(c0 + k);
                }
                for (int k = 0; k < 8; k++) next.push_back(g.child0(u) + k);
            } else if (g.hasChildren(u)) {
                job.stats.nCropped += subtree_size(g, u);
                g.detachChildren(u);
            }
        }
        job.iter++;
        job.replayWork.swap(next);
        if (!pending.empty()) {
            // The pass that computes the missed children also recomputes,
            // with the step's own screening threshold (the pass's), every
            // speculative node of the step that might split: the step is then
            // judged without a pass of its own for them.  A node whose wavelet
            // norm stays below the split threshold of the smallest possible
            // approximate norm (the missed children adding nothing, every
            // speculative norm at the low end of its bound) keeps its value;
            // the uncertainty check of its step still judges it.
            std::vector<std::pair<int, double>> spec;
            double low = job.wNorm - job.normErr;
            for (int u : job.replayWork)
                if (is_speculative(job, u)) {
                    const double B = speculative_bound(job, u);
                    spec.push_back({u, B});
                    if (g.depth(u) >= 0 && std::isfinite(B)) low += std::pow(std::max(0.0, std::sqrt(g.waveletNorm(u)) - B), 2);
                }
            const double normLow = (job.sNorm < 0.0 || low < 0.0) ? -1.0 : job.sNorm + low;
            for (const auto &sb : spec) {
                const int u = sb.first;
                const double thrLow = detail::wavelet_threshold(node_prec(job, u), normLow, job.task.absPrec, g.scale(u));
                const bool surelyNoSplit = std::isfinite(sb.second) && normLow > 0.0 && std::sqrt(g.waveletNorm(u)) + sb.second < thrLow;
                if (surelyNoSplit) continue;
                job.speculative[u] = 0;
                pending.push_back(u);
            }
            job.stats.nFallback += (long long)pending.size();
            job.work.swap(pending);
            return;
        }
    }
    job.mode = Mode::Plain;
}

// After a pass computed job.work: the next work vector of the job.
void advance(Job &job) {
    if (job.mode == Mode::Replay) {
        replay(job);
        return;
    }
    const bool learn = (job.task.predictor || job.task.samples) && job.iter <= kAnchorStep;
    const bool anchor = learn && job.iter == kAnchorStep;
    if (learn) record_exact_step(job);
    finish_step(job);
    if (anchor) job.ctx.treeNorm = job.g->treeSquareNorm;
    if (anchor && job.task.predictor && !job.work.empty()) pregrow(job);
}

// Training samples of a finished apply: every output node beyond the exact
// steps, labelled by whether it was refined.
void collect_samples(Job &job) {
    if (!job.task.samples || job.ctx.out == nullptr) return;
    RefineSamples &rs = *job.task.samples;
    const detail::NodeStore<3> &g = *job.g;
    float f[kPredictorFeatures];
    for (int u : detail::real_nodes(g)) {
        if (u < (int)job.ctx.firstStep.size() && job.ctx.firstStep[u]) continue;
        predictor_features(job.ctx, u, f);
        rs.features.insert(rs.features.end(), f, f + kPredictorFeatures);
        rs.labels.push_back(g.hasRealChildren(u) ? 1.0f : 0.0f);
    }
}

} // namespace

const ApplyStats &last_apply_stats() { return g_stats; }
const std::vector<ApplyStats> &last_batch_stats() { return g_batch_stats; }

void set_sweep_precision(SweepPrecision p) { g_sweep = p; }
SweepPrecision sweep_precision() { return g_sweep; }
void set_level_ahead(bool on) { g_ahead = on; }
bool level_ahead() { return g_ahead; }

void apply_batch(const std::vector<ApplyTask> &tasks) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    const auto t_start = clk::now();
    g_batch_stats.clear();
    if (tasks.empty()) return;
    device::ensure_init();
    std::vector<Job> jobs(tasks.size());
    std::set<const FunctionTree<3> *> outs;
    const int kp1 = tasks[0].out->getMRA().getKp1();
    for (size_t i = 0; i < tasks.size(); i++) {
        const ApplyTask &t = tasks[i];
        if (!t.out || !t.oper || !t.inp) throw Error("apply: null tree or operator");
        if (t.out->getMRA() != t.inp->getMRA()) throw Error("Incompatible MRA");
        if (t.out->getMRA().getKp1() != kp1) throw Error("apply_batch: every tree needs the same Legendre order");
        if (t.out == t.inp || !outs.insert(t.out).second) throw Error("apply_batch: each output tree may appear once and differ from its input");
        if (t.precTrees)
            for (const FunctionTree<3> *p : *t.precTrees) {
                if (!p) throw Error("apply: null precision tree");
                if (p->getMRA() != t.out->getMRA()) throw Error("Incompatible MRA");
            }
    }
    for (const ApplyTask &t : tasks)
        if (t.precTrees)
            for (const FunctionTree<3> *p : *t.precTrees)
                if (outs.count(p)) throw Error("apply: an output tree cannot be a precision tree");
    if (kp1 != 6 && kp1 != 8 && kp1 != 10 && kp1 != 4) throw Error("apply supports Legendre orders 3, 5, 7 and 9");
    const Shape sh{kp1, kp1 * kp1 * kp1, 8 * kp1 * kp1 * kp1};
    for (size_t i = 0; i < tasks.size(); i++) {
        Job &job = jobs[i];
        job.task = tasks[i];
        job.mra = &job.task.out->getMRA();
        job.ht = &job.task.oper->bandTables(job.task.prec);
        job.bt = &job.task.oper->deviceTables(job.task.prec)->viewFor((int)g_sweep);
        job.fm = &job.task.inp->inputMirror();
        job.fstore = &job.task.inp->topology();
        job.g = &job.task.out->nodes();
        job.g->reserve(std::max(job.g->size(), 2 * job.fm->nReal)); // node arrays and index without regrowth
        job.filters = device::filters(job.mra->getOrder());
        // output pool sized for an output tree as large as the input (grown on demand)
        job.task.out->deviceCoefs().reserve((size_t)std::max(job.g->size(), job.fm->nReal) * sh.ncoef);
        job.work = detail::real_nodes(*job.g); // ConvolutionCalculator::getInitialWorkVector
        if (job.task.precTrees) job.precFunc = std::make_unique<PrecFactors>(*job.task.precTrees);
    }
    Scratch &S = scratch();
    S.counters.reserve(2 * jobs.size());
    S.counters.zero(2 * jobs.size());
    const double tSetup = since(t_start);

    Ahead ahead; // one plain tree: levels formed and screened on the device
    for (;;) {
        std::vector<int> active;
        for (size_t i = 0; i < jobs.size(); i++)
            if (!jobs[i].work.empty()) active.push_back((int)i);
        if (active.empty()) break;
        // segments: (tree, scale) runs of this step's work, split into level batches by budget
        std::vector<SegmentSpec> specs;
        for (int j : active) {
            Job &job = jobs[j];
            job.stats.levelNodes.push_back((int)job.work.size());
            eval_prec_factors(job);
            job.gthr = screening_threshold(job);
            job.task.out->deviceCoefs().reserve((size_t)job.g->size() * sh.ncoef, (size_t)job.g->size() * sh.ncoef);
            std::map<int, std::vector<int>> byScale;
            for (int u : job.work) byScale[job.g->scale(u)].push_back(u);
            for (auto &kv : byScale) specs.push_back({j, kv.first, std::move(kv.second)});
        }
        // one tree whose step is one segment within the batch budget: the
        // step runs whole, with its next level formed ahead where it can be
        const bool whole =
                jobs.size() == 1 && specs.size() == 1 && node_budget(jobs[0], specs[0].scale) * (double)specs[0].nodes.size() <= kBatchBudget;
        ahead.want = whole && g_ahead && ahead_eligible(jobs[0], specs[0]);
        if (whole) {
            run_level_batch(jobs, sh, specs, S.counters.get(), &ahead);
            if (ahead.ready) jobs[0].stats.nAhead++;
            specs.clear();
        } else {
            ahead.ready = false;
        }
        std::vector<SegmentSpec> batch;
        double used = 0.0;
        auto flush = [&]() {
            if (!batch.empty()) run_level_batch(jobs, sh, batch, S.counters.get());
            batch.clear();
            used = 0.0;
        };
        for (auto &sp : specs) {
            const double per = node_budget(jobs[sp.job], sp.scale);
            size_t i = 0;
            while (i < sp.nodes.size()) {
                size_t room = (size_t)std::max(1.0, (kBatchBudget - used) / per);
                if (used > 0.0 && room < 64 && sp.nodes.size() - i > room) {
                    flush();
                    continue;
                }
                const size_t take = std::min(room, sp.nodes.size() - i);
                batch.push_back({sp.job, sp.scale, std::vector<int>(sp.nodes.begin() + i, sp.nodes.begin() + i + take)});
                used += per * (double)take;
                i += take;
                if (used >= kBatchBudget) flush();
            }
        }
        flush();
        // the trees' split steps are independent: on the host threads when
        // every tree is a plain one (a predictor's steps call the device)
        bool plainAll = true;
        for (int j : active) plainAll = plainAll && jobs[j].mode == Mode::Plain && !jobs[j].task.predictor && !jobs[j].task.samples;
        auto split_step = [&](int i) {
            const auto th = clk::now();
            advance(jobs[active[i]]);
            jobs[active[i]].stats.tHost += since(th);
        };
        if (plainAll) host_parallel((int)active.size(), split_step);
        else
            for (int i = 0; i < (int)active.size(); i++) split_step(i);
        for (int j : active) {
            Job &job = jobs[j];
            if (!job.work.empty()) {
                // grow the device pool, keeping every slot (recomputed nodes
                // need not be the newest)
                const size_t n = (size_t)job.g->size() * sh.ncoef;
                job.task.out->deviceCoefs().reserve(n, std::min(n, job.task.out->deviceCoefs().capacity()));
            }
        }
        if (ahead.formed) {
            // the device's next level stands when the host formed the same
            // one (and it runs as one level batch); else its screen's table
            // slots are cleared before the next screen
            Job &job = jobs[0];
            ahead.ready = ahead_agrees(job, ahead) &&
                          (job.work.empty() || node_budget(job, job.g->scale(job.work[0])) * (double)job.work.size() <= kBatchBudget);
            ahead.cur = ahead.set;
            ahead.formed = false;
            if (!ahead.ready) S.dedupClean = false;
        } else {
            ahead.ready = false;
        }
    }
    std::vector<unsigned long long> na(2 * jobs.size());
    S.counters.download(na.data(), na.size());
    std::set<device::TreeMirror *> mirrors;
    for (auto &job : jobs) collect_samples(job);
    auto tt = clk::now();
    transform_and_norms(jobs, sh);
    const double tTransform = since(tt);
    std::vector<long long> nOut(jobs.size());
    host_parallel((int)jobs.size(), [&](int i) {
        detail::calc_square_norm(*jobs[i].g);
        nOut[i] = (long long)detail::real_nodes(*jobs[i].g).size();
    });
    for (size_t i = 0; i < jobs.size(); i++) {
        Job &job = jobs[i];
        job.task.out->markDeviceResident();
        job.stats.nApplied = (long long)na[2 * i];
        job.stats.nPairs = (long long)na[2 * i + 1];
        job.stats.tTransform = tTransform;
        job.stats.nOutputNodes = nOut[i];
    }
    const auto tc = clk::now();
    std::vector<int> owner; // first job of every distinct input mirror
    for (size_t j = 0; j < jobs.size(); j++)
        if (mirrors.insert(jobs[j].fm).second) owner.push_back((int)j);
    std::vector<char> dropped(owner.size(), 0);
    host_parallel((int)owner.size(), [&](int i) {
        const Job &job = jobs[owner[i]];
        dropped[i] = device::drop_generated_host(*job.fm, *job.mra, *job.fstore) ? 1 : 0;
    });
    for (size_t i = 0; i < owner.size(); i++)
        if (dropped[i]) device::drop_generated_device(*jobs[owner[i]].fm);
    const double tCleanup = since(tc);
    const double seconds = since(t_start);
    for (auto &job : jobs) {
        job.stats.seconds = seconds;
        job.stats.tSetup = tSetup;
        job.stats.tCleanup = tCleanup;
        g_batch_stats.push_back(job.stats);
    }
}

void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, int maxIter, bool absPrec) {
    if (out.getMRA() != inp.getMRA()) throw Error("Incompatible MRA");
    apply_batch({ApplyTask{prec, &out, &oper, &inp, maxIter, absPrec}});
    g_stats = g_batch_stats[0];
}

void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, const std::vector<FunctionTree<3> *> &precTrees,
           int maxIter, bool absPrec) {
    if (out.getMRA() != inp.getMRA()) throw Error("Incompatible MRA");
    ApplyTask t{prec, &out, &oper, &inp, maxIter, absPrec};
    t.precTrees = &precTrees;
    apply_batch({t});
    g_stats = g_batch_stats[0];
}

void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, FunctionTreeVector<3> &precTrees, int maxIter,
           bool absPrec) {
    std::vector<FunctionTree<3> *> trees;
    for (auto &t : precTrees) trees.push_back(std::get<1>(t));
    apply(prec, out, oper, inp, trees, maxIter, absPrec);
}

} // namespace mrcpp
