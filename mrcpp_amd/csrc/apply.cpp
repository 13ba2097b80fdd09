// apply(): MRCPP's breadth-first operator application with every level's band
// sum, generated band nodes and final multiwavelet transforms on the MI355X.
// The host keeps the tree topology and the split decisions
// (src/treebuilders/apply.cpp:68-93, src/treebuilders/TreeBuilder.cpp:38-87).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "function_tree.h"
#include "keys.h"

namespace mrcpp {

namespace {

thread_local ApplyStats g_stats;
thread_local SweepPrecision g_sweep = SweepPrecision::FP64;

using clk = std::chrono::steady_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

// Device scratch reused across levels and applies.
struct Scratch {
    device::DeviceArray<int> l, pairG, pairF, pairL, gFirst, gCount, count, genIds, genParent, genChild;
    device::DeviceArray<long long> gOff, pOff, cOff;
    device::DeviceArray<double> gNorms;
    device::DeviceArray<unsigned long long> nApplied;
};

Scratch &scratch() {
    static Scratch s;
    return s;
}

// Device bytes per batch of output nodes for the screen candidates and the
// per-pair group lists.
constexpr double kBatchBudget = 16.0e9;

struct LevelContext {
    const MultiResolutionAnalysis<3> &mra;
    ConvolutionOperator<3> &oper;
    const detail::BandTables &ht;
    const mwg_band_tables &bt;
    device::TreeMirror &fm;
    detail::NodeStore<3> *fstore;
    FunctionTree<3> &out;
    const double *filters;
    int kp1, n3, ncoef;
};

// Extend the input mirror with the generated nodes some candidate pairs need
// and patch those pairs to point at them.
void generate_band_nodes(LevelContext &c, std::vector<int> &pairF, const std::vector<int> &pairL, int scale) {
    device::TreeMirror &fm = c.fm;
    struct NewNode {
        int id, parent, child, scale;
    };
    std::vector<NewNode> made;
    for (size_t p = 0; p < pairF.size(); p++) {
        if (pairF[p] >= 0) continue;
        const int anc = -pairF[p] - 1;
        const std::array<int, 3> l = {pairL[3 * p], pairL[3 * p + 1], pairL[3 * p + 2]};
        int prev = anc;
        const int prevScale = (anc < fm.nReal) ? c.fstore->scale(anc) : fm.genScale[anc - fm.nReal];
        for (int s = prevScale + 1; s <= scale; s++) {
            std::array<int, 3> ls;
            for (int d = 0; d < 3; d++) ls[d] = l[d] >> (scale - s);
            const uint64_t key = device::key_of(c.mra, s, ls);
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
        pairF[p] = prev;
    }
    if (made.empty()) return;
    // device storage for the new nodes
    const int total = fm.nReal + fm.nGen;
    const int firstNew = total - (int)made.size();
    fm.hostOff.resize(total);
    for (const auto &m : made) fm.hostOff[m.id] = fm.genBase + (long long)(m.id - fm.nReal) * c.n3;
    fm.coefOff.reserve(total, firstNew);
    fm.coefOff.upload(fm.hostOff.data() + firstNew, made.size(), firstNew);
    fm.norms.reserve((size_t)total * 9, (size_t)firstNew * 9);
    fm.isGen.reserve(total, firstNew);
    std::vector<unsigned char> ones(made.size(), 1);
    fm.isGen.upload(ones.data(), ones.size(), firstNew);
    fm.coef.reserve((size_t)fm.genBase + (size_t)fm.nGen * c.n3, (size_t)fm.genBase + (size_t)(fm.nGen - made.size()) * c.n3);
    fm.uploadTable();
    // one launch per scale, coarse to fine
    std::stable_sort(made.begin(), made.end(), [](const NewNode &a, const NewNode &b) { return a.scale < b.scale; });
    Scratch &S = scratch();
    size_t i = 0;
    const mwg_ftree fv = fm.view(c.mra);
    while (i < made.size()) {
        size_t j = i;
        while (j < made.size() && made[j].scale == made[i].scale) j++;
        std::vector<int> ids, par, ch;
        for (size_t q = i; q < j; q++) {
            ids.push_back(made[q].id);
            par.push_back(made[q].parent);
            ch.push_back(made[q].child);
        }
        S.genIds.reserve(ids.size());
        S.genParent.reserve(ids.size());
        S.genChild.reserve(ids.size());
        S.genIds.upload(ids.data(), ids.size());
        S.genParent.upload(par.data(), par.size());
        S.genChild.upload(ch.data(), ch.size());
        device::check(mwg_generate_nodes(&fv, c.filters, S.genIds.get(), S.genParent.get(), S.genChild.get(), (int)ids.size()),
                      "generate band nodes");
        i = j;
    }
    g_stats.nGenerated += (long long)made.size();
}

// One batch of output nodes of one scale: screen, generate, band sum, reduce.
void run_batch(LevelContext &c, const std::vector<int> &nodes, int scale, double gthr) {
    Scratch &S = scratch();
    detail::NodeStore<3> &g = c.out.topology();
    const int L = (int)nodes.size();
    const int o_depth = scale - c.oper.getOperatorRoot();
    const bool empty_band = o_depth < 0 || o_depth >= c.ht.nDepth || c.ht.bandMax[o_depth] < 0;

    std::vector<int> lv(3 * (size_t)L);
    std::vector<long long> off(L);
    for (int i = 0; i < L; i++) {
        for (int d = 0; d < 3; d++) lv[3 * i + d] = g.transl(nodes[i])[d];
        off[i] = (long long)nodes[i] * c.ncoef;
    }
    S.l.reserve(3 * (size_t)L);
    S.l.upload(lv.data(), lv.size());
    S.gOff.reserve(L);
    S.gOff.upload(off.data(), L);
    S.gNorms.reserve(9 * (size_t)L);
    S.gFirst.reserve(L);
    S.gCount.reserve(L);
    S.count.reserve(2);

    mwg_level lvl{};
    lvl.kp1 = c.kp1;
    lvl.count = L;
    lvl.scale = scale;
    lvl.o_depth = o_depth;
    lvl.g_thr = gthr;
    lvl.band_w = empty_band ? 0 : c.ht.bandMax[o_depth];
    lvl.l = S.l.get();
    lvl.g_coef = c.out.deviceCoefs().get();
    lvl.g_off = S.gOff.get();
    lvl.g_norms = S.gNorms.get();

    mwg_ftree fv = c.fm.view(c.mra);
    mwg_pairs pr{};
    int npairs = 0, nmissing = 0;
    size_t cap = std::max<size_t>(S.pairF.capacity(), 4096);
    auto bind_pairs = [&]() {
        S.pairG.reserve(cap);
        S.pairF.reserve(cap);
        S.pairL.reserve(3 * cap);
        pr.capacity = (int)std::min<size_t>(S.pairF.capacity(), S.pairL.capacity() / 3);
        pr.count = S.count.get();
        pr.pair_g = S.pairG.get();
        pr.pair_f = S.pairF.get();
        pr.pair_l = S.pairL.get();
        pr.g_first = S.gFirst.get();
        pr.g_npairs = S.gCount.get();
    };
    bind_pairs();
    if (!empty_band) {
        auto t0 = clk::now();
        for (;;) {
            device::check(mwg_pair_screen(&lvl, &fv, &c.bt, &pr), "pair screen");
            int cnt[2];
            S.count.download(cnt, 2);
            npairs = cnt[0];
            nmissing = cnt[1];
            if (npairs < 0) throw Error("pair screen: band width does not match the operator tables");
            if (npairs <= pr.capacity) break;
            cap = (size_t)npairs + npairs / 4;
            bind_pairs();
        }
        g_stats.tScreen += since(t0);
        g_stats.nPairs += npairs;
    }
    if (npairs > 0 && nmissing > 0) {
        auto t0 = clk::now();
        std::vector<int> pf(npairs), pl(3 * (size_t)npairs);
        S.pairF.download(pf.data(), npairs);
        {
            S.pairL.download(pl.data(), pl.size());
            generate_band_nodes(c, pf, pl, scale);
            S.pairF.upload(pf.data(), npairs);
            fv = c.fm.view(c.mra);
        }
        g_stats.tGenerate += since(t0);
    }
    auto t0 = clk::now();
    // band sums, output coefficients and norms (zero when no pair survives)
    device::check(mwg_band_apply(&lvl, &fv, &c.bt, &pr, npairs, S.nApplied.get(), (int)g_sweep), "band apply");
    std::vector<double> nrm(9 * (size_t)L);
    S.gNorms.download(nrm.data(), nrm.size());
    g_stats.tBand += since(t0);
    for (int i = 0; i < L; i++) {
        double *cn = g.norms(nodes[i]);
        for (int t = 0; t < 8; t++) cn[t] = nrm[9 * (size_t)i + t];
        g.setSquareNorm(nodes[i], nrm[9 * (size_t)i + 8]);
        g.setHasCoefs(nodes[i], true);
    }
}

// One scale of the work vector, in batches that bound the device scratch.
void run_level(LevelContext &c, const std::vector<int> &nodes, int scale, double gthr) {
    // screen candidates take 8 bytes per band position; group lists 32 M
    // bytes per pair, with a generous per-node pair estimate (buffers grow
    // when a batch needs more)
    const int o_depth = scale - c.oper.getOperatorRoot();
    const int W = (o_depth >= 0 && o_depth < c.ht.nDepth) ? std::max(c.ht.bandMax[o_depth], 0) : 0;
    const double box = std::pow(2.0 * W + 1.0, 3.0);
    const double per_node = 8.0 * (box + 256.0) + 128.0 * (32.0 * c.ht.M + 24.0);
    size_t batch = std::max<size_t>(1, (size_t)(kBatchBudget / per_node));
    for (size_t i = 0; i < nodes.size(); i += batch) {
        std::vector<int> part(nodes.begin() + i, nodes.begin() + std::min(nodes.size(), i + batch));
        run_batch(c, part, scale, gthr);
    }
}

void transform_and_norms(LevelContext &c) {
    Scratch &S = scratch();
    detail::NodeStore<3> &g = c.out.topology();
    auto tab = detail::real_nodes_by_depth(g);
    auto launch = [&](const std::vector<int> &parents, bool up) {
        if (parents.empty()) return;
        std::vector<long long> po, co;
        for (int u : parents) {
            po.push_back((long long)u * c.ncoef);
            co.push_back((long long)g.child0(u) * c.ncoef);
        }
        S.pOff.reserve(po.size());
        S.cOff.reserve(co.size());
        S.pOff.upload(po.data(), po.size());
        S.cOff.upload(co.data(), co.size());
        double *coef = c.out.deviceCoefs().get();
        if (up)
            device::check(mwg_compress(c.kp1, c.filters, coef, S.pOff.get(), S.cOff.get(), (int)po.size(), c.ncoef), "compress");
        else
            device::check(mwg_reconstruct_add(c.kp1, c.filters, coef, S.pOff.get(), S.cOff.get(), (int)po.size(), c.ncoef),
                          "reconstruct");
    };
    auto branches = [&](const std::vector<int> &v) {
        std::vector<int> b;
        for (int u : v)
            if (g.hasRealChildren(u)) b.push_back(u);
        return b;
    };
    // TopDown: add coarse-scale contributions (MWTree::mwTransformDown, overwrite=false)
    for (size_t d = 0; d < tab.size(); d++) launch(branches(tab[d]), false);
    // BottomUp: recompress (MWTree::mwTransformUp)
    for (int d = (int)tab.size() - 2; d >= 0; d--) launch(branches(tab[d]), true);
    // component norms of every node
    std::vector<int> all = detail::real_nodes(g);
    std::vector<long long> off;
    for (int u : all) off.push_back((long long)u * c.ncoef);
    S.pOff.reserve(off.size());
    S.pOff.upload(off.data(), off.size());
    S.gNorms.reserve(9 * off.size());
    device::check(mwg_node_norms(c.kp1, c.out.deviceCoefs().get(), S.pOff.get(), S.gNorms.get(), (int)off.size()), "norms");
    std::vector<double> nrm(9 * off.size());
    S.gNorms.download(nrm.data(), nrm.size());
    for (size_t i = 0; i < all.size(); i++) {
        double *cn = g.norms(all[i]);
        for (int t = 0; t < 8; t++) cn[t] = nrm[9 * i + t];
        g.setSquareNorm(all[i], nrm[9 * i + 8]);
    }
}

// Drop the generated band nodes from the input mirror (FunctionTree::deleteGenerated).
void drop_generated(device::TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, detail::NodeStore<3> &f) {
    if (fm.nGen == 0) return;
    fm.table.reset(4 * fm.nReal + 1024);
    for (int u = 0; u < fm.nReal; u++) fm.table.insert(device::key_of(mra, f.scale(u), f.transl(u)), u);
    fm.nGen = 0;
    fm.genParent.clear();
    fm.genScale.clear();
    fm.genL.clear();
    fm.hostOff.resize(fm.nReal);
    fm.uploadTable();
}

} // namespace

const ApplyStats &last_apply_stats() { return g_stats; }

void set_sweep_precision(SweepPrecision p) { g_sweep = p; }
SweepPrecision sweep_precision() { return g_sweep; }

void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, int maxIter, bool absPrec) {
    if (out.getMRA() != inp.getMRA()) throw Error("Incompatible MRA");
    const auto t_start = clk::now();
    g_stats = ApplyStats{};
    device::ensure_init();
    const MultiResolutionAnalysis<3> &mra = out.getMRA();
    const int kp1 = mra.getKp1();
    if (kp1 != 6 && kp1 != 8 && kp1 != 10 && kp1 != 4) throw Error("apply supports Legendre orders 3, 5, 7 and 9");
    detail::DeviceBandTables *dbt = oper.deviceTables(prec);
    const detail::BandTables &ht = oper.bandTables(prec);
    device::TreeMirror &fm = inp.inputMirror();
    detail::NodeStore<3> &g = out.nodes();
    LevelContext c{mra, oper, ht, dbt->viewFor((int)g_sweep), fm, &inp.topology(), out, device::filters(mra.getOrder()), kp1, kp1 * kp1 * kp1, 8 * kp1 * kp1 * kp1};

    Scratch &S = scratch();
    S.nApplied.reserve(1);
    S.nApplied.zero(1);

    // device output storage for the nodes that exist now
    // output pool sized for an output tree as large as the input (grown on demand)
    out.deviceCoefs().reserve((size_t)std::max(g.size(), fm.nReal) * c.ncoef);

    const int M = oper.size();
    const int maxScale = mra.getMaxScale();
    std::vector<int> work = detail::real_nodes(g); // ConvolutionCalculator::getInitialWorkVector
    double sNorm = 0.0, wNorm = 0.0;
    for (int iter = 0; !work.empty(); iter++) {
        g_stats.levelNodes.push_back((int)work.size());
        double gthr = g.treeSquareNorm;
        if (gthr > 0.0) gthr = prec * 1.0 * std::sqrt(gthr / (double)M);
        out.deviceCoefs().reserve((size_t)g.size() * c.ncoef, (size_t)g.size() * c.ncoef);
        std::map<int, std::vector<int>> byScale;
        for (int u : work) byScale[g.scale(u)].push_back(u);
        for (auto &kv : byScale) run_level(c, kv.second, kv.first, gthr);
        auto th = clk::now();
        // TreeBuilder::build norm bookkeeping and WaveletAdaptor splits
        if (iter == 0) {
            sNorm = 0.0;
            for (int u : work)
                if (g.depth(u) >= 0) sNorm += g.scalingNorm(u);
        }
        for (int u : work)
            if (g.depth(u) >= 0) wNorm += g.waveletNorm(u);
        g.treeSquareNorm = (sNorm < 0.0 || wNorm < 0.0) ? -1.0 : sNorm + wNorm;
        if (iter >= maxIter && maxIter >= 0) work.clear();
        std::vector<int> next;
        for (int u : work) {
            if (g.hasChildren(u)) continue;
            if (g.scale(u) + 2 > maxScale) continue;
            if (detail::wavelet_split(prec, g.treeSquareNorm, absPrec, g.scale(u), g.waveletNorm(u))) {
                const int c0 = g.createChildren(u, false);
                for (int k = 0; k < 8; k++) next.push_back(c0 + k);
            }
        }
        work.swap(next);
        if (!work.empty()) {
            // grow the device pool, keeping the computed nodes
            const size_t keep = (size_t)(g.size() - (int)work.size()) * c.ncoef;
            out.deviceCoefs().reserve((size_t)g.size() * c.ncoef, keep);
        }
        g_stats.tHost += since(th);
    }
    auto tt = clk::now();
    transform_and_norms(c);
    detail::calc_square_norm(g);
    out.markDeviceResident();
    unsigned long long na = 0;
    S.nApplied.download(&na, 1);
    g_stats.nApplied = (long long)na;
    g_stats.tTransform = since(tt);
    drop_generated(fm, mra, inp.topology());
    g_stats.nOutputNodes = (long long)detail::real_nodes(g).size();
    g_stats.seconds = since(t_start);
}

} // namespace mrcpp
