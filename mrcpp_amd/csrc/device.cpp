// Host-side device management: initialisation, band tables, filter banks,
// the key table mirrored on the device.
#include "device.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "keys.h"

namespace mrcpp {
namespace device {

void check(mwg_status s, const char *what) {
    if (s != MWG_OK) throw Error(std::string(what) + ": " + mwg_last_error());
}

namespace {
mwg_status init_status() {
    static std::once_flag once;
    static mwg_status st = MWG_OK;
    std::call_once(once, [] { st = mwg_init(0); });
    return st;
}
} // namespace

std::recursive_mutex &api_mutex() {
    static std::recursive_mutex *m = new std::recursive_mutex(); // never destroyed
    return *m;
}

void ensure_init() { check(init_status(), "device initialisation"); }

bool available() { return init_status() == MWG_OK; }

detail::DeviceBandTables *create_band_tables(const detail::BandTables &t) {
    ensure_init();
    auto *d = new detail::DeviceBandTables();
    auto put_i = [](DeviceArray<int> &a, const std::vector<int> &v) {
        a.reserve(std::max<size_t>(v.size(), 1));
        a.upload(v.data(), v.size());
    };
    auto put_d = [](DeviceArray<double> &a, const std::vector<double> &v) {
        a.reserve(std::max<size_t>(v.size(), 1));
        a.upload(v.data(), v.size());
    };
    put_i(d->bandMax, t.bandMax);
    put_i(d->widths, t.widths);
    put_i(d->bandSize, t.bandSize);
    d->matStart.reserve(std::max<size_t>(t.matStart.size(), 1));
    d->matStart.upload(t.matStart.data(), t.matStart.size());
    put_d(d->mats, t.mats);
    put_d(d->matNorm, t.matNorm);
    put_d(d->pairBound, t.pairBound);
    put_d(d->pairBoundW, t.pairBoundW);
    put_d(d->termNorm, t.termNorm);
    put_i(d->bandSizeMax, t.bandSizeMax);
    mwg_band_tables &v = d->view;
    v.M = t.M;
    v.n_depth = t.nDepth;
    v.kp1 = t.kp1;
    v.wmax = t.Wmax;
    v.band_max = d->bandMax.get();
    v.widths = d->widths.get();
    v.band_size = d->bandSize.get();
    v.mat_start = d->matStart.get();
    v.mats = d->mats.get();
    v.mat_norm = d->matNorm.get();
    v.pair_bound = d->pairBound.get();
    v.pair_bound_w = d->pairBoundW.get();
    v.band_size_max = d->bandSizeMax.get();
    v.term_norm = d->termNorm.get();
    v.n_mat = (long long)t.matNorm.size();
    // screening table: (2 band_max[d] + 1)^3 translation differences per depth
    std::vector<long long> off(t.nDepth + 1, 0);
    for (int dd = 0; dd < t.nDepth; dd++) {
        const long long nb = t.bandMax[dd] >= 0 ? 2ll * t.bandMax[dd] + 1 : 0;
        off[dd + 1] = off[dd] + nb * nb * nb;
    }
    d->screenOff.reserve(off.size());
    d->screenOff.upload(off.data(), off.size());
    d->screenBound.reserve(std::max<long long>(8 * off.back(), 1));
    v.screen_off = d->screenOff.get();
    check(mwg_build_screen_table(&v, off.back(), d->screenBound.get()), "screening table");
    v.screen_bound = d->screenBound.get();
    return d;
}

void destroy_band_tables(detail::DeviceBandTables *t) { delete t; }

const double *filters(int order) {
    static std::map<int, double *> memo;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto it = memo.find(order);
    if (it != memo.end()) return it->second;
    ensure_init();
    const auto &fb = core::filters(order);
    const int n2 = (order + 1) * (order + 1);
    std::vector<double> h(8 * (size_t)n2);
    for (int i = 0; i < 4; i++) {
        std::copy(fb.sub(i, Compression).a.begin(), fb.sub(i, Compression).a.end(), h.begin() + (size_t)i * n2);
        std::copy(fb.sub(i, Reconstruction).a.begin(), fb.sub(i, Reconstruction).a.end(), h.begin() + (size_t)(4 + i) * n2);
    }
    double *dp = nullptr;
    check(mwg_filters_upload(order + 1, h.data(), &dp), "filter upload");
    memo[order] = dp;
    return dp;
}

void KeyTable::reset(int min_entries) {
    log2cap_ = 10;
    while ((1 << log2cap_) < 2 * std::max(min_entries, 1)) log2cap_++;
    keys_.assign((size_t)2 << log2cap_, 0ull);
    vals_.assign((size_t)1 << log2cap_, -1);
    count_ = 0;
    full_ = true;
    dirty_.clear();
    sinceMark_.clear();
    markValid_ = false;
}

bool KeyTable::rollback() {
    if (!markValid_) return false;
    for (int slot : sinceMark_) {
        keys_[2 * (size_t)slot] = 0;
        keys_[2 * (size_t)slot + 1] = 0;
        vals_[slot] = -1;
        count_--;
        if (!full_) dirty_.push_back(slot);
    }
    sinceMark_.clear();
    return true;
}

int KeyTable::insert(const mwgkey::Key &key, int id) {
    if (2 * (count_ + 1) > capacity()) {
        // grow and rehash
        std::vector<unsigned long long> ok = keys_;
        std::vector<int> ov = vals_;
        log2cap_++;
        keys_.assign((size_t)2 << log2cap_, 0ull);
        vals_.assign((size_t)1 << log2cap_, -1);
        count_ = 0;
        full_ = true;
        dirty_.clear();
        markValid_ = false; // entries moved: a rollback would not find them
        sinceMark_.clear();
        for (size_t i = 0; i < ov.size(); i++)
            if (ok[2 * i]) insert(mwgkey::Key{ok[2 * i], ok[2 * i + 1]}, ov[i]);
    }
    const uint32_t mask = (uint32_t)capacity() - 1;
    uint32_t s = mwgkey::slot(key, log2cap_);
    while (true) {
        if (keys_[2 * (size_t)s] == 0) {
            keys_[2 * (size_t)s] = key.hi;
            keys_[2 * (size_t)s + 1] = key.lo;
            vals_[s] = id;
            count_++;
            if (!full_) dirty_.push_back((int)s);
            if (markValid_) sinceMark_.push_back((int)s);
            return id;
        }
        if (keys_[2 * (size_t)s] == key.hi && keys_[2 * (size_t)s + 1] == key.lo) return vals_[s];
        s = (s + 1) & mask;
    }
}

int KeyTable::find(const mwgkey::Key &key) const {
    const uint32_t mask = (uint32_t)capacity() - 1;
    uint32_t s = mwgkey::slot(key, log2cap_);
    while (true) {
        if (keys_[2 * (size_t)s] == 0) return -1;
        if (keys_[2 * (size_t)s] == key.hi && keys_[2 * (size_t)s + 1] == key.lo) return vals_[s];
        s = (s + 1) & mask;
    }
}

mwgkey::Key key_of(const MultiResolutionAnalysis<3> &mra, int scale, const std::array<int, 3> &l) {
    const int d = scale - mra.getRootScale();
    if (d < 0 || d > mwgkey::kMaxDepth) throw Error("node depth outside the device key range (max depth 30)");
    uint32_t o[3];
    for (int i = 0; i < 3; i++) {
        const long long v = (long long)l[i] - (long long)mra.getWorldBox().getCornerIndex()[i] * (1ll << d);
        if (v < 0 || v >= (1ll << 32)) throw Error("node translation outside the device key range");
        o[i] = (uint32_t)v;
    }
    return mwgkey::pack(d, o[0], o[1], o[2]);
}

// The whole table after a reset or rehash; otherwise only the slots filled
// since the last upload, scattered on the device.
void TreeMirror::uploadTable() {
    const size_t cap = table.capacity();
    if (table.fullUpload() || keys.capacity() < 2 * cap || vals.capacity() < cap) {
        keys.reserve(2 * cap);
        vals.reserve(cap);
        keys.upload(table.keys().data(), 2 * cap);
        vals.upload(table.vals().data(), cap);
    } else if (!table.dirty().empty()) {
        const std::vector<int> &d = table.dirty();
        const size_t n = d.size();
        std::vector<long long> upd(4 * n);
        for (size_t i = 0; i < n; i++) {
            upd[i] = d[i];
            upd[n + 2 * i] = (long long)table.keys()[2 * (size_t)d[i]];
            upd[n + 2 * i + 1] = (long long)table.keys()[2 * (size_t)d[i] + 1];
            upd[3 * n + i] = table.vals()[d[i]];
        }
        keyUpdates.reserve(upd.size());
        keyUpdates.upload(upd.data(), upd.size());
        check(mwg_hash_update(keys.get(), vals.get(), keyUpdates.get(), (int)n), "key table update");
    }
    table.markUploaded();
}

mwg_ftree TreeMirror::view(const MultiResolutionAnalysis<3> &mra) const {
    mwg_ftree f{};
    f.kp1 = kp1;
    f.root_scale = mra.getRootScale();
    for (int i = 0; i < 3; i++) {
        f.corner[i] = mra.getWorldBox().getCornerIndex()[i];
        f.nbox[i] = mra.getWorldBox().size(i);
    }
    f.hash_keys = keys.get();
    f.hash_vals = vals.get();
    f.hash_cap = table.capacity();
    f.coef = coef.get();
    f.coef_off = coefOff.get();
    f.norms = norms.get();
    f.is_gen = isGen.get();
    f.norm_max = normMax;
    return f;
}

int plan_mirror_node(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, int n, const std::array<int, 3> &l, std::vector<GenNode> &made) {
    int id = fm.table.find(key_of(mra, n, l));
    if (id >= 0) return id;
    // nearest present ancestor
    int s = n - 1, anc = -1;
    for (; s >= mra.getRootScale(); s--) {
        std::array<int, 3> ls;
        for (int d = 0; d < 3; d++) ls[d] = l[d] >> (n - s);
        anc = fm.table.find(key_of(mra, s, ls));
        if (anc >= 0) break;
    }
    if (anc < 0) throw Error("node outside the world box");
    int prev = anc;
    for (int t = s + 1; t <= n; t++) {
        std::array<int, 3> lt;
        for (int d = 0; d < 3; d++) lt[d] = l[d] >> (n - t);
        id = fm.nReal + fm.nGen;
        fm.table.insert(key_of(mra, t, lt), id);
        fm.nGen++;
        fm.genParent.push_back(prev);
        fm.genScale.push_back(t);
        fm.genL.push_back(lt);
        int child = 0;
        for (int d = 0; d < 3; d++) child |= (lt[d] & 1) << d;
        made.push_back({id, prev, child, t});
        prev = id;
    }
    return prev;
}

void MirrorBatch::add(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const double *filters, std::vector<GenNode> &made) {
    if (made.empty()) return;
    const int n3 = fm.kp1 * fm.kp1 * fm.kp1;
    const int total = fm.nReal + fm.nGen;
    const int firstNew = total - (int)made.size();
    fm.hostOff.resize(total);
    for (const auto &m : made) fm.hostOff[m.id] = fm.genBase + (long long)(m.id - fm.nReal) * n3;
    fm.coefOff.reserve(total, firstNew);
    fm.norms.reserve((size_t)total * 9, (size_t)firstNew * 9);
    fm.isGen.reserve(total, firstNew);
    fm.coef.reserve((size_t)fm.genBase + (size_t)fm.nGen * n3, (size_t)fm.genBase + (size_t)(fm.nGen - made.size()) * n3);
    Entry e{&fm, &mra, filters, 0, 0, 0, 0, made.size(), {}};
    // key table: a full upload now (grown or first), else its changed slots
    // go with the batch (re-applying them after a later full upload of the
    // same table writes the same values: slots are only ever filled)
    const size_t cap = fm.table.capacity();
    if (fm.table.fullUpload() || fm.keys.capacity() < 2 * cap || fm.vals.capacity() < cap) {
        fm.uploadTable();
    } else {
        const std::vector<int> &d = fm.table.dirty();
        e.upd = words_.size();
        e.nupd = d.size();
        words_.resize(e.upd + 4 * d.size());
        long long *u = words_.data() + e.upd;
        const size_t nd = d.size();
        for (size_t k = 0; k < nd; k++) {
            u[k] = d[k];
            u[nd + 2 * k] = (long long)fm.table.keys()[2 * (size_t)d[k]];
            u[nd + 2 * k + 1] = (long long)fm.table.keys()[2 * (size_t)d[k] + 1];
            u[3 * nd + k] = fm.table.vals()[d[k]];
        }
        fm.table.markUploaded();
    }
    std::stable_sort(made.begin(), made.end(), [](const GenNode &a, const GenNode &b) { return a.scale < b.scale; });
    e.offs = words_.size();
    for (const auto &m : made) words_.push_back(fm.hostOff[m.id]);
    e.ids = ints_.size();
    ints_.resize(e.ids + 3 * made.size());
    int *g = ints_.data() + e.ids;
    const size_t n = made.size();
    for (size_t k = 0; k < n; k++) {
        g[k] = made[k].id;
        g[n + k] = made[k].parent;
        g[2 * n + k] = made[k].child;
    }
    for (size_t i = 0; i < n;) {
        size_t j = i;
        while (j < n && made[j].scale == made[i].scale) j++;
        e.runs.push_back({i, j - i});
        e.runScale.push_back(made[i].scale);
        i = j;
    }
    entries_.push_back(std::move(e));
}

size_t MirrorBatch::addInts(const int *p, size_t n) {
    const size_t at = ints_.size();
    ints_.insert(ints_.end(), p, p + n);
    return at;
}

const int *MirrorBatch::run(DeviceArray<long long> &staging) {
    // trees of one order and filter bank: one key-table launch for all of
    // them and one generation launch per scale (coarse to fine)
    bool one = entries_.size() > 1;
    for (const Entry &e : entries_) one = one && e.filters == entries_[0].filters && e.fm->kp1 == entries_[0].fm->kp1;
    if (one) return runTogether(staging);
    const size_t nw = words_.size(), total = nw + (ints_.size() + 1) / 2;
    if (total == 0) return nullptr;
    std::vector<long long> buf(total, 0);
    std::copy(words_.begin(), words_.end(), buf.begin());
    if (!ints_.empty()) std::memcpy(buf.data() + nw, ints_.data(), sizeof(int) * ints_.size());
    staging.reserve(total);
    staging.upload(buf.data(), total);
    const long long *W = staging.get();
    const int *I = reinterpret_cast<const int *>(W + nw);
    for (const Entry &e : entries_) {
        TreeMirror &fm = *e.fm;
        if (e.nupd) check(mwg_hash_update(fm.keys.get(), fm.vals.get(), W + e.upd, (int)e.nupd), "key table update");
        const mwg_ftree fv = fm.view(*e.mra); // after every reserve of the batch
        const int *ids = I + e.ids, *par = ids + e.n, *ch = ids + 2 * e.n;
        for (const auto &r : e.runs)
            check(mwg_generate_nodes_recorded(&fv, e.filters, ids + r.first, par + r.first, ch + r.first, W + e.offs + r.first, fm.coefOff.get(),
                                              fm.isGen.get(), (int)r.second),
                  "generate nodes");
    }
    entries_.clear();
    words_.clear();
    ints_.clear();
    return I;
}

const int *MirrorBatch::runTogether(DeviceArray<long long> &staging) {
    const size_t nEnt = entries_.size();
    std::vector<mwg_gen_tree> trees(nEnt);
    std::vector<mwg_hash_table> tables(nEnt);
    std::vector<long long> upd; // (tree, slot, key hi, key lo, value)
    struct G {
        int scale, tree;
        size_t k; // node k of the tree's entry
    };
    std::vector<G> gen;
    for (size_t t = 0; t < nEnt; t++) {
        const Entry &e = entries_[t];
        TreeMirror &fm = *e.fm;
        trees[t] = mwg_gen_tree{fm.view(*e.mra), fm.coefOff.get(), fm.isGen.get()}; // after every reserve of the batch
        tables[t] = mwg_hash_table{fm.keys.get(), fm.vals.get()};
        const long long *u = words_.data() + e.upd;
        for (size_t k = 0; k < e.nupd; k++)
            upd.insert(upd.end(), {(long long)t, u[k], u[e.nupd + 2 * k], u[e.nupd + 2 * k + 1], u[3 * e.nupd + k]});
        for (size_t r = 0; r < e.runs.size(); r++)
            for (size_t k = e.runs[r].first; k < e.runs[r].first + e.runs[r].second; k++) gen.push_back({e.runScale[r], (int)t, k});
    }
    std::stable_sort(gen.begin(), gen.end(), [](const G &a, const G &b) { return a.scale < b.scale; });
    // one upload: [updates | node offsets | tree views | tables | integers of addInts | tree, id, parent, child of every node]
    static_assert(sizeof(mwg_gen_tree) % 8 == 0 && sizeof(mwg_hash_table) % 8 == 0, "upload layout");
    const size_t ng = gen.size(), wTree = sizeof(mwg_gen_tree) / 8, wTab = sizeof(mwg_hash_table) / 8;
    const size_t oOffs = upd.size(), oTrees = oOffs + ng, oTabs = oTrees + wTree * nEnt, oInts = oTabs + wTab * nEnt;
    const size_t nInts = ints_.size() + 4 * ng;
    std::vector<long long> buf(oInts + (nInts + 1) / 2, 0);
    std::copy(upd.begin(), upd.end(), buf.begin());
    std::memcpy(buf.data() + oTrees, trees.data(), sizeof(mwg_gen_tree) * nEnt);
    std::memcpy(buf.data() + oTabs, tables.data(), sizeof(mwg_hash_table) * nEnt);
    int *I = reinterpret_cast<int *>(buf.data() + oInts);
    if (!ints_.empty()) std::memcpy(I, ints_.data(), sizeof(int) * ints_.size());
    int *gt = I + ints_.size(), *gid = gt + ng, *gpar = gid + ng, *gch = gpar + ng;
    for (size_t i = 0; i < ng; i++) {
        const Entry &e = entries_[gen[i].tree];
        const size_t k = gen[i].k;
        buf[oOffs + i] = words_[e.offs + k];
        gt[i] = gen[i].tree;
        gid[i] = ints_[e.ids + k];
        gpar[i] = ints_[e.ids + e.n + k];
        gch[i] = ints_[e.ids + 2 * e.n + k];
    }
    staging.reserve(buf.size());
    staging.upload(buf.data(), buf.size());
    const long long *W = staging.get();
    const int *ID = reinterpret_cast<const int *>(W + oInts);
    const auto *treesDev = reinterpret_cast<const mwg_gen_tree *>(W + oTrees);
    if (!upd.empty())
        check(mwg_hash_update_multi(reinterpret_cast<const mwg_hash_table *>(W + oTabs), (int)nEnt, W, (int)(upd.size() / 5)), "key table update");
    const int *dt = ID + ints_.size(), *did = dt + ng, *dpar = did + ng, *dch = dpar + ng;
    for (size_t a = 0; a < ng;) {
        size_t b = a;
        while (b < ng && gen[b].scale == gen[a].scale) b++;
        check(mwg_generate_nodes_multi(entries_[0].fm->kp1, treesDev, (int)nEnt, entries_[0].filters, dt + a, did + a, dpar + a, dch + a,
                                       W + oOffs + a, (int)(b - a)),
              "generate nodes");
        a = b;
    }
    entries_.clear();
    words_.clear();
    ints_.clear();
    return ID;
}

void generate_mirror_nodes(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const double *filters, std::vector<GenNode> &made,
                           DeviceArray<long long> &staging) {
    MirrorBatch b;
    b.add(fm, mra, filters, made);
    b.run(staging);
}

bool drop_generated_host(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const detail::NodeStore<3> &f) {
    if (fm.nGen == 0) return false;
    // the generated nodes' slots emptied (only those go to the device), or
    // after a rehash the real nodes' table rebuilt
    if (!fm.table.rollback()) {
        fm.table.reset(4 * fm.nReal + 1024);
        for (int u = 0; u < fm.nReal; u++) fm.table.insert(key_of(mra, f.scale(u), f.transl(u)), u);
        fm.table.mark();
    }
    fm.nGen = 0;
    fm.genParent.clear();
    fm.genScale.clear();
    fm.genL.clear();
    fm.hostOff.resize(fm.nReal);
    return true;
}

void drop_generated_device(TreeMirror &fm) { fm.uploadTable(); }

void drop_generated(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const detail::NodeStore<3> &f) {
    if (drop_generated_host(fm, mra, f)) drop_generated_device(fm);
}

} // namespace device

namespace detail {

const mwg_band_tables &DeviceBandTables::viewFor(int precision) {
    if (precision != MWG_PREC_FP64 && precision != MWG_PREC_FP32 && precision != MWG_PREC_BF16X3) return view;
    if (!view.frag[precision] && view.n_mat > 0) {
        const size_t bytes = mwg_fragment_bytes(view.kp1, view.n_mat, precision);
        if (bytes == 0) throw Error("no matrix-core operand table for this sweep precision");
        frag[precision].reserve(bytes);
        device::check(mwg_build_fragments(&view, precision, frag[precision].get()), "operand table");
        view.frag[precision] = frag[precision].get();
    }
    return view;
}

} // namespace detail
} // namespace mrcpp
