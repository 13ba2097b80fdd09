// FunctionTree<3> host methods, input mirroring and the host-side builders.
#include "function_tree.h"

#include <map>
#include <memory>
#include <mutex>

#include <cmath>
#include <cstring>

namespace mrcpp {

FunctionTree<3>::FunctionTree(const MultiResolutionAnalysis<3> &mra, const std::string &name)
        : mra_(mra)
        , name_(name)
        , store_(std::make_unique<detail::NodeStore<3>>(mra_)) {}

void FunctionTree<3>::clear() {
    store_ = std::make_unique<detail::NodeStore<3>>(mra_);
    host_stale_ = false;
    mirror_.reset();
    dev_coef_.release();
}

int FunctionTree<3>::getNNodes() const { return (int)detail::real_nodes(*store_).size(); }

int FunctionTree<3>::getNEndNodes() const { return (int)detail::end_nodes(*store_).size(); }

void FunctionTree<3>::syncHost() {
    if (!host_stale_) return;
    detail::NodeStore<3> &s = *store_;
    const size_t nc = s.ncoef();
    std::vector<double> buf((size_t)s.size() * nc);
    dev_coef_.download(buf.data(), buf.size());
    for (int u = 0; u < s.size(); u++) std::memcpy(s.coefs(u), buf.data() + (size_t)u * nc, sizeof(double) * nc);
    host_stale_ = false;
}

detail::NodeStore<3> &FunctionTree<3>::nodes() {
    syncHost();
    return *store_;
}

// Legendre integral: root scaling coefficient 0 times 2^{-3n/2}, times the
// square roots of the scaling factors (src/trees/FunctionTree.cpp:438-454).
double FunctionTree<3>::integrate() {
    detail::NodeStore<3> &s = *store_;
    double r = 0.0;
    for (int u : s.roots()) {
        double c0 = s.coefs(u)[0];
        // coefficients in HBM: only the roots' first scaling coefficient comes back
        if (host_stale_) dev_coef_.download(&c0, 1, (size_t)u * s.ncoef());
        r += std::pow(2.0, -(3 * s.scale(u)) / 2.0) * c0;
    }
    double jac = 1.0;
    for (int d = 0; d < 3; d++) jac *= std::sqrt(mra_.getWorldBox().getScalingFactor(d));
    return jac * r;
}

// Scaling part of the end node containing r (src/trees/FunctionTree.cpp:524-546,
// src/trees/FunctionNode.cpp:60-88).
double FunctionTree<3>::evalf(const Coord<3> &r) {
    detail::NodeStore<3> &s = nodes();
    const auto &box = mra_.getWorldBox();
    Coord<3> x;
    for (int d = 0; d < 3; d++) {
        x[d] = r[d] / box.getScalingFactor(d);
        const double lo = box.getCornerIndex()[d] * std::pow(2.0, -box.getScale());
        const double hi = lo + box.size(d) * std::pow(2.0, -box.getScale());
        if (x[d] < lo || x[d] > hi) return 0.0;
    }
    std::array<int, 3> l;
    for (int d = 0; d < 3; d++) {
        int li = (int)std::floor(x[d] * std::pow(2.0, box.getScale()));
        l[d] = std::min(li, box.getCornerIndex()[d] + box.size(d) - 1);
    }
    int u = s.roots()[box.rootIndex(box.getScale(), l)];
    while (s.hasRealChildren(u)) {
        int c = 0;
        for (int d = 0; d < 3; d++) {
            int li = (int)std::floor(x[d] * std::pow(2.0, s.scale(u) + 1));
            li = std::min(std::max(li, 2 * s.transl(u)[d]), 2 * s.transl(u)[d] + 1);
            c |= (li - 2 * s.transl(u)[d]) << d;
        }
        u = s.child0(u) + c;
    }
    const auto &bas = mra_.basis();
    const int kp1 = s.kp1();
    const double nf = std::pow(2.0, s.scale(u));
    std::vector<double> v(3 * kp1);
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < kp1; k++) v[d * kp1 + k] = bas.phi(k, x[d] * nf - (double)s.transl(u)[d]);
    double res = 0.0;
    const double *c = s.coefs(u);
    for (int i = 0; i < s.kp1_d(); i++) res += c[i] * v[i % kp1] * v[kp1 + (i / kp1) % kp1] * v[2 * kp1 + i / (kp1 * kp1)];
    double coef = 1.0;
    for (int d = 0; d < 3; d++) coef /= std::sqrt(box.getScalingFactor(d));
    return coef * std::pow(2.0, (3 * s.scale(u)) / 2.0) * res;
}

device::TreeMirror &FunctionTree<3>::inputMirror() {
    if (mirror_ && mirror_->valid) return *mirror_;
    device::ensure_init();
    // coefficients already in this tree's device pool (slot = node id, the
    // mirror's layout) are copied on the device
    const bool onDevice = host_stale_;
    detail::NodeStore<3> &s = onDevice ? *store_ : nodes();
    auto m = std::make_unique<device::TreeMirror>();
    m->kp1 = s.kp1();
    m->ncoef = s.ncoef();
    const int N = s.size();
    for (int u = 0; u < N; u++)
        if (s.isGen(u)) throw Error("input tree holds generated nodes");
    m->nReal = N;
    m->nGen = 0;
    m->genBase = (long long)N * m->ncoef;
    m->table.reset(4 * N + 1024);
    for (int u = 0; u < N; u++) m->table.insert(device::key_of(mra_, s.scale(u), s.transl(u)), u);
    m->table.mark(); // generated nodes inserted later are removed by rollback
    const size_t nc = m->ncoef;
    std::vector<double> nrm((size_t)N * 9);
    m->hostOff.resize(N);
    double normMax = 0.0;
    for (int u = 0; u < N; u++) {
        for (int t = 0; t < 8; t++) {
            nrm[(size_t)u * 9 + t] = s.compNorm(u, t);
            normMax = std::max(normMax, nrm[(size_t)u * 9 + t]);
        }
        nrm[(size_t)u * 9 + 8] = s.squareNorm(u);
        normMax = std::max(normMax, std::sqrt(std::max(nrm[(size_t)u * 9 + 8], 0.0)));
        m->hostOff[u] = (long long)u * nc;
    }
    m->normMax = std::isfinite(normMax) ? normMax * (1.0 + 1e-10) : HUGE_VAL;
    m->coef.reserve((size_t)N * nc + 1024 * (size_t)s.kp1_d());
    if (onDevice) {
        device::check(mwg_copy(m->coef.get(), dev_coef_.get(), sizeof(double) * (size_t)N * nc), "mirror copy");
    } else {
        std::vector<double> stage((size_t)N * nc);
        for (int u = 0; u < N; u++) std::memcpy(stage.data() + (size_t)u * nc, s.coefs(u), sizeof(double) * nc);
        m->coef.upload(stage.data(), stage.size());
    }
    m->norms.reserve(nrm.size() + 9 * 1024);
    m->norms.upload(nrm.data(), nrm.size());
    m->coefOff.reserve(N + 1024);
    m->coefOff.upload(m->hostOff.data(), N);
    std::vector<unsigned char> g(N, 0);
    m->isGen.reserve(N + 1024);
    m->isGen.upload(g.data(), N);
    m->uploadTable();
    m->valid = true;
    mirror_ = std::move(m);
    return *mirror_;
}

namespace device {

// kept until the process ends (never destroyed: no device call runs from a
// static destructor)
const BasisTables &basis_tables(const MultiResolutionAnalysis<3> &mra) {
    static std::map<int, BasisTables *> memo;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto &slot = memo[mra.getOrder()];
    if (!slot) {
        slot = new BasisTables();
        const auto &b = mra.basis();
        slot->nodes.reserve(b.nodes().size());
        slot->nodes.upload(b.nodes().data(), b.nodes().size());
        slot->vcmap.reserve(b.vcMap().a.size());
        slot->vcmap.upload(b.vcMap().a.data(), b.vcMap().a.size());
        slot->cvmap.reserve(b.cvMap().a.size());
        slot->cvmap.upload(b.cvMap().a.data(), b.cvMap().a.size());
    }
    return *slot;
}

} // namespace device

void build_grid(FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter) {
    build_grid_nodes(out.nodes(), inp, maxIter);
    out.invalidate();
}

void build_grid(FunctionTree<3> &out, const GaussFunc<3> &inp, int maxIter) {
    GaussExp<3> g;
    g.append(inp);
    build_grid(out, g, maxIter);
}

namespace {

thread_local ProjectionTarget g_proj_target = ProjectionTarget::Auto;
thread_local bool g_proj_device = false;


// project() with every level's nodes computed on the device (ProjectionCalculator
// per level, then MWTree::mwTransform(BottomUp) and calcSquareNorm,
// src/treebuilders/project.cpp:85-102); coefficients stay in the tree's device pool.
void project_device(double prec, FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter, bool absPrec) {
    device::ensure_init();
    detail::NodeStore<3> &s = out.topology();
    const auto &mra = out.getMRA();
    const int nc = s.ncoef();
    const device::BasisTables &bt = device::basis_tables(mra);
    const double *filt = device::filters(mra.getOrder());
    std::vector<mwg_gauss> gs(inp.size());
    for (int i = 0; i < inp.size(); i++) {
        const auto &g = inp.getFunc(i);
        gs[i].coef = g.getCoef();
        for (int d = 0; d < 3; d++) {
            gs[i].alpha[d] = g.getExp()[d];
            gs[i].pos[d] = g.getPos()[d];
        }
    }
    device::DeviceArray<mwg_gauss> dg;
    dg.reserve(std::max<size_t>(gs.size(), 1));
    dg.upload(gs.data(), gs.size());
    mwg_projection pj{};
    pj.kp1 = s.kp1();
    for (int d = 0; d < 3; d++) pj.sf[d] = mra.getWorldBox().getScalingFactor(d);
    pj.nodes = bt.nodes.get();
    pj.vcmap = bt.vcmap.get();
    pj.filters = filt;
    pj.gauss = dg.get();
    pj.ngauss = (int)gs.size();
    device::DeviceArray<double> &pool = out.deviceCoefs();
    device::DeviceArray<long long> lvl; // per level: offsets [n] | scale and translation ints [4n]
    device::DeviceArray<double> nrm;
    pool.reserve((size_t)s.size() * nc);
    auto set_norms = [&](const std::vector<int> &v, const std::vector<double> &h) {
        for (size_t i = 0; i < v.size(); i++) {
            double *cn = s.norms(v[i]);
            for (int t = 0; t < 8; t++) cn[t] = h[9 * i + t];
            s.setSquareNorm(v[i], h[9 * i + 8]);
            s.setHasCoefs(v[i], true);
        }
    };
    detail::BuildHooks<3> hk;
    hk.initial = [](detail::NodeStore<3> &st) { return detail::end_nodes(st); };
    hk.compute = [&](detail::NodeStore<3> &st, const std::vector<int> &v) {
        const size_t n = v.size();
        if (!n) return;
        pool.reserve((size_t)st.size() * nc, pool.capacity());
        std::vector<long long> packed(n + (4 * n + 1) / 2, 0);
        int *ints = reinterpret_cast<int *>(packed.data() + n);
        for (size_t i = 0; i < n; i++) {
            packed[i] = (long long)v[i] * nc;
            ints[i] = st.scale(v[i]);
            for (int d = 0; d < 3; d++) ints[n + 3 * i + d] = st.transl(v[i])[d];
        }
        lvl.reserve(packed.size());
        lvl.upload(packed.data(), packed.size());
        nrm.reserve(9 * n);
        const int *dints = reinterpret_cast<const int *>(lvl.get() + n);
        device::check(mwg_project_nodes(&pj, dints, dints + n, lvl.get(), pool.get(), nrm.get(), (int)n), "project nodes");
        std::vector<double> h(9 * n);
        nrm.download(h.data(), h.size());
        set_norms(v, h);
    };
    hk.split = [prec, absPrec](const detail::NodeStore<3> &st, int u) {
        return detail::wavelet_split(prec, st.treeSquareNorm, absPrec, st.scale(u), st.waveletNorm(u));
    };
    detail::build_tree(s, hk, maxIter);
    // BottomUp (MWTree::mwTransformUp): parents from the deepest level up, then their norms
    const auto tab = detail::real_nodes_by_depth(s);
    std::vector<int> parents;
    device::DeviceArray<long long> offs;
    for (int d = (int)tab.size() - 2; d >= 0; d--) {
        std::vector<int> v;
        for (int u : tab[d])
            if (s.hasRealChildren(u)) v.push_back(u);
        if (v.empty()) continue;
        std::vector<long long> o(2 * v.size());
        for (size_t i = 0; i < v.size(); i++) {
            o[i] = (long long)v[i] * nc;
            o[v.size() + i] = (long long)s.child0(v[i]) * nc;
        }
        offs.reserve(o.size());
        offs.upload(o.data(), o.size());
        device::check(mwg_compress(s.kp1(), filt, pool.get(), offs.get(), offs.get() + v.size(), (int)v.size(), nc), "compress");
        parents.insert(parents.end(), v.begin(), v.end());
    }
    if (!parents.empty()) {
        std::vector<long long> o(parents.size());
        for (size_t i = 0; i < parents.size(); i++) o[i] = (long long)parents[i] * nc;
        offs.reserve(o.size());
        offs.upload(o.data(), o.size());
        nrm.reserve(9 * parents.size());
        device::check(mwg_node_norms(s.kp1(), pool.get(), offs.get(), nrm.get(), (int)parents.size()), "node norms");
        std::vector<double> h(9 * parents.size());
        nrm.download(h.data(), h.size());
        set_norms(parents, h);
    }
    detail::calc_square_norm(s);
    out.markDeviceResident();
}

} // namespace

void set_projection_target(ProjectionTarget t) { g_proj_target = t; }
bool last_projection_on_device() { return g_proj_device; }

void project(double prec, FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter, bool absPrec) {
    const bool dev = g_proj_target == ProjectionTarget::Device || (g_proj_target == ProjectionTarget::Auto && device::available());
    g_proj_device = dev;
    if (dev) {
        project_device(prec, out, inp, maxIter, absPrec);
        return;
    }
    project_nodes(out.nodes(), prec, inp, maxIter, absPrec);
    out.invalidate();
}

void project(double prec, FunctionTree<3> &out, const GaussFunc<3> &inp, int maxIter, bool absPrec) {
    GaussExp<3> g;
    g.append(inp);
    project(prec, out, g, maxIter, absPrec);
}

namespace {
struct PointFunction {
    const std::function<double(const Coord<3> &)> &f;
    double evalf(const Coord<3> &r) const { return f(r); }
};
} // namespace

void project(double prec, FunctionTree<3> &out, std::function<double(const Coord<3> &r)> func, int maxIter, bool absPrec) {
    if (!func) throw Error("project: empty function");
    g_proj_device = false;
    // node values come from arbitrary host code: one node per call, in order
    // within a node, nodes of a level in parallel as MRCPP does
    project_nodes(out.nodes(), prec, PointFunction{func}, maxIter, absPrec);
    out.invalidate();
}

// Root scaling products plus wavelet products of every node present in both
// trees, summed serially.
double dot(FunctionTree<3> &bra, FunctionTree<3> &ket) {
    if (bra.getMRA() != ket.getMRA()) throw Error("Trees not compatible");
    // coefficients in HBM: the node products there, no copy to the host
    if (bra.hostStale() || ket.hostStale()) return detail::dot_device(bra, ket);
    detail::NodeStore<3> &a = bra.nodes();
    detail::NodeStore<3> &b = ket.nodes();
    const int kd = a.kp1_d();
    double result = 0.0;
    for (int u : detail::real_nodes(a)) {
        const int v = b.findReal(a.scale(u), a.transl(u));
        if (v < 0) continue;
        const double *x = a.coefs(u), *y = b.coefs(v);
        if (a.depth(u) == 0) {
            double s = 0.0;
            for (int i = 0; i < kd; i++) s += x[i] * y[i];
            result += s;
        }
        double w = 0.0;
        for (int i = kd; i < 8 * kd; i++) w += x[i] * y[i];
        result += w;
    }
    return result;
}

} // namespace mrcpp
