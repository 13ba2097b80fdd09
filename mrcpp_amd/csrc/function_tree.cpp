// FunctionTree<3> host methods, input mirroring and the host-side builders.
#include "function_tree.h"

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
    detail::NodeStore<3> &s = nodes();
    double r = 0.0;
    for (int u : s.roots()) r += std::pow(2.0, -(3 * s.scale(u)) / 2.0) * s.coefs(u)[0];
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
    detail::NodeStore<3> &s = nodes();
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
    const size_t nc = m->ncoef;
    std::vector<double> stage((size_t)N * nc);
    std::vector<double> nrm((size_t)N * 9);
    m->hostOff.resize(N);
    for (int u = 0; u < N; u++) {
        std::memcpy(stage.data() + (size_t)u * nc, s.coefs(u), sizeof(double) * nc);
        for (int t = 0; t < 8; t++) nrm[(size_t)u * 9 + t] = s.compNorm(u, t);
        nrm[(size_t)u * 9 + 8] = s.squareNorm(u);
        m->hostOff[u] = (long long)u * nc;
    }
    m->coef.reserve(stage.size() + 1024 * (size_t)s.kp1_d());
    m->coef.upload(stage.data(), stage.size());
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

void build_grid(FunctionTree<3> &out, const GaussExp<3> &inp) {
    build_grid_nodes(out.nodes(), inp);
    out.invalidate();
}

void project(double prec, FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter, bool absPrec) {
    project_nodes(out.nodes(), prec, inp, maxIter, absPrec);
    out.invalidate();
}

// Root scaling products plus wavelet products of every node present in both
// trees, summed serially.
double dot(FunctionTree<3> &bra, FunctionTree<3> &ket) {
    if (bra.getMRA() != ket.getMRA()) throw Error("Trees not compatible");
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
