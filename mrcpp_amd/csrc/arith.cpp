// Sums and products of trees on the MI355X: add (src/treebuilders/add.cpp:43-80)
// and multiply (src/treebuilders/multiply.cpp:52-100).  The host keeps the
// TreeBuilder loop (TreeBuilder::build, src/treebuilders/TreeBuilder.cpp:38-87):
// the output's end nodes, then after each step the approximate tree norm and
// the WaveletAdaptor split test.  Every step's nodes are computed in one launch
// (mwg_combine_nodes) from the operands' device mirrors, with operand nodes
// missing below their trees' leaves generated on the device first
// (MWTree::getNode); the bottom-up transform and the norms run on the device
// and the result stays in HBM.
#include <set>

#include "function_tree.h"
#include "max_norms.h"

namespace mrcpp {

namespace {

enum Mode { Sum = 0, Product = 1 };

// Device buffers of the steps, kept for the process (never destroyed: no
// device call from a static destructor).
struct ArithScratch {
    device::DeviceArray<long long> offs;
    device::DeviceArray<int> scale;
    device::DeviceArray<long long> staging;
    device::DeviceArray<double> norms;
};
ArithScratch &arith_scratch() {
    static ArithScratch *s = new ArithScratch();
    return *s;
}

// MWTree::mwTransform(BottomUp) and MWTree::calcSquareNorm on the device copy
// of the output (src/trees/MWTree.cpp:109-118,166-181).
void compress_and_norms(FunctionTree<3> &out, const double *filters) {
    ArithScratch &S = arith_scratch();
    detail::NodeStore<3> &g = out.topology();
    const int kp1 = g.kp1(), ncoef = g.ncoef();
    const auto tab = detail::real_nodes_by_depth(g);
    std::vector<long long> po, co;
    std::vector<std::pair<size_t, size_t>> passes; // (first, count) per depth, deepest parents first
    for (int d = (int)tab.size() - 2; d >= 0; d--) {
        const size_t at = po.size();
        for (int u : tab[d])
            if (g.hasRealChildren(u)) {
                po.push_back((long long)u * ncoef);
                co.push_back((long long)g.child0(u) * ncoef);
            }
        if (po.size() > at) passes.push_back({at, po.size() - at});
    }
    const std::vector<int> all = detail::real_nodes(g);
    const size_t normAt = po.size(), coAt = po.size() + all.size();
    for (int u : all) po.push_back((long long)u * ncoef);
    po.insert(po.end(), co.begin(), co.end());
    S.offs.reserve(po.size());
    S.offs.upload(po.data(), po.size());
    double *coef = out.deviceCoefs().get();
    for (const auto &p : passes)
        device::check(mwg_compress(kp1, filters, coef, S.offs.get() + p.first, S.offs.get() + coAt + p.first, (int)p.second, ncoef), "compress");
    S.norms.reserve(9 * all.size());
    device::check(mwg_node_norms(kp1, coef, S.offs.get() + normAt, S.norms.get(), (int)all.size()), "norms");
    std::vector<double> nrm(9 * all.size());
    S.norms.download(nrm.data(), nrm.size());
    for (size_t i = 0; i < all.size(); i++) {
        double *cn = g.norms(all[i]);
        for (int t = 0; t < 8; t++) cn[t] = nrm[9 * i + t];
        g.setSquareNorm(all[i], nrm[9 * i + 8]);
        g.setHasCoefs(all[i], true);
    }
    detail::calc_square_norm(g);
}

// MultiplicationAdaptor::splitNode (src/treebuilders/MultiplicationAdaptor.h:46-66):
// split where maxW0 maxS1 + maxW1 maxS0 + maxW0 maxW1 > prec (square roots of
// the operands' largest scaled norms at the node's index) unless both
// operands are leaves there.
struct MaxNormSplit {
    detail::MaxNorms a, b;
    bool operator()(double prec, int n, const std::array<int, 3> &l) {
        const detail::MaxNorms::Info p0 = a.at(n, l), p1 = b.at(n, l);
        const double maxW0 = std::sqrt(p0.maxWSq), maxW1 = std::sqrt(p1.maxWSq);
        const double maxS0 = std::sqrt(p0.maxSq), maxS1 = std::sqrt(p1.maxSq);
        const double multNorm = maxW0 * maxS1 + maxW1 * maxS0 + maxW0 * maxW1;
        return multNorm > prec && !(p0.leaf && p1.leaf);
    }
};

void combine(Mode mode, double prec, FunctionTree<3> &out, const FunctionTreeVector<3> &inp, int maxIter, bool absPrec, bool useMaxNorms = false) {
    const char *what = mode == Sum ? "add" : "multiply";
    if (inp.empty()) throw Error(std::string(what) + ": no input trees");
    if (inp.size() > MWG_MAX_OPERANDS) throw Error(std::string(what) + ": at most 8 input trees");
    for (const auto &t : inp) {
        FunctionTree<3> *f = std::get<1>(t);
        if (!f) throw Error(std::string(what) + ": null input tree");
        if (f == &out) throw Error(std::string(what) + ": the output tree cannot be an input");
        if (out.getMRA() != f->getMRA()) throw Error("Incompatible MRA");
    }
    const MultiResolutionAnalysis<3> &mra = out.getMRA();
    const int kp1 = mra.getKp1();
    if (kp1 != 4 && kp1 != 6 && kp1 != 8 && kp1 != 10) throw Error(std::string(what) + " supports Legendre orders 3, 5, 7 and 9");
    device::ensure_init();
    const int nin = (int)inp.size(), ncoef = 8 * kp1 * kp1 * kp1;
    const double *filters = device::filters(mra.getOrder());
    const device::BasisTables &bt = device::basis_tables(mra);
    std::vector<device::TreeMirror *> fm(nin);
    for (int i = 0; i < nin; i++) fm[i] = &std::get<1>(inp[i])->inputMirror();

    mwg_combine cb{};
    cb.kp1 = kp1;
    cb.mode = mode;
    cb.nin = nin;
    for (int d = 0; d < 3; d++) cb.sf[d] = mra.getWorldBox().getScalingFactor(d);
    for (int i = 0; i < nin; i++) cb.c[i] = std::get<0>(inp[i]);
    cb.filters = filters;
    cb.vcmap = bt.vcmap.get();
    cb.cvmap = bt.cvmap.get();

    // the operands' largest norms, before any node is generated below their leaves
    std::unique_ptr<MaxNormSplit> maxNorms;
    if (useMaxNorms) {
        if (inp.size() != 2) throw Error("multiply: the max-norm refinement needs exactly two input trees");
        maxNorms.reset(new MaxNormSplit{detail::MaxNorms(std::get<1>(inp[0])->nodes()), detail::MaxNorms(std::get<1>(inp[1])->nodes())});
    }
    ArithScratch &S = arith_scratch();
    detail::NodeStore<3> &g = out.topology();
    std::vector<int> work = detail::end_nodes(g); // TreeCalculator::getInitialWorkVector
    double sNorm = 0.0, wNorm = 0.0;
    const int maxScale = mra.getMaxScale();
    for (int iter = 0; !work.empty(); iter++) {
        const size_t count = work.size();
        // operand nodes, generated below the operands' leaves where missing
        std::vector<long long> offs(count * nin + count);
        std::vector<int> scales(count);
        for (int i = 0; i < nin; i++) {
            std::vector<device::GenNode> made;
            std::vector<int> ids(count);
            for (size_t k = 0; k < count; k++) ids[k] = device::plan_mirror_node(*fm[i], mra, g.scale(work[k]), g.transl(work[k]), made);
            device::generate_mirror_nodes(*fm[i], mra, filters, made, S.staging);
            for (size_t k = 0; k < count; k++) {
                const long long off = fm[i]->hostOff[ids[k]];
                offs[k * nin + i] = ids[k] >= fm[i]->nReal ? -off - 1 : off;
            }
        }
        for (int i = 0; i < nin; i++) cb.pool[i] = fm[i]->coef.get();
        for (size_t k = 0; k < count; k++) {
            offs[count * nin + k] = (long long)work[k] * ncoef;
            scales[k] = g.scale(work[k]);
        }
        out.deviceCoefs().reserve((size_t)g.size() * ncoef, (size_t)g.size() * ncoef);
        S.offs.reserve(offs.size());
        S.offs.upload(offs.data(), offs.size());
        S.scale.reserve(count);
        S.scale.upload(scales.data(), count);
        S.norms.reserve(9 * count);
        device::check(mwg_combine_nodes(&cb, S.offs.get(), S.scale.get(), S.offs.get() + count * nin, out.deviceCoefs().get(), S.norms.get(),
                                        (int)count),
                      what);
        std::vector<double> nrm(9 * count);
        S.norms.download(nrm.data(), nrm.size());
        for (size_t k = 0; k < count; k++) {
            double *cn = g.norms(work[k]);
            for (int t = 0; t < 8; t++) cn[t] = nrm[9 * k + t];
            g.setSquareNorm(work[k], nrm[9 * k + 8]);
            g.setHasCoefs(work[k], true);
        }
        // TreeBuilder::build bookkeeping and WaveletAdaptor splits
        if (iter == 0) {
            sNorm = 0.0;
            for (int u : work)
                if (g.depth(u) >= 0) sNorm += g.scalingNorm(u);
        }
        // the step's wavelet norms summed first, then added once
        // (wNorm += calcWaveletNorm(*workVec), src/treebuilders/TreeBuilder.cpp:58)
        double stepW = 0.0;
        for (int u : work)
            if (g.depth(u) >= 0) stepW += g.waveletNorm(u);
        wNorm += stepW;
        g.treeSquareNorm = (sNorm < 0.0 || wNorm < 0.0) ? -1.0 : sNorm + wNorm;
        if (iter >= maxIter && maxIter >= 0) work.clear();
        std::vector<int> next;
        for (int u : work) {
            if (g.hasChildren(u) || g.scale(u) + 2 > maxScale) continue;
            const bool split = maxNorms ? (*maxNorms)(prec, g.scale(u), g.transl(u))
                                        : detail::wavelet_split(prec, g.treeSquareNorm, absPrec, g.scale(u), g.waveletNorm(u));
            if (split) {
                const int c0 = g.createChildren(u, false);
                for (int c = 0; c < 8; c++) next.push_back(c0 + c);
            }
        }
        work.swap(next);
    }
    out.deviceCoefs().reserve((size_t)g.size() * ncoef, (size_t)g.size() * ncoef);
    compress_and_norms(out, filters);
    out.markDeviceResident();
    // FunctionTree::deleteGenerated on the inputs
    std::set<device::TreeMirror *> seen;
    for (int i = 0; i < nin; i++)
        if (seen.insert(fm[i]).second) device::drop_generated(*fm[i], mra, std::get<1>(inp[i])->topology());
}

} // namespace

double detail::dot_device(FunctionTree<3> &bra, FunctionTree<3> &ket) {
    device::ensure_init();
    const detail::NodeStore<3> &a = bra.topology(), &b = ket.topology();
    device::TreeMirror &ma = bra.inputMirror();
    device::TreeMirror &mb = ket.inputMirror();
    std::vector<long long> offs;
    std::vector<int> root;
    std::vector<long long> boff;
    for (int u : detail::real_nodes(a)) {
        const int v = b.findReal(a.scale(u), a.transl(u));
        if (v < 0) continue;
        offs.push_back(ma.hostOff[u]);
        boff.push_back(mb.hostOff[v]);
        root.push_back(a.depth(u) == 0 ? 1 : 0);
    }
    const size_t n = root.size();
    if (n == 0) return 0.0;
    offs.insert(offs.end(), boff.begin(), boff.end());
    ArithScratch &S = arith_scratch();
    S.offs.reserve(offs.size());
    S.offs.upload(offs.data(), offs.size());
    S.scale.reserve(n);
    S.scale.upload(root.data(), n);
    S.norms.reserve(2 * n);
    device::check(mwg_dot_nodes(a.kp1(), ma.coef.get(), S.offs.get(), mb.coef.get(), S.offs.get() + n, S.scale.get(), S.norms.get(), (int)n), "dot");
    std::vector<double> part(2 * n);
    S.norms.download(part.data(), part.size());
    double result = 0.0;
    for (size_t k = 0; k < n; k++) {
        if (root[k]) result += part[2 * k];
        result += part[2 * k + 1];
    }
    return result;
}

void FunctionTree<3>::rescale(double c) {
    detail::NodeStore<3> &s = *store_;
    const std::vector<int> ends = detail::end_nodes(s);
    if (host_stale_) {
        // end nodes times c in place on the device (a one-term sum), then the
        // bottom-up transform and the norms there
        device::ensure_init();
        const int kp1 = s.kp1(), ncoef = s.ncoef();
        mwg_combine cb{};
        cb.kp1 = kp1;
        cb.mode = Sum;
        cb.nin = 1;
        cb.c[0] = c;
        cb.pool[0] = dev_coef_.get();
        cb.filters = device::filters(mra_.getOrder());
        for (int d = 0; d < 3; d++) cb.sf[d] = mra_.getWorldBox().getScalingFactor(d);
        const size_t n = ends.size();
        std::vector<long long> offs(2 * n);
        std::vector<int> scales(n);
        for (size_t k = 0; k < n; k++) {
            offs[k] = offs[n + k] = (long long)ends[k] * ncoef;
            scales[k] = s.scale(ends[k]);
        }
        ArithScratch &S = arith_scratch();
        S.offs.reserve(offs.size());
        S.offs.upload(offs.data(), offs.size());
        S.scale.reserve(n);
        S.scale.upload(scales.data(), n);
        S.norms.reserve(9 * n);
        device::check(mwg_combine_nodes(&cb, S.offs.get(), S.scale.get(), S.offs.get() + n, dev_coef_.get(), S.norms.get(), (int)n), "rescale");
        compress_and_norms(*this, cb.filters);
        markDeviceResident();
        return;
    }
    const int nc = s.ncoef();
    for (int u : ends) {
        double *x = s.coefs(u);
        for (int j = 0; j < nc; j++) x[j] *= c;
        detail::compute_norms(s, u);
    }
    detail::transform_up(s);
    detail::calc_square_norm(s);
    invalidate();
}

void add(double prec, FunctionTree<3> &out, double a, FunctionTree<3> &inp_a, double b, FunctionTree<3> &inp_b, int maxIter, bool absPrec,
         bool /*conjugate*/) {
    combine(Sum, prec, out, {{a, &inp_a}, {b, &inp_b}}, maxIter, absPrec);
}

void add(double prec, FunctionTree<3> &out, const FunctionTreeVector<3> &inp, int maxIter, bool absPrec, bool /*conjugate*/) {
    combine(Sum, prec, out, inp, maxIter, absPrec);
}

void add(double prec, FunctionTree<3> &out, const std::vector<FunctionTree<3> *> &inp, int maxIter, bool absPrec, bool conjugate) {
    FunctionTreeVector<3> v;
    for (auto *t : inp) v.push_back({1.0, t});
    add(prec, out, v, maxIter, absPrec, conjugate);
}

void multiply(double prec, FunctionTree<3> &out, double c, FunctionTree<3> &inp_a, FunctionTree<3> &inp_b, int maxIter, bool absPrec,
              bool useMaxNorms, bool conjugate) {
    multiply(prec, out, {{c, &inp_a}, {1.0, &inp_b}}, maxIter, absPrec, useMaxNorms, conjugate);
}

void multiply(double prec, FunctionTree<3> &out, const FunctionTreeVector<3> &inp, int maxIter, bool absPrec, bool useMaxNorms,
              bool /*conjugate*/) {
    combine(Product, prec, out, inp, maxIter, absPrec, useMaxNorms);
}

void multiply(double prec, FunctionTree<3> &out, const std::vector<FunctionTree<3> *> &inp, int maxIter, bool absPrec, bool useMaxNorms,
              bool conjugate) {
    FunctionTreeVector<3> v;
    for (auto *t : inp) v.push_back({1.0, t});
    multiply(prec, out, v, maxIter, absPrec, useMaxNorms, conjugate);
}

} // namespace mrcpp
