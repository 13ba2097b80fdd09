// Operator construction (host FP64) and band tables for the device kernels.
#include "operators.h"

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "device.h"

namespace mrcpp {

namespace {
thread_local OperatorTarget g_operator_target = OperatorTarget::Auto;
} // namespace

void set_operator_target(OperatorTarget t) { g_operator_target = t; }

std::vector<KernelTerm> poisson_kernel(double epsilon, double r_min, double r_max) {
    // Trapezoidal rule on the sinh-substituted integral representation of
    // 1/r, built on [r_min/r_max, 1] and rescaled to [r_min, r_max].  The
    // long-double literals reproduce the reference's mixed-precision steps.
    const double r0 = r_min / r_max;
    const double r1 = r_max;
    double t1 = 1.0L;
    while ((2.0 * t1 * std::exp(-t1)) > epsilon) t1 *= 1.1L;
    double t2 = 1.0L;
    while ((std::sqrt(t2) * std::exp(-t2) / r0) > epsilon) t2 *= 1.1L;
    const double s1 = -std::log(2.0 * t1);
    const double s2 = std::log(t2 / (r0 * r0)) / 2.0;
    const double h = 1.0 / (0.2L - 0.47L * std::log10(epsilon));
    const int n = static_cast<int>(std::ceil((s2 - s1) / h) + 1);
    if (n > MaxSepRank) throw Error("Maximum separation rank exceeded.");
    std::vector<KernelTerm> out;
    out.reserve(n);
    for (int i = 0; i < n; i++) {
        const double s = s1 + h * i;
        const double sh = std::sinh(s), ch = std::cosh(s);
        const double e = 1.0 + std::exp(-sh);
        double expo = 4.0L * (sh + std::log(e)) * (sh + std::log(e));
        double coef = h * (4.0L / root_pi) * ch / e;
        expo *= 1.0 / (r1 * r1);
        coef *= 1.0 / r1;
        if (i == 0 || i == n - 1) coef *= 1.0 / 2.0;
        out.push_back({expo, coef});
    }
    return out;
}

std::vector<KernelTerm> helmholtz_kernel(double mu, double epsilon, double r_min, double r_max) {
    const double r0 = r_min / r_max;
    const double r1 = r_max;
    const double mt = mu * r1;
    const double t = std::max((-2.5L * std::log(epsilon)), 5.0L);
    const double s1 = -std::log(4 * t / (mt * mt)) / 2;
    const double s2 = std::log(t / (r0 * r0)) / 2;
    const double h = 1.0 / (0.20L - 0.47L * std::log10(epsilon));
    const int n = static_cast<int>(std::ceil((s2 - s1) / h) + 1);
    if (n > MaxSepRank) throw Error("Maximum separation rank exceeded.");
    std::vector<KernelTerm> out;
    out.reserve(n);
    for (int i = 0; i < n; i++) {
        const double s = s1 + h * i;
        const double e2 = -s * 2.0;
        const double e3 = -mt * mt * std::exp(e2) / 4.0 + s;
        double beta = (h * (2.0 / root_pi) * std::exp(e3));
        const double two_s = 2.0L * s;
        double alpha = std::exp(two_s);
        alpha *= 1.0 / (r1 * r1);
        beta *= 1.0 / r1;
        if (i == 0 || i == n - 1) beta *= 1.0 / 2.0;
        out.push_back({alpha, beta});
    }
    return out;
}

namespace {

using detail::NodeStore;

// Cross-correlation of the projected 1D kernel into the children of one 2D
// node, then local compression (CrossCorrelationCalculator::calcNode/applyCcc,
// src/treebuilders/CrossCorrelationCalculator.cpp:36-96).
void cross_correlate_node(NodeStore<2> &ot, int u, NodeStore<1> &kt) {
    ot.zeroCoefs(u);
    const int K = ot.kp1(), kd = ot.kp1_d(), K2 = 2 * K;
    const auto &cc = core::cross_correlation(ot.mra().getOrder());
    const int scale = ot.scale(u) + 1;
    const auto &l = ot.transl(u);
    std::vector<double> v(4 * kd, 0.0);
    for (int c = 0; c < 4; c++) {
        const int l0 = 2 * l[0] + (c & 1), l1 = 2 * l[1] + ((c >> 1) & 1);
        const int a = detail::get_or_generate(kt, scale, {l1 - l0 - 1});
        const int b = detail::get_or_generate(kt, scale, {l1 - l0});
        const double *sa = kt.coefs(a), *sb = kt.coefs(b);
        for (int r = 0; r < kd; r++) {
            const double *L = cc.left() + (size_t)r * K2, *R = cc.right() + (size_t)r * K2;
            double xa = 0.0, xb = 0.0;
            for (int p = 0; p < K2; p++) {
                xa += L[p] * sa[p];
                xb += R[p] * sb[p];
            }
            v[(size_t)c * kd + r] = xa + xb;
        }
    }
    const double two_n = std::pow(2.0, -scale / 2.0);
    const double sf = ot.mra().getWorldBox().getScalingFactor(0);
    double *coef = ot.coefs(u);
    for (int i = 0; i < 4 * kd; i++) coef[i] = std::sqrt(sf) * two_n * v[i];
    detail::local_transform(ot, u, Compression);
    ot.setHasCoefs(u, true);
    detail::compute_norms(ot, u);
}

std::vector<int> max_translations(const NodeStore<2> &t) {
    std::vector<int> mt(t.getDepth(), 0);
    for (int u : detail::real_nodes(t)) {
        const int d = t.depth(u);
        mt[d] = std::max(mt[d], std::max(std::abs(t.transl(u)[0]), std::abs(t.transl(u)[1])));
    }
    return mt;
}

// The 1D kernel tree of one Gaussian term, projected at the kernel order
// (ConvolutionOperator::initialize, src/operators/ConvolutionOperator.cpp:78-108).
std::unique_ptr<NodeStore<1>> kernel_tree(const MultiResolutionAnalysis<1> &kmra, const KernelTerm &term, double k_prec) {
    GaussFunc<1> kf(term.alpha, std::copysign(std::pow(std::abs(term.coef), 1.0 / 3), term.coef));
    auto kt = std::make_unique<NodeStore<1>>(kmra);
    GaussExp<1> ge;
    ge.append(kf);
    build_grid_nodes(*kt, ge);
    project_nodes(*kt, k_prec, kf);
    return kt;
}

// An operator term with its 2D tree at the root boxes.
std::unique_ptr<detail::OperatorTerm> term_shell(const MultiResolutionAnalysis<3> &mra, double o_prec, int reach) {
    auto term_out = std::make_unique<detail::OperatorTerm>();
    const double sf = mra.getWorldBox().getScalingFactor(0);
    term_out->mra = std::make_unique<MultiResolutionAnalysis<2>>(
        BoundingBox<2>(mra.getRootScale(), {0, 0}, {reach, reach}, {sf, sf}), LegendreBasis(mra.getOrder()));
    term_out->tree = std::make_unique<NodeStore<2>>(*term_out->mra, true, o_prec);
    return term_out;
}

// OperatorAdaptor::splitNode (src/treebuilders/OperatorAdaptor.h:38-48)
bool operator_split(const NodeStore<2> &st, int u) {
    const auto &l = st.transl(u);
    if (!(l[0] == 0 || l[1] == 0)) return false;
    for (int c = 1; c < 4; c++)
        if (st.compNorm(u, c) > 0.0) return true;
    return false;
}

void finish_term(detail::OperatorTerm &term);

// One Gaussian term: 1D kernel projection, 2D tree by cross-correlation,
// bottom-up transform and translation cache (ConvolutionOperator::initialize,
// src/operators/ConvolutionOperator.cpp:78-108).
std::unique_ptr<detail::OperatorTerm> build_term(const MultiResolutionAnalysis<3> &mra, const MultiResolutionAnalysis<1> &kmra,
                                                  const KernelTerm &term, double k_prec, double o_prec, int reach) {
    std::unique_ptr<NodeStore<1>> kt = kernel_tree(kmra, term, k_prec);
    auto term_out = term_shell(mra, o_prec, reach);
    NodeStore<2> &ot = *term_out->tree;
    detail::BuildHooks<2> h;
    h.initial = [](NodeStore<2> &st) { return detail::end_nodes(st); };
    h.compute = [&kt](NodeStore<2> &st, const std::vector<int> &v) {
        for (int u : v) cross_correlate_node(st, u, *kt);
    };
    h.split = operator_split;
    detail::build_tree(ot, h, -1);
    finish_term(*term_out);
    return term_out;
}

// Bottom-up transform, tree norm and translation cache of a built term.
void finish_term(detail::OperatorTerm &term) {
    auto *term_out = &term;
    NodeStore<2> &ot = *term.tree;
    detail::transform_up(ot);
    detail::calc_square_norm(ot);

    // translation cache: l > 0 -> node (l, 0), l <= 0 -> node (0, -l)
    term_out->maxTransl = max_translations(ot);
    const int ns = ot.getDepth();
    term_out->cache.resize(ns);
    for (int d = 0; d < ns; d++) {
        const int nt = term_out->maxTransl[d], scale = ot.rootScale() + d;
        auto &row = term_out->cache[d];
        row.resize(2 * nt + 1);
        for (int i = nt; i >= 0; i--) row[nt - i] = detail::get_or_generate(ot, scale, {0, i});
        for (int i = 1; i <= nt; i++) row[nt + i] = detail::get_or_generate(ot, scale, {i, 0});
    }
}

// Device tables of the cross-correlation for an order (left, right), kept
// until the process ends (never destroyed: no device call from a static destructor).
struct CccTables {
    device::DeviceArray<double> left, right;
};
const CccTables &ccc_tables(int order) {
    static std::map<int, CccTables *> memo;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto &slot = memo[order];
    if (!slot) {
        slot = new CccTables();
        const auto &cc = core::cross_correlation(order);
        const size_t n = (size_t)(order + 1) * (order + 1) * 2 * (order + 1);
        slot->left.reserve(n);
        slot->left.upload(cc.left(), n);
        slot->right.reserve(n);
        slot->right.upload(cc.right(), n);
    }
    return *slot;
}

// The 2D trees of all terms built together, one refinement level of every
// term per device launch (TreeBuilder::build per term with
// CrossCorrelationCalculator and OperatorAdaptor).  The host keeps each
// term's topology, approximate norm and split test; the device forms the
// nodes bit-identically to cross_correlate_node.
// Returns the seconds spent packing, on the device round trips and on the
// host's per-level work.
std::array<double, 3> build_trees_device(const std::vector<detail::OperatorTerm *> &terms, const std::vector<NodeStore<1> *> &kts,
                                         const MultiResolutionAnalysis<3> &mra) {
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); };
    std::array<double, 3> times{};
    device::ensure_init();
    const int M = (int)terms.size();
    const int K = mra.getOrder() + 1, kd = K * K, K2 = 2 * K;
    const CccTables &cc = ccc_tables(mra.getOrder());
    const double *filt = device::filters(mra.getOrder());
    const double sf = mra.getWorldBox().getScalingFactor(0);
    struct State {
        std::vector<int> work;
        double sNorm = 0.0, wNorm = 0.0;
        int iter = 0;
    };
    std::vector<State> st(M);
    for (int i = 0; i < M; i++) st[i].work = detail::end_nodes(*terms[i]->tree);
    // Per level one upload from page-locked memory, [kernel blocks | factors |
    // thresholds | block slots], and one download [coefficients (4 K^2 per
    // node) | norms (5 per node)].  A node at (l0, l1) reads the kernel blocks
    // at translations D - 2 .. D + 1, D = 2 (l1 - l0); a term's nodes with
    // values of D less than 5 apart share one run of consecutive blocks, so a
    // level uploads about one block per node instead of eight.
    static std::mutex pin_mutex;
    static device::PinnedArray<double> &pin_in = *new device::PinnedArray<double>(), &pin_out = *new device::PinnedArray<double>();
    std::lock_guard<std::mutex> pin_lock(pin_mutex);
    device::DeviceArray<double> din, dout;
    std::vector<std::vector<int>> uniqD(M), blockOf(M);
    std::vector<size_t> kbase(M + 1, 0);
    for (;;) {
        std::vector<size_t> base(M + 1, 0);
        for (int i = 0; i < M; i++) base[i + 1] = base[i] + st[i].work.size();
        const size_t n = base[M];
        if (n == 0) break;
        auto t0 = clk::now();
        // runs of kernel blocks per term
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < M; i++) {
            const NodeStore<2> &ot = *terms[i]->tree;
            std::vector<int> &U = uniqD[i], &B = blockOf[i];
            U.clear();
            for (int u : st[i].work) U.push_back(2 * (ot.transl(u)[1] - ot.transl(u)[0]));
            std::sort(U.begin(), U.end());
            U.erase(std::unique(U.begin(), U.end()), U.end());
            B.resize(U.size());
            int total = 0;
            for (size_t q = 0; q < U.size(); q++) {
                B[q] = (q > 0 && U[q] - U[q - 1] <= 4) ? B[q - 1] + (U[q] - U[q - 1]) : total;
                total = B[q] + 4;
            }
            kbase[i + 1] = (size_t)total;
        }
        for (int i = 0; i < M; i++) kbase[i + 1] += kbase[i];
        const size_t nblk = kbase[M];
        if (nblk > (size_t)INT32_MAX) throw Error("operator level too large");
        const size_t nin = nblk * K2 + 2 * n + (n + 1) / 2;
        pin_in.reserve(nin);
        double *blocks = pin_in.data(), *fac = blocks + nblk * K2, *thr = fac + n;
        int *slot = reinterpret_cast<int *>(thr + n);
        int mixed = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : mixed)
        for (int i = 0; i < M; i++) {
            NodeStore<2> &ot = *terms[i]->tree;
            NodeStore<1> &kt = *kts[i];
            const std::vector<int> &U = uniqD[i], &B = blockOf[i];
            if (st[i].work.empty()) continue;
            const int scale = ot.scale(st[i].work[0]) + 1;
            for (size_t q = 0; q < U.size(); q++) {
                // blocks of this run not yet written by the previous value of D
                const int first = (q > 0 && U[q] - U[q - 1] <= 4) ? 4 - (U[q] - U[q - 1]) : 0;
                for (int off = first; off < 4; off++) {
                    const int a = detail::get_or_generate(kt, scale, {U[q] - 2 + off});
                    std::memcpy(blocks + (kbase[i] + B[q] + off) * K2, kt.coefs(a), sizeof(double) * K2);
                }
            }
            for (size_t j = 0; j < st[i].work.size(); j++) {
                const int u = st[i].work[j];
                const size_t k = base[i] + j;
                mixed |= ot.scale(u) + 1 != scale;
                const int D = 2 * (ot.transl(u)[1] - ot.transl(u)[0]);
                const size_t q = std::lower_bound(U.begin(), U.end(), D) - U.begin();
                slot[k] = (int)(kbase[i] + B[q]);
                const double two_n = std::pow(2.0, -scale / 2.0);
                fac[k] = std::sqrt(sf) * two_n;
                thr[k] = std::max(MachinePrec, ot.normPrec() / (8.0 * (1 << ot.depth(u))));
            }
        }
        // a term's work nodes are the children of one level, so one scale
        if (mixed) throw Error("operator tree level mixes scales");
        times[0] += since(t0);
        t0 = clk::now();
        din.reserve(nin);
        din.upload(pin_in.data(), nin);
        dout.reserve(n * (4 * kd + 5));
        const double *dblocks = din.get(), *dfac = dblocks + nblk * K2, *dthr = dfac + n;
        device::check(mwg_operator_nodes(K, cc.left.get(), cc.right.get(), filt, dblocks, reinterpret_cast<const int *>(dthr + n), dfac, dthr,
                                         dout.get(), dout.get() + n * 4 * kd, (int)n),
                      "operator nodes");
        pin_out.reserve(n * (4 * kd + 5));
        dout.download(pin_out.data(), n * (4 * kd + 5));
        const double *coef = pin_out.data(), *nrm = pin_out.data() + n * 4 * kd;
        times[1] += since(t0);
        t0 = clk::now();
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < M; i++) {
            NodeStore<2> &ot = *terms[i]->tree;
            State &S = st[i];
            if (S.work.empty()) continue;
            for (size_t j = 0; j < S.work.size(); j++) {
                const int u = S.work[j];
                const size_t k = base[i] + j;
                std::memcpy(ot.coefs(u), coef + k * 4 * kd, sizeof(double) * 4 * kd);
                double *cn = ot.norms(u);
                for (int t = 0; t < 4; t++) cn[t] = nrm[k * 5 + t];
                ot.setSquareNorm(u, nrm[k * 5 + 4]);
                ot.setHasCoefs(u, true);
            }
            // TreeBuilder::build bookkeeping (src/treebuilders/TreeBuilder.cpp:38-87)
            if (S.iter == 0) {
                S.sNorm = 0.0;
                for (int u : S.work)
                    if (ot.depth(u) >= 0) S.sNorm += ot.scalingNorm(u);
            }
            double stepW = 0.0; // the step's sum first, then added once (TreeBuilder.cpp:58)
            for (int u : S.work)
                if (ot.depth(u) >= 0) stepW += ot.waveletNorm(u);
            S.wNorm += stepW;
            ot.treeSquareNorm = (S.sNorm < 0.0 || S.wNorm < 0.0) ? -1.0 : S.sNorm + S.wNorm;
            std::vector<int> next;
            const int maxScale = ot.mra().getMaxScale();
            for (int u : S.work) {
                if (ot.hasChildren(u)) continue;
                if (ot.scale(u) + 2 > maxScale) continue;
                if (operator_split(ot, u)) {
                    const int c0 = ot.createChildren(u, false);
                    for (int c = 0; c < 4; c++) next.push_back(c0 + c);
                }
            }
            S.work.swap(next);
            S.iter++;
        }
        times[2] += since(t0);
    }
    return times;
}

// OperatorTree::calcBandWidth (src/trees/OperatorTree.cpp:109-134)
void term_band_width(detail::OperatorTerm &t, double prec) {
    NodeStore<2> &ot = *t.tree;
    const int depthN = ot.getDepth();
    t.widths.assign(depthN + 1, {-1, -1, -1, -1, -1});
    const std::vector<int> mt = max_translations(ot);
    if (prec < 0.0) prec = ot.normPrec();
    for (int d = 0; d < depthN; d++) {
        const double thr = std::max(MachinePrec, prec / (8.0 * (1 << d)));
        for (int l = 0;; l++) {
            if (l > t.maxTransl[d]) throw Error("band width beyond the operator translation cache");
            const int u = t.node(d, l);
            bool any = false;
            for (int c = 0; c < 4; c++)
                if (ot.compNorm(u, c) > thr) {
                    t.widths[d][c] = l;
                    t.widths[d][4] = std::max(t.widths[d][4], l);
                    any = true;
                }
            if (!any || l + 1 > mt[d]) break;
        }
    }
}

} // namespace

ConvolutionOperator<3>::ConvolutionOperator(const MultiResolutionAnalysis<3> &mra)
        : mra_(mra) {}

ConvolutionOperator<3>::ConvolutionOperator(const MultiResolutionAnalysis<3> &mra, const std::vector<KernelTerm> &kernel, double prec)
        : mra_(mra) {
    initialize(kernel, prec);
}

ConvolutionOperator<3>::~ConvolutionOperator() { device::destroy_band_tables(dev_tables_); }

namespace {

using clock_type = std::chrono::steady_clock;
double seconds_since(clock_type::time_point t0) { return std::chrono::duration<double>(clock_type::now() - t0).count(); }

// Root-box reach of the kernel MRA: order 2k+1 on [-reach, reach) root boxes
// (ConvolutionOperator::getKernelMRA, src/operators/ConvolutionOperator.cpp:110-142).
int kernel_reach(const MultiResolutionAnalysis<3> &mra) {
    int reach = 0;
    for (int d = 0; d < 3; d++) reach = std::max(reach, mra.getWorldBox().size(d));
    return reach;
}

MultiResolutionAnalysis<1> kernel_mra(const MultiResolutionAnalysis<3> &mra) {
    const int reach = kernel_reach(mra);
    const double sf = mra.getWorldBox().getScalingFactor(0);
    return MultiResolutionAnalysis<1>(BoundingBox<1>(mra.getRootScale(), {-reach}, {2 * reach}, {sf}), LegendreBasis(2 * mra.getOrder() + 1));
}

// the shared filter and cross-correlation caches, filled before a parallel region
void prime_tables(const MultiResolutionAnalysis<3> &mra) {
    core::cross_correlation(mra.getOrder());
    core::filters(2 * mra.getOrder() + 1);
}

} // namespace

void ConvolutionOperator<3>::prepare(const std::vector<KernelTerm> &kernel, double prec) {
    kernel_ = kernel;
    build_prec_ = prec;
    terms_.clear();
    terms_.resize(kernel.size());
    band_max_.clear();
    build_phases_ = {};
}

void ConvolutionOperator<3>::build_host() {
    const auto t0 = clock_type::now();
    const MultiResolutionAnalysis<1> kmra = kernel_mra(mra_);
    const int reach = kernel_reach(mra_);
    prime_tables(mra_);
    const double o_prec = build_prec_, k_prec = build_prec_ / 10.0;
    const int M = (int)kernel_.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < M; i++) terms_[i] = build_term(mra_, kmra, kernel_[i], k_prec, o_prec, reach);
    built_on_device_ = false;
    build_seconds_ = seconds_since(t0);
    build_phases_ = {0.0, 0.0, 0.0, build_seconds_, 0.0};
}

// The terms of all operators (one MRA) in one list: kernel trees and term
// shells on the host cores, the 2D trees one device launch per level for all,
// then the bottom-up transforms and caches on the host cores.
void ConvolutionOperator<3>::build_device(const std::vector<ConvolutionOperator<3> *> &ops) {
    if (ops.empty()) return;
    const auto t0 = clock_type::now();
    const MultiResolutionAnalysis<3> &mra = ops[0]->mra_;
    const MultiResolutionAnalysis<1> kmra = kernel_mra(mra);
    const int reach = kernel_reach(mra);
    prime_tables(mra);
    std::vector<std::pair<ConvolutionOperator<3> *, int>> all;
    for (auto *op : ops)
        for (int i = 0; i < (int)op->kernel_.size(); i++) all.push_back({op, i});
    const int T = (int)all.size();
    std::vector<std::unique_ptr<NodeStore<1>>> kts(T);
    std::vector<NodeStore<1> *> kraw(T);
    std::vector<detail::OperatorTerm *> traw(T);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < T; j++) {
        ConvolutionOperator<3> &op = *all[j].first;
        const int i = all[j].second;
        kts[j] = kernel_tree(kmra, op.kernel_[i], op.build_prec_ / 10.0);
        op.terms_[i] = term_shell(mra, op.build_prec_, reach);
        kraw[j] = kts[j].get();
        traw[j] = op.terms_[i].get();
    }
    const double tKernel = seconds_since(t0);
    const std::array<double, 3> tLevels = build_trees_device(traw, kraw, mra);
    const auto t2 = clock_type::now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < T; j++) finish_term(*traw[j]);
    const double tFinish = seconds_since(t2), total = seconds_since(t0);
    for (auto *op : ops) {
        op->built_on_device_ = true;
        op->build_seconds_ = total;
        op->build_phases_ = {tKernel, tLevels[0], tLevels[1], tLevels[2], tFinish};
    }
}

void ConvolutionOperator<3>::initialize(const std::vector<KernelTerm> &kernel, double prec) {
    prepare(kernel, prec);
    // Auto builds on the host cores: the device path pays one host round trip
    // per refinement level (about 25 for these kernels), which one operator's
    // few hundred nodes per level do not amortise
    if (g_operator_target == OperatorTarget::Device)
        build_device({this});
    else
        build_host();
}

std::vector<std::unique_ptr<ConvolutionOperator<3>>> build_operators(const MultiResolutionAnalysis<3> &mra, const std::vector<OperatorSpec> &specs,
                                                                     OperatorTarget target) {
    std::vector<std::unique_ptr<ConvolutionOperator<3>>> out;
    std::vector<ConvolutionOperator<3> *> raw;
    for (const OperatorSpec &sp : specs) {
        if (!(sp.prec > 0.0)) throw Error("operator precision must be positive");
        const double k_prec = sp.prec / 10.0;
        std::unique_ptr<ConvolutionOperator<3>> op(new ConvolutionOperator<3>(mra));
        op->mu_ = sp.mu > 0.0 ? sp.mu : 0.0;
        op->prepare(sp.mu > 0.0 ? helmholtz_kernel(sp.mu, k_prec, mra.calcMinDistance(k_prec), mra.calcMaxDistance())
                                : poisson_kernel(k_prec, mra.calcMinDistance(k_prec), mra.calcMaxDistance()),
                    sp.prec);
        raw.push_back(op.get());
        out.push_back(std::move(op));
    }
    if (target == OperatorTarget::Device) {
        ConvolutionOperator<3>::build_device(raw);
    } else {
        const auto t0 = clock_type::now();
        for (auto *op : raw) op->build_host();
        const double total = seconds_since(t0);
        for (auto *op : raw) op->build_seconds_ = total;
    }
    return out;
}

void ConvolutionOperator<3>::calcBandWidths(double prec) {
    int maxDepth = 0;
    for (auto &t : terms_) {
        term_band_width(*t, prec);
        maxDepth = std::max(maxDepth, t->bwDepth());
    }
    band_max_.assign(maxDepth + 1, -1);
    for (auto &t : terms_)
        for (int n = 0; n <= t->bwDepth(); n++)
            for (int c = 0; c < 4; c++) band_max_[n] = std::max(band_max_[n], t->width(n, c));
}

int ConvolutionOperator<3>::getMaxBandWidth(int depth) const {
    if (depth < 0) return band_max_.empty() ? -1 : *std::max_element(band_max_.begin(), band_max_.end());
    if (depth < (int)band_max_.size()) return band_max_[depth];
    return -1;
}

const detail::BandTables &ConvolutionOperator<3>::bandTables(double prec) {
    if (tables_ && table_prec_ == prec) return *tables_;
    calcBandWidths(prec);
    auto T = std::make_unique<detail::BandTables>();
    const int M = size(), kp1 = mra_.getKp1(), n2 = kp1 * kp1;
    T->M = M;
    T->kp1 = kp1;
    T->nDepth = (int)band_max_.size();
    T->bandMax = band_max_;
    T->Wmax = std::max(0, getMaxBandWidth(-1));
    const int ND = T->nDepth, W = T->Wmax;
    T->widths.assign((size_t)ND * M * 5, -1);
    T->bandSize.assign((size_t)ND * M * 64, 0);
    T->matStart.assign((size_t)ND * M, -1);
    T->pairBound.assign((size_t)ND * (2 * W + 1), 0.0);
    T->pairBoundW.assign((size_t)ND * (2 * W + 1), 0.0);
    T->bandSizeMax.assign((size_t)ND * M, 0);
    long long nmat = 0;
    for (int d = 0; d < ND; d++)
        for (int i = 0; i < M; i++) {
            const auto &t = *terms_[i];
            for (int c = 0; c < 5; c++) T->widths[((size_t)d * M + i) * 5 + c] = (c < 4) ? t.width(d, c) : t.maxWidth(d);
            // ConvolutionCalculator::calcBandSizeFactor (:116-139): a negative
            // width drops only its own factor
            int bsmax = 0;
            for (int gt = 0; gt < 8; gt++)
                for (int ft = 0; ft < 8; ft++) {
                    int tot = 1;
                    for (int dim = 0; dim < 3; dim++) {
                        const int w = t.width(d, 2 * ((gt >> dim) & 1) + ((ft >> dim) & 1));
                        if (w < 0) continue;
                        tot *= 2 * w + 1;
                    }
                    T->bandSize[((size_t)d * M + i) * 64 + gt * 8 + ft] = tot * 64;
                    bsmax = std::max(bsmax, tot * 64);
                }
            T->bandSizeMax[(size_t)d * M + i] = bsmax;
            const int w = t.maxWidth(d);
            if (w < 0) continue;
            T->matStart[(size_t)d * M + i] = nmat;
            nmat += (long long)(2 * w + 1) * 4;
            const double cb = std::cbrt((double)bsmax);
            for (int dl = -w; dl <= w; dl++) {
                const int u = t.node(d, dl);
                double mx = 0.0, mw = 0.0;
                for (int c = 0; c < 4; c++) mx = std::max(mx, t.tree->compNorm(u, c));
                for (int c = 1; c < 4; c++) mw = std::max(mw, t.tree->compNorm(u, c));
                double &pb = T->pairBound[(size_t)d * (2 * W + 1) + dl + W];
                pb = std::max(pb, cb * mx);
                double &pw = T->pairBoundW[(size_t)d * (2 * W + 1) + dl + W];
                pw = std::max(pw, cb * mw);
            }
        }
    T->mats.resize((size_t)nmat * n2);
    T->matNorm.resize(nmat);
    T->termNorm.assign(nmat / 4, 0.0);
    for (int d = 0; d < ND; d++)
        for (int i = 0; i < M; i++) {
            const long long s = T->matStart[(size_t)d * M + i];
            if (s < 0) continue;
            const auto &t = *terms_[i];
            const int w = t.maxWidth(d);
            for (int dl = -w; dl <= w; dl++) {
                const int u = t.node(d, dl);
                for (int c = 0; c < 4; c++) {
                    const long long m = s + (long long)(dl + w) * 4 + c;
                    std::memcpy(&T->mats[(size_t)m * n2], t.tree->coefs(u) + (size_t)c * n2, sizeof(double) * n2);
                    T->matNorm[m] = t.tree->compNorm(u, c);
                    T->termNorm[m / 4] = std::max(T->termNorm[m / 4], T->matNorm[m]);
                }
            }
        }
    tables_ = std::move(T);
    table_prec_ = prec;
    device::destroy_band_tables(dev_tables_);
    dev_tables_ = nullptr;
    return *tables_;
}

detail::DeviceBandTables *ConvolutionOperator<3>::deviceTables(double prec) {
    if (!(tables_ && table_prec_ == prec && dev_tables_)) {
        bandTables(prec);
        dev_tables_ = device::create_band_tables(*tables_);
    }
    return dev_tables_;
}

PoissonOperator::PoissonOperator(const MultiResolutionAnalysis<3> &mra, double prec)
        : ConvolutionOperator<3>(mra) {
    const double k_prec = prec / 10.0;
    initialize(poisson_kernel(k_prec, mra.calcMinDistance(k_prec), mra.calcMaxDistance()), prec);
}

HelmholtzOperator::HelmholtzOperator(const MultiResolutionAnalysis<3> &mra, double mu, double prec)
        : ConvolutionOperator<3>(mra) {
    mu_ = mu;
    const double k_prec = prec / 10.0;
    initialize(helmholtz_kernel(mu, k_prec, mra.calcMinDistance(k_prec), mra.calcMaxDistance()), prec);
}

} // namespace mrcpp
