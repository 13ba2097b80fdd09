// Analytic Gaussian sources and their projection onto a multiwavelet tree
// (host FP64): GaussFunc / GaussExp, build_grid and project.
#pragma once

#include <array>
#include <cmath>
#include <vector>

#include "trees.h"

namespace mrcpp {

// Isotropic Cartesian Gaussian coef * exp(-alpha |r - pos|^2)
// (src/functions/GaussFunc.cpp:47-66, src/functions/Gaussian.cpp:116-135).
template <int D> class GaussFunc {
public:
    // GaussFunc(beta, alpha, pos, pow) (src/functions/GaussFunc.h:56): exponent,
    // coefficient, centre; Cartesian powers other than zero are not provided
    GaussFunc(double alpha, double coef, const Coord<D> &pos = {}, const std::array<int, D> &pow = {})
            : coef_(coef)
            , pos_(pos) {
        alpha_.fill(alpha);
        for (int p : pow)
            if (p != 0) throw Error("GaussFunc: only zero Cartesian powers are supported");
    }
    // Coulomb interaction of two unit-charge Gaussians
    // (GaussFunc::calcCoulombEnergy, src/functions/GaussFunc.cpp:210-237):
    // sqrt(4 a / pi) F0(a R^2) with a = p q / (p + q) and the Boys function
    // F0(x) = sqrt(pi / x) erf(sqrt(x)) / 2, F0(0) = 1.
    double calcCoulombEnergy(const GaussFunc<D> &g) const {
        if (D != 3) throw Error("calcCoulombEnergy: three dimensions only");
        const double p = alpha_[0], q = g.alpha_[0];
        const double a = p * q / (p + q);
        double r2 = 0.0;
        for (int d = 0; d < D; d++) r2 += (pos_[d] - g.pos_[d]) * (pos_[d] - g.pos_[d]);
        const double x = a * r2;
        const double f0 = x < 1e-14 ? 1.0 - x / 3.0 : 0.5 * std::sqrt(pi / x) * std::erf(std::sqrt(x));
        return std::sqrt(4.0 * a / pi) * f0;
    }
    double evalf(const Coord<D> &r) const {
        double q2 = 0.0;
        for (int d = 0; d < D; d++) {
            const double q = r[d] - pos_[d];
            q2 += alpha_[d] * q * q;
        }
        return coef_ * 1.0 * std::exp(-q2);
    }
    // visible once the node is finer than the Gaussian's width
    bool isVisibleAtScale(int scale, int nQuadPts) const {
        for (double a : alpha_) {
            const double sd = std::pow(2.0 * a, -0.5);
            const int vis = static_cast<int>(-std::floor(std::log2(nQuadPts * 0.5 * sd)));
            if (scale < vis) return false;
        }
        return true;
    }
    // zero outside pos +- 5 standard deviations
    bool isZeroOnInterval(const double *a, const double *b) const {
        for (int d = 0; d < D; d++) {
            const double sd = std::pow(2.0 * alpha_[d], -0.5);
            if (a[d] > pos_[d] + 5.0 * sd || b[d] < pos_[d] - 5.0 * sd) return true;
        }
        return false;
    }
    double getCoef() const { return coef_; }
    void setCoef(double c) { coef_ = c; }
    const std::array<double, D> &getExp() const { return alpha_; }
    const Coord<D> &getPos() const { return pos_; }

private:
    std::array<double, D> alpha_;
    double coef_;
    Coord<D> pos_;
};

// Sum of Gaussians (src/functions/GaussExp.cpp:80-98)
template <int D> class GaussExp {
public:
    void append(const GaussFunc<D> &g) { funcs_.push_back(g); }
    int size() const { return (int)funcs_.size(); }
    const GaussFunc<D> &getFunc(int i) const { return funcs_[i]; }
    double evalf(const Coord<D> &r) const {
        double v = 0.0;
        for (const auto &g : funcs_) v += g.evalf(r);
        return v;
    }

private:
    std::vector<GaussFunc<D>> funcs_;
};

namespace detail {

// Child quadrature points, function values, value->coefficient transform and
// compression for one node (ProjectionCalculator::calcNode,
// src/treebuilders/ProjectionCalculator.cpp:34-51; MWNode::cvTransform,
// src/trees/MWNode.cpp:448-490).
template <int D, class F> void project_node(NodeStore<D> &s, int u, const F &func) {
    constexpr int TD = NodeStore<D>::TD;
    const auto &m = s.mra();
    const auto &bas = m.basis();
    const int kp1 = s.kp1(), kd = s.kp1_d();
    const auto &sf = m.getWorldBox().getScalingFactors();
    const double h = std::pow(2.0, -(s.scale(u) + 1));
    const auto &l = s.transl(u);
    std::vector<double> pts[D][2];
    for (int d = 0; d < D; d++)
        for (int half = 0; half < 2; half++) {
            pts[d][half].resize(kp1);
            for (int i = 0; i < kp1; i++) {
                const double base = bas.nodes()[i] + 2.0 * static_cast<double>(l[d]);
                pts[d][half][i] = h * (half ? base + 1.0 : base);
            }
        }
    double *c = s.coefs(u);
    for (int t = 0; t < TD; t++)
        for (int p = 0; p < kd; p++) {
            Coord<D> r;
            int q = p;
            for (int d = 0; d < D; d++) {
                r[d] = sf[d] * pts[d][(t >> d) & 1][q % kp1];
                q /= kp1;
            }
            c[(size_t)t * kd + p] = func.evalf(r);
        }
    std::vector<double> tmp((size_t)TD * kd);
    double *src = c, *dst = tmp.data();
    const int m1 = kd / kp1;
    for (int pass = 0; pass < D; pass++) {
        for (int t = 0; t < TD; t++) core::filter_pass(dst + (size_t)t * kd, src + (size_t)t * kd, bas.vcMap(), m1, false);
        std::swap(src, dst);
    }
    double sf_prod = 1.0;
    for (int d = 0; d < D; d++) sf_prod *= sf[d];
    if (sf_prod <= MachineZero) sf_prod = 1.0;
    const double fac = std::sqrt(1.0 / (std::pow(2.0, D * (s.scale(u) + 1)) / sf_prod));
    if (D % 2 == 1) {
        for (int i = 0; i < TD * kd; i++) c[i] = fac * src[i];
    } else {
        for (int i = 0; i < TD * kd; i++) c[i] *= fac;
    }
    local_transform(s, u, Compression);
    s.setHasCoefs(u, true);
    compute_norms(s, u);
}

template <int D> bool analytic_split(const NodeStore<D> &s, int u, const GaussFunc<D> &g) {
    const auto &m = s.mra();
    if (g.isVisibleAtScale(s.scale(u), s.kp1())) return false;
    double lb[D], ub[D];
    const double w = std::pow(2.0, -s.scale(u));
    for (int d = 0; d < D; d++) {
        const double sfd = m.getWorldBox().getScalingFactor(d);
        lb[d] = sfd * w * s.transl(u)[d];
        ub[d] = sfd * w * (s.transl(u)[d] + 1);
    }
    return !g.isZeroOnInterval(lb, ub);
}

} // namespace detail

// build_grid(tree, GaussExp): refine the empty grid once per Gaussian
// (src/treebuilders/grid.cpp:106-123 with AnalyticAdaptor.h:42-50).
template <int D> void build_grid_nodes(detail::NodeStore<D> &s, const GaussExp<D> &g, int maxIter = -1) {
    for (int i = 0; i < g.size(); i++) {
        const GaussFunc<D> &gi = g.getFunc(i);
        detail::BuildHooks<D> h;
        h.initial = [](detail::NodeStore<D> &st) { return detail::end_nodes(st); };
        h.compute = [](detail::NodeStore<D> &st, const std::vector<int> &v) {
            for (int u : v) {
                st.setHasCoefs(u, false);
                st.clearNorms(u);
            }
        };
        h.split = [&gi](const detail::NodeStore<D> &st, int u) { return detail::analytic_split(st, u, gi); };
        detail::build_tree(s, h, maxIter);
    }
}

// project(prec, tree, func) (src/treebuilders/project.cpp:85-102)
template <int D, class F> void project_nodes(detail::NodeStore<D> &s, double prec, const F &func, int maxIter = -1, bool absPrec = false) {
    detail::BuildHooks<D> h;
    h.initial = [](detail::NodeStore<D> &st) { return detail::end_nodes(st); };
    h.compute = [&func](detail::NodeStore<D> &st, const std::vector<int> &v) {
        const int nn = (int)v.size();
#pragma omp parallel for schedule(dynamic, 4) if (nn > 8)
        for (int i = 0; i < nn; i++) detail::project_node(st, v[i], func);
    };
    h.split = [prec, absPrec](const detail::NodeStore<D> &st, int u) {
        return detail::wavelet_split(prec, st.treeSquareNorm, absPrec, st.scale(u), st.waveletNorm(u));
    };
    detail::build_tree(s, h, maxIter);
    detail::transform_up(s);
    detail::calc_square_norm(s);
}

} // namespace mrcpp
