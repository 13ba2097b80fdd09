// Numerical core of the MI355X MRCPP path: Legendre scaling basis, Gauss-Legendre
// nodes, multiwavelet filters and cross-correlation tables (host, FP64).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mrcpp {

// api/constants.h:30-57
constexpr double MachinePrec = 1.0e-15;
constexpr double MachineZero = 1.0e-14;
constexpr int MaxOrder = 41;
constexpr int MaxDepth = 30;
constexpr int MaxSepRank = 1000;
constexpr double pi = 3.1415926535897932384626433832795;
constexpr double root_pi = 1.7724538509055160273;

enum FuncType { Legendre = 0, Interpol = 1 };
enum MWDir { Compression = 0, Reconstruction = 1 };

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Directory holding the L_H0_k / L_G0_k / L_c_{left,right}_k tables.
void set_filter_dir(const std::string &dir);
const std::string &filter_dir();

namespace core {

// Small dense column-major square matrix.
struct SqMat {
    int n = 0;
    std::vector<double> a; // a[i + n*j] = M(i, j)
    SqMat() = default;
    explicit SqMat(int n_)
            : n(n_)
            , a((size_t)n_ * n_, 0.0) {}
    double &operator()(int i, int j) { return a[i + n * j]; }
    double operator()(int i, int j) const { return a[i + n * j]; }
    const double *data() const { return a.data(); }
};

// Scaling basis phi_j(x) = sqrt(2j+1) P_j(2x-1) on [0,1], j = 0..order, held as
// power-series coefficients in q = 2x - 1 (src/core/LegendreBasis.cpp:67-73,
// src/functions/LegendrePoly.cpp:62-86), plus the quadrature value->coefficient
// map vc(i,j) = phi_j(x_i) w_i (src/core/LegendreBasis.cpp:98-109).
class LegendreTables {
public:
    explicit LegendreTables(int order);
    int order() const { return order_; }
    double phi(int j, double x) const;
    const std::vector<double> &nodes() const { return nodes_; }
    const std::vector<double> &weights() const { return weights_; }
    const SqMat &vcMap() const { return vc_; }
    // coefficient -> value map, the inverse of vcMap formed numerically as
    // LegendreBasis::calcCVMaps does (src/core/LegendreBasis.cpp:98-109); in
    // exact arithmetic cv(i, j) = phi_i(x_j)
    const SqMat &cvMap() const { return cv_; }

private:
    int order_;
    std::vector<std::vector<double>> power_; // per j: coefficients of q^m
    std::vector<double> nodes_, weights_;
    SqMat vc_, cv_;
};

// Gauss-Legendre nodes/weights on [0,1] with the Newton iteration of
// src/core/GaussQuadrature.cpp:157-194 (tolerance 3e-12, at most 10 steps).
void gauss_legendre_01(int npts, std::vector<double> &x, std::vector<double> &w);

// Power-series coefficients of the Legendre polynomial P_k on [-1,1].
const std::vector<double> &legendre_power_coefs(int k);

// Legendre multiwavelet filter bank of one order (src/core/MWFilter.cpp:200-251).
// sub(idx, dir) returns the (k+1)x(k+1) block indexed by 2*outbit + inbit.
class FilterBank {
public:
    explicit FilterBank(int order);
    int order() const { return order_; }
    const SqMat &sub(int idx, int dir) const { return dir == Compression ? comp_[idx] : reco_[idx]; }

private:
    int order_;
    SqMat comp_[4], reco_[4];
};

// Cross-correlation coefficient tables (src/core/CrossCorrelation.cpp:99-124):
// rows (k+1)^2, columns 2(k+1), row-major; |c| < MachinePrec flushed to 0.
class CrossCorrTables {
public:
    explicit CrossCorrTables(int order);
    int order() const { return order_; }
    const double *left() const { return left_.data(); }
    const double *right() const { return right_.data(); }

private:
    int order_;
    std::vector<double> left_, right_;
};

// Process-wide caches, thread safe.
const LegendreTables &legendre(int order);
const FilterBank &filters(int order);
const CrossCorrTables &cross_correlation(int order);

// out(r, j) (=|+=) sum_i in(i, r) F(i, j); in is (n x m) and out (m x n),
// column-major, n = F.n (math_utils::apply_filter, src/utils/math_utils.cpp:175-194).
void filter_pass(double *out, const double *in, const SqMat &F, int m, bool accumulate);

inline int ipow(int b, int e) {
    int r = 1;
    while (e-- > 0) r *= b;
    return r;
}

} // namespace core
} // namespace mrcpp
