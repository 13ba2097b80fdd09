// Convolution operators: Gaussian-expansion kernels (Poisson, Helmholtz), their
// projection into per-term 2D non-standard-form operator trees (host FP64), band
// widths, and the flat band tables the MI355X kernels read.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "functions.h"
#include "trees.h"

namespace mrcpp {

struct KernelTerm {
    double alpha; // Gaussian exponent
    double coef;  // expansion coefficient
};

// Sinh-quadrature expansions of 1/r and exp(-mu r)/r on [r_min, r_max]
// (src/operators/PoissonKernel.cpp:48-86, src/operators/HelmholtzKernel.cpp:47-80).
std::vector<KernelTerm> poisson_kernel(double epsilon, double r_min, double r_max);
std::vector<KernelTerm> helmholtz_kernel(double mu, double epsilon, double r_min, double r_max);

namespace detail {

// One separable term: a 2D operator tree over (l_f, l_g) and its translation
// cache (src/trees/OperatorTree.cpp:200-238) plus band widths per depth
// (src/trees/BandWidth.h:37-61; columns T, C, B, A, max).
struct OperatorTerm {
    std::unique_ptr<MultiResolutionAnalysis<2>> mra;
    std::unique_ptr<NodeStore<2>> tree;
    std::vector<int> maxTransl;               // per depth, at cache setup
    std::vector<std::vector<int>> cache;      // cache[depth][l + maxTransl[depth]] -> node id
    std::vector<std::array<int, 5>> widths;   // rows 0..tree depth
    int bwDepth() const { return (int)widths.size() - 1; }
    int width(int depth, int c) const { return depth > bwDepth() ? -1 : widths[depth][c]; }
    int maxWidth(int depth) const { return depth > bwDepth() ? -1 : widths[depth][4]; }
    int node(int depth, int l) const { return cache[depth][l + maxTransl[depth]]; }
};

// Flat tables for one apply precision.  For depth d and term i only
// translations |dl| <= maxWidth_i(d) are stored, 4 components each.
struct BandTables {
    int M = 0, nDepth = 0, kp1 = 0, Wmax = 0;
    std::vector<int> bandMax;        // [nDepth]
    std::vector<int> widths;         // [nDepth][M][5]
    std::vector<int> bandSize;       // [nDepth][M][64]  (gt*8+ft)
    std::vector<long long> matStart; // [nDepth][M] index of matrix (dl=-w, comp 0), -1 if none
    std::vector<double> mats;        // matrices, column-major (k+1)^2 each
    std::vector<double> matNorm;     // per matrix
    std::vector<double> pairBound;   // [nDepth][2*Wmax+1] max over terms of cbrt(bs_max) x max component norm
    std::vector<double> pairBoundW;  // same over the wavelet components (C, B, A) only
    std::vector<int> bandSizeMax;    // [nDepth][M] row maximum of bandSize
    std::vector<double> termNorm;    // [matrix / 4]: max over the 4 components
};

struct DeviceBandTables; // owned by the device layer

} // namespace detail

template <int D> class ConvolutionOperator;

// Where later operator constructions on this thread build the 2D operator
// trees: Auto (the host cores), Host, Device (one launch per refinement
// level for all terms).  The trees are bit-identical either way.
enum class OperatorTarget { Auto = 0, Host = 1, Device = 2 };
void set_operator_target(OperatorTarget t);

// One operator of a batch construction: Poisson when mu <= 0, else Helmholtz(mu).
struct OperatorSpec {
    double mu = 0.0;
    double prec = 1e-6;
};

// Several operators on one MRA constructed together (no reference
// counterpart; MRCPP builds each operator in its constructor).  With
// OperatorTarget::Device the 2D trees of every term of every operator are
// formed in one device launch per refinement level; Host (or Auto) builds them
// one after the other on the host cores.  Each operator equals the one its own
// constructor builds, bit for bit; buildSeconds() of each is the batch's time.
std::vector<std::unique_ptr<ConvolutionOperator<3>>> build_operators(const MultiResolutionAnalysis<3> &mra, const std::vector<OperatorSpec> &specs,
                                                                     OperatorTarget target = OperatorTarget::Device);

// Isotropic separated 3D convolution operator (src/operators/ConvolutionOperator.h:62-84).
template <> class ConvolutionOperator<3> {
public:
    ConvolutionOperator(const MultiResolutionAnalysis<3> &mra, const std::vector<KernelTerm> &kernel, double prec);
    virtual ~ConvolutionOperator();
    ConvolutionOperator(const ConvolutionOperator &) = delete;
    ConvolutionOperator &operator=(const ConvolutionOperator &) = delete;

    const MultiResolutionAnalysis<3> &getMRA() const { return mra_; }
    int size() const { return (int)terms_.size(); }
    int getOperatorRoot() const { return mra_.getRootScale(); }
    double getBuildPrec() const { return build_prec_; }
    const std::vector<KernelTerm> &kernel() const { return kernel_; }
    // screening parameter of the kernel (Helmholtz mu; 0 for Poisson)
    double kernelMu() const { return mu_; }
    const detail::OperatorTerm &term(int i) const { return *terms_[i]; }
    double buildSeconds() const { return build_seconds_; }
    // seconds of the construction's phases: 1D kernel trees and term set-up;
    // per refinement level the packing of the device's inputs, the device
    // round trip (upload, launch, download) and the host's share (storing the
    // nodes, norms, split tests; a host build counts its whole 2D tree
    // construction here); bottom-up transforms and translation caches
    const std::array<double, 5> &buildPhases() const { return build_phases_; }
    bool builtOnDevice() const { return built_on_device_; }

    // MWOperator::calcBandWidths / getMaxBandWidth (src/operators/MWOperator.cpp:62-108)
    void calcBandWidths(double prec);
    void clearBandWidths() { band_max_.clear(); }
    int getMaxBandWidth(int depth) const;
    const std::vector<int> &bandMax() const { return band_max_; }

    // Tables for the MI355X band kernels at the given apply precision
    // (built on first use, cached per precision).
    const detail::BandTables &bandTables(double prec);
    detail::DeviceBandTables *deviceTables(double prec);

protected:
    ConvolutionOperator(const MultiResolutionAnalysis<3> &mra);
    void initialize(const std::vector<KernelTerm> &kernel, double prec);
    double mu_ = 0.0;

private:
    void prepare(const std::vector<KernelTerm> &kernel, double prec);
    void build_host();
    static void build_device(const std::vector<ConvolutionOperator<3> *> &ops);

    friend std::vector<std::unique_ptr<ConvolutionOperator<3>>> build_operators(const MultiResolutionAnalysis<3> &,
                                                                                const std::vector<OperatorSpec> &, OperatorTarget);

    MultiResolutionAnalysis<3> mra_;
    std::vector<KernelTerm> kernel_;
    std::vector<std::unique_ptr<detail::OperatorTerm>> terms_;
    std::vector<int> band_max_;
    double build_prec_ = 0.0;
    double build_seconds_ = 0.0;
    std::array<double, 5> build_phases_{};
    bool built_on_device_ = false;
    double table_prec_ = -1.0;
    std::unique_ptr<detail::BandTables> tables_;
    detail::DeviceBandTables *dev_tables_ = nullptr;
};

// PoissonOperator (src/operators/PoissonOperator.cpp:40-55)
class PoissonOperator : public ConvolutionOperator<3> {
public:
    PoissonOperator(const MultiResolutionAnalysis<3> &mra, double prec);
};

// HelmholtzOperator (src/operators/HelmholtzOperator.cpp:41-56)
class HelmholtzOperator : public ConvolutionOperator<3> {
public:
    HelmholtzOperator(const MultiResolutionAnalysis<3> &mra, double mu, double prec);
};

} // namespace mrcpp
