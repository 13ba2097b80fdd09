// FunctionTree<3>: a multiwavelet function on the host node store with an
// optional device-resident copy of its coefficients (src/trees/FunctionTree.h:55-183).
#pragma once

#include <complex>
#include <functional>
#include <memory>
#include <string>
#include <tuple>

#include "device.h"
#include "functions.h"
#include "operators.h"
#include "trees.h"

namespace mrcpp {

namespace device {
// Quadrature nodes and the value <-> coefficient maps of an order on the device.
struct BasisTables {
    DeviceArray<double> nodes, vcmap, cvmap;
};
const BasisTables &basis_tables(const MultiResolutionAnalysis<3> &mra);
} // namespace device

// FunctionTree<D, T> (src/trees/FunctionTree.h:55): D = 3 and real (double)
// coefficients are provided; FunctionTree<3> names FunctionTree<3, double>.
template <int D, typename T = double> class FunctionTree;

template <> class FunctionTree<3, double> {
public:
    explicit FunctionTree(const MultiResolutionAnalysis<3> &mra, const std::string &name = "nn");
    FunctionTree(const FunctionTree &) = delete;
    FunctionTree &operator=(const FunctionTree &) = delete;

    const MultiResolutionAnalysis<3> &getMRA() const { return mra_; }
    const std::string &getName() const { return name_; }
    double getSquareNorm() const { return store_->treeSquareNorm; }
    int getNNodes() const;
    int getNEndNodes() const;
    double integrate();
    double evalf(const Coord<3> &r);

    // FunctionTree::rescale (src/trees/FunctionTree.cpp:651-668): end-node
    // coefficients times c, then the bottom-up transform and the norms (on the
    // device when the coefficients live in HBM).
    void rescale(double c);

    // FunctionTree::saveTree / loadTree (src/trees/FunctionTree.cpp:375-435):
    // binary file "<file>.tree" of every real node (this library's layout; the
    // tree loading it must have the MRA it was saved with).
    void saveTree(const std::string &file);
    void loadTree(const std::string &file);
    // FunctionTree::saveTreeTXT / loadTreeTXT (src/trees/FunctionTree.cpp:127-373):
    // MADNESS-convention text, the values of every end node's children at
    // their quadrature points.  Loading needs a world cube [-L, L], L = 2^-rootScale.
    void saveTreeTXT(const std::string &file);
    void loadTreeTXT(const std::string &file);

    // host node store with coefficients brought up to date
    detail::NodeStore<3> &nodes();
    // host node store without touching coefficients (topology and norms)
    detail::NodeStore<3> &topology() { return *store_; }
    void clear();

    // device copy of the coefficients (node id u at u * 8 (k+1)^3)
    device::DeviceArray<double> &deviceCoefs() { return dev_coef_; }
    void markDeviceResident() {
        host_stale_ = true;
        mirror_.reset();
    }
    bool hostStale() const { return host_stale_; }
    void syncHost();
    // device mirror used when this tree is the input of apply
    device::TreeMirror &inputMirror();
    void invalidate() { mirror_.reset(); }

private:
    MultiResolutionAnalysis<3> mra_;
    std::string name_;
    std::unique_ptr<detail::NodeStore<3>> store_;
    device::DeviceArray<double> dev_coef_;
    bool host_stale_ = false;
    std::unique_ptr<device::TreeMirror> mirror_;
};

// build_grid(tree, GaussExp, maxIter) (src/treebuilders/grid.cpp:106-123)
// and build_grid(tree, Gaussian, maxIter) (grid.cpp:78-84, the same refinement
// for one term)
void build_grid(FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter = -1);
void build_grid(FunctionTree<3> &out, const GaussFunc<3> &inp, int maxIter = -1);
// project(prec, tree, GaussExp) (src/treebuilders/project.cpp:85-102).  Each
// breadth-first level is computed on the MI355X when one is present (the
// function values, value -> coefficient maps, compression and norms; the
// host keeps the topology and the split test), else on the host cores.
void project(double prec, FunctionTree<3> &out, const GaussExp<3> &inp, int maxIter = -1, bool absPrec = false);
// project(prec, tree, GaussFunc) (src/treebuilders/project.h:33): a one-term expansion
void project(double prec, FunctionTree<3> &out, const GaussFunc<3> &inp, int maxIter = -1, bool absPrec = false);
// project(prec, tree, std::function) (src/treebuilders/project.h:34,
// project.cpp:39-43): any function of a point, evaluated on the host cores
void project(double prec, FunctionTree<3> &out, std::function<double(const Coord<3> &r)> func, int maxIter = -1, bool absPrec = false);
enum class ProjectionTarget { Auto = 0, Host = 1, Device = 2 };
// Where later project() calls on this thread run (Auto: the GPU if present).
void set_projection_target(ProjectionTarget t);
// Whether the last project() on this thread ran on the GPU.
bool last_projection_on_device();
// dot(bra, ket) (src/treebuilders/multiply.cpp:286-318)
double dot(FunctionTree<3> &bra, FunctionTree<3> &ket);

namespace detail {
// dot on the device copies of both trees (their input mirrors): node products
// in one launch, partial sums added on the host in bra's node order.
double dot_device(FunctionTree<3> &bra, FunctionTree<3> &ket);
} // namespace detail

// (coefficient, tree) terms (src/trees/FunctionTreeVector.h:35-36) and the
// accessors of src/trees/FunctionTreeVector.h:42-104.
template <int D, typename T = double> using CoefsFunctionTree = std::tuple<T, FunctionTree<D, T> *>;
template <int D, typename T = double> using FunctionTreeVector = std::vector<CoefsFunctionTree<D, T>>;
template <int D, typename T> void clear(FunctionTreeVector<D, T> &fs, bool dealloc = false) {
    if (dealloc)
        for (auto &t : fs) delete std::get<1>(t);
    fs.clear();
}
template <int D, typename T> T get_coef(const FunctionTreeVector<D, T> &fs, int i) { return std::get<0>(fs[i]); }
template <int D, typename T> FunctionTree<D, T> &get_func(FunctionTreeVector<D, T> &fs, int i) { return *std::get<1>(fs[i]); }
template <int D, typename T> const FunctionTree<D, T> &get_func(const FunctionTreeVector<D, T> &fs, int i) { return *std::get<1>(fs[i]); }
template <int D, typename T> int get_n_nodes(const FunctionTreeVector<D, T> &fs) {
    int n = 0;
    for (const auto &t : fs)
        if (std::get<1>(t)) n += std::get<1>(t)->getNNodes();
    return n;
}
// add(prec, out, a, inp_a, b, inp_b, maxIter, absPrec) and add(prec, out,
// FunctionTreeVector, ...) (src/treebuilders/add.cpp:43-80): TreeBuilder with
// AdditionCalculator and WaveletAdaptor from out's end nodes, then the
// bottom-up transform.  Each breadth-first step runs on the MI355X (at most 8
// terms); the output's coefficients stay in HBM.
// conjugate (src/treebuilders/add.h:31-33, multiply.h:41-55) conjugates the
// operands of complex trees; real trees are their own conjugates, so it is
// accepted and has no effect.  The std::vector forms take every coefficient as 1.
void add(double prec, FunctionTree<3> &out, double a, FunctionTree<3> &inp_a, double b, FunctionTree<3> &inp_b, int maxIter = -1,
         bool absPrec = false, bool conjugate = false);
void add(double prec, FunctionTree<3> &out, const FunctionTreeVector<3> &inp, int maxIter = -1, bool absPrec = false, bool conjugate = false);
void add(double prec, FunctionTree<3> &out, const std::vector<FunctionTree<3> *> &inp, int maxIter = -1, bool absPrec = false,
         bool conjugate = false);
// multiply(prec, out, c, inp_a, inp_b, maxIter, absPrec, useMaxNorms) and the
// FunctionTreeVector form (src/treebuilders/multiply.cpp:52-119) with
// MultiplicationCalculator and WaveletAdaptor, or with useMaxNorms the
// MultiplicationAdaptor refinement from the two operands' largest norms
// (src/treebuilders/MultiplicationAdaptor.h:46-66).
void multiply(double prec, FunctionTree<3> &out, double c, FunctionTree<3> &inp_a, FunctionTree<3> &inp_b, int maxIter = -1,
              bool absPrec = false, bool useMaxNorms = false, bool conjugate = false);
void multiply(double prec, FunctionTree<3> &out, const FunctionTreeVector<3> &inp, int maxIter = -1, bool absPrec = false,
              bool useMaxNorms = false, bool conjugate = false);
void multiply(double prec, FunctionTree<3> &out, const std::vector<FunctionTree<3> *> &inp, int maxIter = -1, bool absPrec = false,
              bool useMaxNorms = false, bool conjugate = false);

// Statistics of the last apply on this thread.
struct ApplyStats {
    long long nApplied = 0;    // applied (f, ft, gt, term) tuples
    long long nPairs = 0;      // candidate (g, f) pairs after the separable bound
    long long nGenerated = 0;  // generated band nodes
    long long nOutputNodes = 0;
    std::vector<int> levelNodes;
    long long nPredicted = 0;  // output nodes created by the refinement predictor
    long long nCropped = 0;    // predicted nodes the wavelet test rejected (removed)
    long long nFallback = 0;   // nodes the predictor missed, computed in extra passes
    long long nAhead = 0;      // steps whose nodes and pair screen were formed on the device ahead of the host's split
    double seconds = 0.0;      // host wall time of apply()
    double tScreen = 0.0, tGenerate = 0.0, tBand = 0.0, tReduce = 0.0, tTransform = 0.0, tHost = 0.0;
    double tSetup = 0.0;   // tables, mirrors and output pool before the first step
    double tCleanup = 0.0; // generated band nodes dropped, statistics
};
const ApplyStats &last_apply_stats();

// Precision of the band sweeps (FP64 is the drop-in default).  FP32 and
// BF16x3 run the sweep products on the matrix cores in lower precision with
// the band sum accumulated in FP64; the screening, operator construction and
// transforms stay FP64.  FP64Vector runs the FP64 sweeps on the vector ALUs.
enum class SweepPrecision { FP64 = 0, FP32 = 1, BF16x3 = 2, FP64Vector = 3 };
// Sweep precision used by apply() on this thread.
void set_sweep_precision(SweepPrecision p);
SweepPrecision sweep_precision();
// Whether apply() on this thread forms the next breadth-first level and its
// pair screen on the device before the host's split of a single plain tree
// (the same tree either way, one host round trip fewer per step; off by
// default: at the step sizes measured the extra launches cost more than the
// round trip saves).
void set_level_ahead(bool on);
bool level_ahead();

// Launch shape of the band sum chosen by timing run() (the caller's own
// workload, e.g. one apply) under each candidate: layout (MWG_BAND_ROWS or
// MWG_BAND_TILES) x minimum groups per work unit (24, 32, 48) x workgroups
// per CU of the persistent kernels (as many as fit, or 3); run() is called
// once untimed and `repeats` times timed per candidate, the median kept.
// The fastest is left set (mwg_set_band_layout, mwg_set_band_launch) and
// returned; timings, if given, receives every candidate.  Process-wide.
struct BandTuning {
    int layout, unitGroups, wgPerCu;
    double seconds;
};
BandTuning autotune_band(const std::function<void()> &run, int repeats = 3, std::vector<BandTuning> *timings = nullptr);

// apply(prec, out, oper, inp, maxIter, absPrec) (src/treebuilders/apply.cpp:68-93),
// the band sum running on the MI355X.
void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, int maxIter = -1,
           bool absPrec = false);
// apply(prec, out, oper, inp, precTrees, maxIter, absPrec)
// (src/treebuilders/apply.h:43, src/treebuilders/apply.cpp:214-251): the
// precision of output node idx is prec / max over precTrees of the largest
// norm at or below idx (scaled to unit volume).
void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp, FunctionTreeVector<3> &precTrees,
           int maxIter = -1, bool absPrec = false);
// The same with the precision trees as plain pointers (their coefficients in
// a FunctionTreeVector are not used by the reference either).
void apply(double prec, FunctionTree<3> &out, ConvolutionOperator<3> &oper, FunctionTree<3> &inp,
           const std::vector<FunctionTree<3> *> &precTrees, int maxIter = -1, bool absPrec = false);

class RefinePredictor;
struct RefineSamples;

// One application of a batch: the arguments of apply(), and optionally a
// refinement predictor that grows the output grid after the second step (the
// wavelet test stays the judge: predicted nodes it rejects are removed,
// nodes it adds that were not predicted are computed in further passes),
// source centres for the predictor's features, and a sink for training
// samples (every output node beyond the second step: features, refined or not).
struct ApplyTask {
    double prec;
    FunctionTree<3> *out;
    ConvolutionOperator<3> *oper;
    FunctionTree<3> *inp;
    int maxIter = -1;
    bool absPrec = false;
    const RefinePredictor *predictor = nullptr;
    const std::vector<std::array<double, 3>> *centers = nullptr;
    RefineSamples *samples = nullptr;
    // precision trees (nullptr: the apply without them)
    const std::vector<FunctionTree<3> *> *precTrees = nullptr;
};
// Many applications at once (no reference counterpart): every breadth-first
// step of every unfinished tree shares one screen and one band-sum launch.
// Each output tree equals what apply() alone produces.  All trees must have
// the same Legendre order; output trees must be distinct.
void apply_batch(const std::vector<ApplyTask> &tasks);
// Per-task statistics of the last apply_batch on this thread (seconds is the
// whole batch's wall time).
const std::vector<ApplyStats> &last_batch_stats();

using ComplexDouble = std::complex<double>;
constexpr ComplexDouble defaultMetric[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};

template <int D> class CompFunction;
// CompFunction<3> (src/utils/CompFunction.h): a function of up to 4 real
// components, each a FunctionTree<3> owned by it (the reference's complex
// components are not provided).
template <> class CompFunction<3> {
public:
    explicit CompFunction(const MultiResolutionAnalysis<3> &mra, int ncomp = 1);
    CompFunction(const CompFunction &) = delete;
    CompFunction &operator=(const CompFunction &) = delete;
    const MultiResolutionAnalysis<3> &getMRA() const { return mra_; }
    int Ncomp() const { return ncomp_; }
    bool isreal() const { return true; }
    // a new empty tree for component i (Ncomp grows to at least i + 1)
    void alloc_comp(int i);
    // ncomp new empty components, the others dropped (CompFunction::paramCopy(true))
    void reset(int ncomp);
    FunctionTree<3> *CompD[4] = {nullptr, nullptr, nullptr, nullptr};

private:
    MultiResolutionAnalysis<3> mra_;
    int ncomp_ = 0;
    std::unique_ptr<FunctionTree<3>> own_[4];
};

// apply(prec, CompFunction out, oper, CompFunction inp, metric, maxIter,
// absPrec) (src/treebuilders/apply.h:42, apply.cpp:119-139): out gets inp's
// component count; for every (input, output) component pair with a nonzero
// metric entry the operator is applied and the result rescaled by the entry.
// When the selected pairs write distinct output components they run as one
// apply_batch.  The metric must be real.
void apply(double prec, CompFunction<3> &out, ConvolutionOperator<3> &oper, const CompFunction<3> &inp,
           const ComplexDouble (*metric)[4] = defaultMetric, int maxIter = -1, bool absPrec = false);
// apply(prec, CompFunction out, oper, CompFunction inp, precTrees, metric,
// maxIter, absPrec) (src/treebuilders/apply.h:44, apply.cpp:252-270): as
// above with precTrees[i] the precision trees of input component i; a null
// metric is the identity.
void apply(double prec, CompFunction<3> &out, ConvolutionOperator<3> &oper, CompFunction<3> &inp, FunctionTreeVector<3> *precTrees,
           const ComplexDouble (*metric)[4] = nullptr, int maxIter = -1, bool absPrec = false);
// The same over caller-owned trees (out[i], inp[i] for i < ncomp); the
// output components a nonzero metric entry selects must be present.
void apply_components(double prec, FunctionTree<3> *const *out, ConvolutionOperator<3> &oper, FunctionTree<3> *const *inp, int ncomp,
                      const ComplexDouble (*metric)[4], int maxIter = -1, bool absPrec = false);

} // namespace mrcpp
