// Refinement predictor for apply(): a small perceptron that predicts, per
// output node, whether the adaptive apply refines it (splits it into
// children), so the output grid can be grown before the band sums run.
// Inference and training run on the device (mwg_mlp_* in include/mwg.h).
// No reference counterpart: MRCPP refines level by level only.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "function_tree.h"

namespace mrcpp {

constexpr int kPredictorFeatures = MWG_MLP_FEATURES;

// What the features of a node see: the input tree, the operator at the apply
// precision, the source centres (optional), and the output after the exact
// breadth-first steps (their nodes' wavelet norms and the approximate tree norm).
struct PredictorContext {
    const detail::NodeStore<3> *inp = nullptr;
    const detail::NodeStore<3> *out = nullptr;
    const detail::BandTables *bands = nullptr;
    const MultiResolutionAnalysis<3> *mra = nullptr;
    double prec = 0.0, mu = 0.0;
    double inpNorm = 0.0;      // square norm of the input tree
    double treeNorm = 0.0;     // approximate output square norm after the exact steps
    const std::vector<std::array<double, 3>> *centers = nullptr;
    std::vector<char> firstStep; // output node computed in an exact step
    std::vector<double> firstWavelet; // its wavelet norm (by output node id)
};

// The 12 features of output node u (computed or not):
//  0 depth below the root scale
//  1 log2 of the distance from the node centre to the nearest source centre in node widths (0 without centres)
//  2 log10 of the input scaling norm at the node, or at its nearest existing input ancestor
//  3 log10 of the input wavelet norm at the node (-30 when the node is not in the input)
//  4 levels between the node and that input ancestor
//  5 log10 of the wavelet norm of the nearest computed output ancestor (the parent)
//  6 levels between the node and that ancestor
//  7 log10 of the wavelet split threshold at the node's scale (approximate tree norm after the exact steps)
//  8 log10 of the apply precision
//  9 Helmholtz mu (0 for Poisson)
// 10 log10 of the largest operator pair bound at the node's depth
// 11 log10 of the input square norm
void predictor_features(const PredictorContext &ctx, int u, float *out);

class RefinePredictor {
public:
    RefinePredictor();
    ~RefinePredictor();
    RefinePredictor(const RefinePredictor &) = delete;
    RefinePredictor &operator=(const RefinePredictor &) = delete;

    void setParams(const std::vector<float> &p);
    std::vector<float> params() const;
    // Probability of refinement for n samples of 12 features (host arrays).
    std::vector<float> predict(const float *features, int n) const;
    // Adam on minibatches of the weighted cross-entropy; returns the mean
    // loss of every epoch.
    std::vector<float> train(const float *features, const float *labels, int n, int epochs, int batch, float lr, float posWeight,
                             uint64_t seed);

    float threshold = 0.5f; // refine when p > threshold

private:
    float *params_ = nullptr, *m_ = nullptr, *v_ = nullptr, *grad_ = nullptr;
    int step_ = 0;
};

// Training samples of a finished apply: features of every output node and
// whether the adaptive apply refined it.
struct RefineSamples {
    std::vector<float> features;
    std::vector<float> labels;
};

} // namespace mrcpp
