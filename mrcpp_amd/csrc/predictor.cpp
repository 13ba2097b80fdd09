#include "predictor.h"

#include <algorithm>
#include <cmath>

#include "device.h"

namespace mrcpp {

namespace {

// SplitMix64 stream (parameter initialisation and epoch shuffles).
struct Mix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = std::max(uniform(), 1e-300), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

float lg(double x) { return (float)(std::log10(std::max(x, 1e-300)) / 10.0); }

float *device_floats(size_t n) {
    void *p = nullptr;
    device::check(mwg_malloc(&p, std::max<size_t>(n, 1) * sizeof(float)), "predictor allocation");
    return static_cast<float *>(p);
}

} // namespace

void predictor_features(const PredictorContext &c, int u, float *out) {
    const detail::NodeStore<3> &g = *c.out;
    const auto &box = c.mra->getWorldBox();
    const int n = g.scale(u);
    const auto &l = g.transl(u);
    out[0] = (float)g.depth(u) / 10.0f;
    // distance to the nearest source centre, in node widths
    out[1] = 0.0f;
    if (c.centers && !c.centers->empty()) {
        const double w = box.getScalingFactor(0) * std::ldexp(1.0, -n);
        double best = 1e300;
        for (const auto &r : *c.centers) {
            double d2 = 0.0;
            for (int d = 0; d < 3; d++) {
                const double x = box.getScalingFactor(d) * std::ldexp(l[d] + 0.5, -n) - r[d];
                d2 += x * x;
            }
            best = std::min(best, d2);
        }
        out[1] = (float)(std::log2(std::max(std::sqrt(best) / w, 1.0 / 16.0)) / 10.0);
    }
    // nearest existing input node at or above (n, l)
    int nn = n, gap = 0, v = -1;
    std::array<int, 3> ll = l;
    for (;;) {
        v = c.inp->findReal(nn, ll);
        if (v >= 0 || nn <= c.inp->rootScale()) break;
        nn--;
        for (int d = 0; d < 3; d++) ll[d] >>= 1;
        gap++;
    }
    out[2] = v >= 0 ? lg(c.inp->scalingNorm(v)) : -3.0f;
    out[3] = (v >= 0 && gap == 0) ? lg(c.inp->waveletNorm(v)) : -3.0f;
    out[4] = (float)gap / 10.0f;
    // nearest output ancestor computed in an exact step
    int a = u, up = 0;
    while (a >= 0 && !(a < (int)c.firstStep.size() && c.firstStep[a])) {
        a = g.parent(a);
        up++;
    }
    out[5] = a >= 0 ? lg(c.firstWavelet[a]) : -3.0f;
    out[6] = (float)up / 10.0f;
    out[7] = lg(std::max(2e-15, c.prec * std::sqrt(std::max(c.treeNorm, 0.0)) * std::pow(2.0, -0.5 * (n + 1))));
    out[8] = lg(c.prec);
    out[9] = (float)c.mu;
    // largest operator pair bound at this depth
    const detail::BandTables &bt = *c.bands;
    const int depth = g.depth(u);
    double pb = 0.0;
    if (depth >= 0 && depth < bt.nDepth)
        for (int k = 0; k < 2 * bt.Wmax + 1; k++) pb = std::max(pb, bt.pairBound[(size_t)depth * (2 * bt.Wmax + 1) + k]);
    out[10] = lg(pb);
    out[11] = lg(c.inpNorm);
}

RefinePredictor::RefinePredictor() {
    device::ensure_init();
    params_ = device_floats(MWG_MLP_PARAMS);
    m_ = device_floats(MWG_MLP_PARAMS);
    v_ = device_floats(MWG_MLP_PARAMS);
    grad_ = device_floats(MWG_MLP_PARAMS + 1);
    // He initialisation, zero biases
    std::vector<float> p(MWG_MLP_PARAMS, 0.0f);
    Mix64 r{0x5EED5EEDull};
    const int F = MWG_MLP_FEATURES, H = MWG_MLP_HIDDEN;
    for (int i = 0; i < F * H; i++) p[i] = (float)(r.normal() * std::sqrt(2.0 / F));
    for (int i = 0; i < H * H; i++) p[F * H + H + i] = (float)(r.normal() * std::sqrt(2.0 / H));
    for (int i = 0; i < H; i++) p[F * H + H + H * H + H + i] = (float)(r.normal() * std::sqrt(1.0 / H));
    setParams(p);
}

RefinePredictor::~RefinePredictor() {
    mwg_free(params_);
    mwg_free(m_);
    mwg_free(v_);
    mwg_free(grad_);
}

void RefinePredictor::setParams(const std::vector<float> &p) {
    if (p.size() != (size_t)MWG_MLP_PARAMS) throw Error("predictor: wrong parameter count");
    device::check(mwg_upload(params_, p.data(), p.size() * sizeof(float)), "predictor upload");
    device::check(mwg_zero(m_, MWG_MLP_PARAMS * sizeof(float)), "predictor reset");
    device::check(mwg_zero(v_, MWG_MLP_PARAMS * sizeof(float)), "predictor reset");
    step_ = 0;
}

std::vector<float> RefinePredictor::params() const {
    std::vector<float> p(MWG_MLP_PARAMS);
    device::check(mwg_download(p.data(), params_, p.size() * sizeof(float)), "predictor download");
    return p;
}

std::vector<float> RefinePredictor::predict(const float *features, int n) const {
    std::vector<float> prob(std::max(n, 0));
    if (n <= 0) return prob;
    float *x = device_floats((size_t)n * MWG_MLP_FEATURES), *pr = device_floats(n);
    device::check(mwg_upload(x, features, (size_t)n * MWG_MLP_FEATURES * sizeof(float)), "predictor upload");
    device::check(mwg_mlp_forward(params_, x, n, pr), "predictor forward");
    device::check(mwg_download(prob.data(), pr, (size_t)n * sizeof(float)), "predictor download");
    mwg_free(x);
    mwg_free(pr);
    return prob;
}

std::vector<float> RefinePredictor::train(const float *features, const float *labels, int n, int epochs, int batch, float lr, float posWeight,
                                          uint64_t seed) {
    std::vector<float> losses;
    if (n <= 0 || epochs <= 0) return losses;
    batch = std::max(1, std::min(batch, n));
    const int F = MWG_MLP_FEATURES;
    float *x = device_floats((size_t)n * F), *y = device_floats(n);
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    std::vector<float> px((size_t)n * F), py(n);
    Mix64 r{seed};
    for (int e = 0; e < epochs; e++) {
        for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int)(r.next() % (uint64_t)(i + 1))]);
        for (int i = 0; i < n; i++) {
            std::copy(features + (size_t)perm[i] * F, features + (size_t)perm[i] * F + F, px.begin() + (size_t)i * F);
            py[i] = labels[perm[i]];
        }
        device::check(mwg_upload(x, px.data(), px.size() * sizeof(float)), "predictor upload");
        device::check(mwg_upload(y, py.data(), py.size() * sizeof(float)), "predictor upload");
        for (int off = 0; off < n; off += batch) {
            const int b = std::min(batch, n - off);
            device::check(mwg_mlp_gradient(params_, x + (size_t)off * F, y + off, b, posWeight, grad_), "predictor gradient");
            device::check(mwg_mlp_adam(params_, grad_, m_, v_, ++step_, lr, 0.9f, 0.999f, 1e-8f), "predictor step");
        }
        // loss of the whole set after the epoch
        device::check(mwg_mlp_gradient(params_, x, y, n, posWeight, grad_), "predictor loss");
        float loss = 0.0f;
        device::check(mwg_download(&loss, grad_ + MWG_MLP_PARAMS, sizeof(float)), "predictor download");
        losses.push_back(loss);
    }
    mwg_free(x);
    mwg_free(y);
    return losses;
}

} // namespace mrcpp
