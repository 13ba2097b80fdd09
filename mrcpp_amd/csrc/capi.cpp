// include/mrcpp_amd.h over the C++ API.
#include "../../include/mrcpp_amd.h"

#include <cstring>
#include <string>

#include "function_tree.h"
#include "predictor.h"

using namespace mrcpp;

namespace {
thread_local std::string g_error;

// Every entry point runs under the process-wide device lock: two host threads
// calling in at once (ctypes releases the GIL) are served one after the other.
template <typename F> int guarded(F &&f) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    try {
        f();
        return 0;
    } catch (std::exception &e) {
        g_error = e.what();
        return -1;
    }
}
} // namespace

extern "C" {

const char *mrcpp_last_error(void) { return g_error.c_str(); }

void mrcpp_set_filter_dir(const char *dir) { set_filter_dir(dir ? dir : ""); }

int mrcpp_mra_create(int box_lo, int box_hi, int order, int depth, void **mra) {
    return guarded([&] { *mra = new MultiResolutionAnalysis<3>({box_lo, box_hi}, order, depth); });
}

int mrcpp_mra_create_box(int scale, const int *corner, const int *nboxes, const double *sf, int order, int depth, void **mra) {
    return guarded([&] {
        if (!corner || !nboxes) throw Error("null box argument");
        std::array<int, 3> c, nb;
        std::array<double, 3> s;
        for (int d = 0; d < 3; d++) {
            if (nboxes[d] < 1) throw Error("Invalid number of root boxes");
            c[d] = corner[d];
            nb[d] = nboxes[d];
            s[d] = sf ? sf[d] : 1.0;
            if (!(s[d] > 0.0)) throw Error("Invalid scaling factor");
        }
        *mra = new MultiResolutionAnalysis<3>(BoundingBox<3>(scale, c, nb, s), LegendreBasis(order), depth);
    });
}

void mrcpp_mra_destroy(void *mra) { delete static_cast<MultiResolutionAnalysis<3> *>(mra); }

int mrcpp_mra_info(void *mra, int *root_scale, double *scaling_factor, int *max_scale) {
    return guarded([&] {
        auto *m = static_cast<MultiResolutionAnalysis<3> *>(mra);
        *root_scale = m->getRootScale();
        *scaling_factor = m->getWorldBox().getScalingFactor(0);
        *max_scale = m->getMaxScale();
    });
}

int mrcpp_tree_create(void *mra, void **tree) {
    return guarded([&] { *tree = new FunctionTree<3>(*static_cast<MultiResolutionAnalysis<3> *>(mra)); });
}

void mrcpp_tree_destroy(void *tree) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    delete static_cast<FunctionTree<3> *>(tree);
}

static GaussExp<3> gauss_sum(int n, const double *coefs, const double *alphas, const double *centers) {
    GaussExp<3> ge;
    for (int i = 0; i < n; i++) ge.append(GaussFunc<3>(alphas[i], coefs[i], {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]}));
    return ge;
}

int mrcpp_tree_build_grid_gaussians(void *tree, int n, const double *coefs, const double *alphas, const double *centers) {
    return guarded([&] { build_grid(*static_cast<FunctionTree<3> *>(tree), gauss_sum(n, coefs, alphas, centers)); });
}

int mrcpp_tree_project_gaussians(void *tree, double prec, int n, const double *coefs, const double *alphas, const double *centers) {
    return guarded([&] { project(prec, *static_cast<FunctionTree<3> *>(tree), gauss_sum(n, coefs, alphas, centers)); });
}

int mrcpp_tree_project_gaussians_ex(void *tree, double prec, int n, const double *coefs, const double *alphas, const double *centers, int max_iter,
                                    int abs_prec) {
    return guarded([&] { project(prec, *static_cast<FunctionTree<3> *>(tree), gauss_sum(n, coefs, alphas, centers), max_iter, abs_prec != 0); });
}

long long mrcpp_tree_num_nodes(void *tree) { return static_cast<FunctionTree<3> *>(tree)->getNNodes(); }

double mrcpp_tree_square_norm(void *tree) { return static_cast<FunctionTree<3> *>(tree)->getSquareNorm(); }

int mrcpp_tree_export(void *tree, int *idx, double *coefs, double *norms) {
    int n = -1;
    int rc = guarded([&] {
        auto &t = *static_cast<FunctionTree<3> *>(tree);
        auto &s = coefs ? t.nodes() : t.topology();
        auto all = detail::real_nodes(s);
        const int nc = s.ncoef();
        for (size_t i = 0; i < all.size(); i++) {
            const int u = all[i];
            idx[5 * i] = s.scale(u);
            for (int d = 0; d < 3; d++) idx[5 * i + 1 + d] = s.transl(u)[d];
            idx[5 * i + 4] = s.isEnd(u) ? 1 : 0;
            if (coefs) std::memcpy(coefs + (size_t)i * nc, s.coefs(u), sizeof(double) * nc);
            if (norms) {
                for (int c = 0; c < 8; c++) norms[9 * i + c] = s.compNorm(u, c);
                norms[9 * i + 8] = s.squareNorm(u);
            }
        }
        n = (int)all.size();
    });
    return rc < 0 ? -1 : n;
}

namespace {
FunctionTreeVector<3> tree_vector(int n, const double *c, void **trees) {
    if (n < 0 || (n > 0 && (!c || !trees))) throw Error("bad input tree list");
    FunctionTreeVector<3> v;
    for (int i = 0; i < n; i++) v.emplace_back(c[i], static_cast<FunctionTree<3> *>(trees[i]));
    return v;
}
} // namespace

int mrcpp_add(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec) {
    return guarded([&] { add(prec, *static_cast<FunctionTree<3> *>(out), tree_vector(n, c, trees), max_iter, abs_prec != 0); });
}

int mrcpp_multiply(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec) {
    return guarded([&] { multiply(prec, *static_cast<FunctionTree<3> *>(out), tree_vector(n, c, trees), max_iter, abs_prec != 0); });
}

int mrcpp_multiply_ex(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec, int use_max_norms) {
    return guarded([&] {
        multiply(prec, *static_cast<FunctionTree<3> *>(out), tree_vector(n, c, trees), max_iter, abs_prec != 0, use_max_norms != 0);
    });
}

int mrcpp_tree_rescale(void *tree, double c) {
    return guarded([&] { static_cast<FunctionTree<3> *>(tree)->rescale(c); });
}

int mrcpp_apply_components(double prec, void **out, void *oper, void **inp, int ncomp, const double *metric, int max_iter, int abs_prec) {
    return guarded([&] {
        if (ncomp < 0 || ncomp > 4 || !out || !inp) throw Error("apply: bad component arrays");
        ComplexDouble m[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) m[i][j] = metric ? ComplexDouble(metric[4 * i + j], 0.0) : defaultMetric[i][j];
        FunctionTree<3> *o[4] = {}, *in[4] = {};
        for (int i = 0; i < 4; i++) o[i] = static_cast<FunctionTree<3> *>(out[i]);
        for (int i = 0; i < ncomp; i++) in[i] = static_cast<FunctionTree<3> *>(inp[i]);
        apply_components(prec, o, *static_cast<ConvolutionOperator<3> *>(oper), in, ncomp, m, max_iter, abs_prec != 0);
    });
}

int mrcpp_tree_save(void *tree, const char *file) {
    return guarded([&] { static_cast<FunctionTree<3> *>(tree)->saveTree(file); });
}

int mrcpp_tree_load(void *tree, const char *file) {
    return guarded([&] { static_cast<FunctionTree<3> *>(tree)->loadTree(file); });
}

int mrcpp_tree_save_txt(void *tree, const char *file) {
    return guarded([&] { static_cast<FunctionTree<3> *>(tree)->saveTreeTXT(file); });
}

int mrcpp_tree_load_txt(void *tree, const char *file) {
    return guarded([&] { static_cast<FunctionTree<3> *>(tree)->loadTreeTXT(file); });
}

double mrcpp_tree_evalf(void *tree, double x, double y, double z) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    return static_cast<FunctionTree<3> *>(tree)->evalf({x, y, z});
}

double mrcpp_tree_integrate(void *tree) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    return static_cast<FunctionTree<3> *>(tree)->integrate();
}

double mrcpp_tree_dot(void *bra, void *ket) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    return dot(*static_cast<FunctionTree<3> *>(bra), *static_cast<FunctionTree<3> *>(ket));
}

int mrcpp_operator_create(void *mra, int kind, double mu, double prec, void **oper) {
    return guarded([&] {
        auto &m = *static_cast<MultiResolutionAnalysis<3> *>(mra);
        if (kind == 0)
            *oper = static_cast<ConvolutionOperator<3> *>(new PoissonOperator(m, prec));
        else if (kind == 1)
            *oper = static_cast<ConvolutionOperator<3> *>(new HelmholtzOperator(m, mu, prec));
        else
            throw Error("unknown operator kind");
    });
}

void mrcpp_operator_destroy(void *oper) {
    std::lock_guard<std::recursive_mutex> lock(device::api_mutex());
    delete static_cast<ConvolutionOperator<3> *>(oper);
}

int mrcpp_operator_size(void *oper) { return static_cast<ConvolutionOperator<3> *>(oper)->size(); }

double mrcpp_operator_build_seconds(void *oper) { return static_cast<ConvolutionOperator<3> *>(oper)->buildSeconds(); }

int mrcpp_operator_built_on_device(void *oper) { return static_cast<ConvolutionOperator<3> *>(oper)->builtOnDevice() ? 1 : 0; }

int mrcpp_operators_create(void *mra, int n, const double *mu, const double *prec, int target, void **opers) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!mu || !prec || !opers))) throw Error("operators: bad arguments");
        if (target < 0 || target > 2) throw Error("operator target: 0 auto, 1 host, 2 device");
        std::vector<OperatorSpec> specs(n);
        for (int i = 0; i < n; i++) specs[i] = {mu[i], prec[i]};
        auto ops = build_operators(*static_cast<MultiResolutionAnalysis<3> *>(mra), specs, static_cast<OperatorTarget>(target));
        for (int i = 0; i < n; i++) opers[i] = ops[i].release();
    });
}

int mrcpp_operator_build_phases(void *oper, double *out) {
    return guarded([&] {
        const auto &ph = static_cast<ConvolutionOperator<3> *>(oper)->buildPhases();
        for (int i = 0; i < 5; i++) out[i] = ph[i];
    });
}

int mrcpp_set_operator_target(int target) {
    return guarded([&] {
        if (target < 0 || target > 2) throw Error("operator target: 0 auto, 1 host, 2 device");
        set_operator_target(static_cast<OperatorTarget>(target));
    });
}

int mrcpp_operator_kernel(void *oper, double *alpha, double *coef) {
    return guarded([&] {
        const auto &k = static_cast<ConvolutionOperator<3> *>(oper)->kernel();
        for (size_t i = 0; i < k.size(); i++) {
            alpha[i] = k[i].alpha;
            coef[i] = k[i].coef;
        }
    });
}

int mrcpp_operator_bandwidths(void *oper, double prec, int *out, int max_depth) {
    int n = -1;
    int rc = guarded([&] {
        auto &o = *static_cast<ConvolutionOperator<3> *>(oper);
        o.calcBandWidths(prec);
        for (int i = 0; i < o.size(); i++)
            for (int d = 0; d < max_depth; d++)
                for (int c = 0; c < 5; c++) out[((size_t)i * max_depth + d) * 5 + c] = (c < 4) ? o.term(i).width(d, c) : o.term(i).maxWidth(d);
        n = (int)o.bandMax().size();
    });
    return rc < 0 ? -1 : n;
}

int mrcpp_operator_node(void *oper, int term, int depth, int l, int comp, double *mat, double *norm) {
    return guarded([&] {
        const auto &t = static_cast<ConvolutionOperator<3> *>(oper)->term(term);
        if (depth < 0 || depth >= (int)t.cache.size() || std::abs(l) > t.maxTransl[depth]) throw Error("no such operator node");
        const int u = t.node(depth, l);
        const int n2 = t.tree->kp1_d();
        std::memcpy(mat, t.tree->coefs(u) + (size_t)comp * n2, sizeof(double) * n2);
        *norm = t.tree->compNorm(u, comp);
    });
}

int mrcpp_kernel_terms(int kind, double mu, double eps, double r_min, double r_max, double *alpha, double *coef, int cap) {
    int n = -1;
    int rc = guarded([&] {
        auto k = (kind == 0) ? poisson_kernel(eps, r_min, r_max) : helmholtz_kernel(mu, eps, r_min, r_max);
        for (size_t i = 0; i < k.size() && (int)i < cap; i++) {
            alpha[i] = k[i].alpha;
            coef[i] = k[i].coef;
        }
        n = (int)k.size();
    });
    return rc < 0 ? -1 : n;
}

int mrcpp_set_sweep_precision(int precision) {
    return guarded([&] {
        if (precision < 0 || precision > 3) throw Error("sweep precision must be 0 (FP64), 1 (FP32), 2 (BF16x3) or 3 (FP64 vector)");
        set_sweep_precision(static_cast<SweepPrecision>(precision));
    });
}

int mrcpp_get_sweep_precision(void) { return (int)sweep_precision(); }

int mrcpp_set_level_ahead(int on) {
    return guarded([&] { set_level_ahead(on != 0); });
}

int mrcpp_level_ahead(void) { return level_ahead() ? 1 : 0; }

int mrcpp_autotune_band(void (*run)(void *), void *ctx, int repeats, double *best, double *table, int table_cap) {
    int n = 0;
    const int rc = guarded([&] {
        if (!run || !best) throw Error("autotune: null argument");
        std::vector<BandTuning> all;
        const BandTuning b = autotune_band([&] { run(ctx); }, repeats, &all);
        best[0] = b.layout;
        best[1] = b.unitGroups;
        best[2] = b.wgPerCu;
        best[3] = b.seconds;
        for (const BandTuning &t : all) {
            if (table && n < table_cap) {
                table[4 * n] = t.layout;
                table[4 * n + 1] = t.unitGroups;
                table[4 * n + 2] = t.wgPerCu;
                table[4 * n + 3] = t.seconds;
            }
            n++;
        }
    });
    return rc < 0 ? -1 : n;
}

namespace {
void pack_stats(const ApplyStats &s, double *out) {
    double v[16] = {(double)s.nApplied, (double)s.nPairs, (double)s.nGenerated, (double)s.nOutputNodes, s.seconds,
                    (double)s.levelNodes.size(), s.tScreen, s.tGenerate, s.tBand, s.tSetup, s.tTransform, s.tHost, s.tCleanup, (double)s.nAhead, 0, 0};
    std::memcpy(out, v, sizeof(v));
}
} // namespace

int mrcpp_apply_batch(int n, const double *prec, void *const *out, void *const *oper, void *const *inp, const int *max_iter, const int *abs_prec,
                      double *stats) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!prec || !out || !oper || !inp))) throw Error("apply_batch: null argument");
        std::vector<ApplyTask> tasks(n);
        for (int i = 0; i < n; i++)
            tasks[i] = ApplyTask{prec[i], static_cast<FunctionTree<3> *>(out[i]), static_cast<ConvolutionOperator<3> *>(oper[i]),
                                 static_cast<FunctionTree<3> *>(inp[i]), max_iter ? max_iter[i] : -1, abs_prec ? abs_prec[i] != 0 : false};
        apply_batch(tasks);
        if (stats)
            for (int i = 0; i < n; i++) pack_stats(last_batch_stats()[i], stats + 16 * (size_t)i);
    });
}

namespace {
thread_local RefineSamples g_samples;

std::vector<std::vector<std::array<double, 3>>> unpack_centers(int n, const double *centers, const int *ncenters) {
    std::vector<std::vector<std::array<double, 3>>> out(n);
    if (!centers || !ncenters) return out;
    size_t at = 0;
    for (int i = 0; i < n; i++) {
        if (ncenters[i] < 0) throw Error("negative centre count");
        for (int c = 0; c < ncenters[i]; c++, at++) out[i].push_back({centers[3 * at], centers[3 * at + 1], centers[3 * at + 2]});
    }
    return out;
}
} // namespace

int mrcpp_predictor_create(void **pred) {
    return guarded([&] {
        if (!pred) throw Error("null argument");
        *pred = new RefinePredictor();
    });
}

int mrcpp_predictor_destroy(void *pred) {
    return guarded([&] { delete static_cast<RefinePredictor *>(pred); });
}

int mrcpp_predictor_num_params(void) { return MWG_MLP_PARAMS; }

int mrcpp_predictor_set_params(void *pred, const float *params, int n) {
    return guarded([&] {
        if (!pred || !params) throw Error("null argument");
        static_cast<RefinePredictor *>(pred)->setParams(std::vector<float>(params, params + n));
    });
}

int mrcpp_predictor_get_params(void *pred, float *params, int n) {
    return guarded([&] {
        if (!pred || !params || n != MWG_MLP_PARAMS) throw Error("bad parameter buffer");
        const auto p = static_cast<RefinePredictor *>(pred)->params();
        std::memcpy(params, p.data(), sizeof(float) * p.size());
    });
}

int mrcpp_predictor_set_threshold(void *pred, double threshold) {
    return guarded([&] {
        if (!pred) throw Error("null argument");
        static_cast<RefinePredictor *>(pred)->threshold = (float)threshold;
    });
}

int mrcpp_predictor_predict(void *pred, const float *features, int n, float *prob) {
    return guarded([&] {
        if (!pred || (n > 0 && (!features || !prob))) throw Error("null argument");
        const auto p = static_cast<RefinePredictor *>(pred)->predict(features, n);
        if (n > 0) std::memcpy(prob, p.data(), sizeof(float) * n);
    });
}

int mrcpp_predictor_train(void *pred, const float *features, const float *labels, int n, int epochs, int batch, double lr, double pos_weight,
                          unsigned long long seed, float *losses) {
    return guarded([&] {
        if (!pred || (n > 0 && (!features || !labels))) throw Error("null argument");
        const auto l = static_cast<RefinePredictor *>(pred)->train(features, labels, n, epochs, batch, (float)lr, (float)pos_weight, seed);
        if (losses) std::memcpy(losses, l.data(), sizeof(float) * l.size());
    });
}

int mrcpp_apply_batch_predicted(int n, const double *prec, void *const *out, void *const *oper, void *const *inp, void *predictor,
                                const double *centers, const int *ncenters, int collect, double *stats) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!prec || !out || !oper || !inp))) throw Error("apply_batch: null argument");
        const auto cs = unpack_centers(n, centers, ncenters);
        g_samples = RefineSamples{};
        std::vector<ApplyTask> tasks(n);
        for (int i = 0; i < n; i++) {
            tasks[i] = ApplyTask{prec[i], static_cast<FunctionTree<3> *>(out[i]), static_cast<ConvolutionOperator<3> *>(oper[i]),
                                 static_cast<FunctionTree<3> *>(inp[i])};
            tasks[i].predictor = static_cast<const RefinePredictor *>(predictor);
            tasks[i].centers = &cs[i];
            if (collect) tasks[i].samples = &g_samples;
        }
        apply_batch(tasks);
        if (stats)
            for (int i = 0; i < n; i++) {
                const ApplyStats &st = last_batch_stats()[i];
                pack_stats(st, stats + 16 * (size_t)i);
                stats[16 * (size_t)i + 13] = (double)st.nPredicted;
                stats[16 * (size_t)i + 14] = (double)st.nCropped;
                stats[16 * (size_t)i + 15] = (double)st.nFallback;
            }
    });
}

int mrcpp_samples_count(void) { return (int)g_samples.labels.size(); }

int mrcpp_samples_fetch(float *features, float *labels, int cap) {
    const int n = (int)g_samples.labels.size();
    if (!features || !labels || cap < n) return -1;
    std::memcpy(features, g_samples.features.data(), sizeof(float) * g_samples.features.size());
    std::memcpy(labels, g_samples.labels.data(), sizeof(float) * n);
    return n;
}

int mrcpp_apply(double prec, void *out, void *oper, void *inp, int max_iter, int abs_prec, double *stats) {
    return guarded([&] {
        apply(prec, *static_cast<FunctionTree<3> *>(out), *static_cast<ConvolutionOperator<3> *>(oper), *static_cast<FunctionTree<3> *>(inp),
              max_iter, abs_prec != 0);
        if (stats) pack_stats(last_apply_stats(), stats);
    });
}

int mrcpp_apply_prec(double prec, void *out, void *oper, void *inp, void *const *prec_trees, int n_prec, int max_iter, int abs_prec, double *stats) {
    return guarded([&] {
        if (n_prec < 0 || (n_prec > 0 && !prec_trees)) throw Error("apply: bad precision tree list");
        std::vector<FunctionTree<3> *> pt;
        for (int i = 0; i < n_prec; i++) pt.push_back(static_cast<FunctionTree<3> *>(prec_trees[i]));
        apply(prec, *static_cast<FunctionTree<3> *>(out), *static_cast<ConvolutionOperator<3> *>(oper), *static_cast<FunctionTree<3> *>(inp), pt,
              max_iter, abs_prec != 0);
        if (stats) pack_stats(last_apply_stats(), stats);
    });
}

int mrcpp_set_projection_target(int target) {
    return guarded([&] {
        if (target < 0 || target > 2) throw Error("projection target: 0 auto, 1 host, 2 device");
        set_projection_target(static_cast<ProjectionTarget>(target));
    });
}

int mrcpp_last_projection_on_device(void) { return last_projection_on_device() ? 1 : 0; }

int mrcpp_last_apply_levels(int *level_nodes, int cap) {
    const ApplyStats &s = last_apply_stats();
    for (size_t i = 0; i < s.levelNodes.size() && (int)i < cap; i++) level_nodes[i] = s.levelNodes[i];
    return (int)s.levelNodes.size();
}

} // extern "C"
