/*
 * mrcpp_amd.h -- C ABI of the MI355X MRCPP operator-application library.
 *
 * A flat binding of the C++ API in namespace mrcpp (mrcpp_amd/csrc), for
 * FFI callers (ctypes, cffi, cgo, JNI).  Each function names the MRCPP C++
 * call it stands for (paths relative to the MRCPP source tree).  Handles are
 * opaque; every function returning int returns 0 on success and -1 on error,
 * with mrcpp_last_error() describing the failure.
 *
 * Threads: every call that may touch the GPU holds one process-wide lock, so
 * calls from several host threads are served one after the other (results
 * equal those of the same calls made in sequence).  Settings documented as
 * "on this thread" (sweep precision, projection and operator targets) and
 * mrcpp_last_error() are per calling thread.
 */
#ifndef MRCPP_AMD_H
#define MRCPP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

const char *mrcpp_last_error(void);
/* Directory with the L_H0_k, L_G0_k, L_c_left_k, L_c_right_k tables
 * (MWFILTERS_DIR / details::find_filters, src/utils/details.cpp:53-69). */
void mrcpp_set_filter_dir(const char *dir);

/* MultiResolutionAnalysis<3>(BoundingBox<3>({lo, hi}), LegendreBasis(order), depth)
 * (src/trees/MultiResolutionAnalysis.h:51-53) */
int mrcpp_mra_create(int box_lo, int box_hi, int order, int depth, void **mra);
/* MultiResolutionAnalysis<3>(BoundingBox<3>(scale, corner, nboxes, sf), LegendreBasis(order), depth)
 * (src/trees/BoundingBox.cpp:100-106, src/trees/MultiResolutionAnalysis.h:53): a world of
 * nboxes[d] root boxes from translation corner[d] at root scale `scale`; sf NULL means 1.0. */
int mrcpp_mra_create_box(int scale, const int *corner, const int *nboxes, const double *sf, int order, int depth, void **mra);
void mrcpp_mra_destroy(void *mra);
/* root scale, scaling factor and max scale of the world box */
int mrcpp_mra_info(void *mra, int *root_scale, double *scaling_factor, int *max_scale);

/* FunctionTree<3>(mra) (src/trees/FunctionTree.h:81) */
int mrcpp_tree_create(void *mra, void **tree);
void mrcpp_tree_destroy(void *tree);
/* Gaussian sums f = sum_i c_i exp(-a_i |r - R_i|^2), centres R_i = centers[3i..3i+2]:
 * build_grid(tree, GaussExp) (src/treebuilders/grid.cpp:106-123) and
 * project(prec, tree, GaussExp) (src/treebuilders/project.cpp:85-102) */
int mrcpp_tree_build_grid_gaussians(void *tree, int n, const double *coefs, const double *alphas, const double *centers);
int mrcpp_tree_project_gaussians(void *tree, double prec, int n, const double *coefs, const double *alphas, const double *centers);
/* the same with project()'s maxIter and absPrec (src/treebuilders/project.cpp:85) */
int mrcpp_tree_project_gaussians_ex(void *tree, double prec, int n, const double *coefs, const double *alphas, const double *centers, int max_iter,
                                    int abs_prec);
long long mrcpp_tree_num_nodes(void *tree);
double mrcpp_tree_square_norm(void *tree);
/* real nodes depth first: idx[5 i] = scale, lx, ly, lz, is_end;
 * coefs[8 (k+1)^3 i ...] (may be NULL); norms[9 i] = 8 component norms, square norm */
int mrcpp_tree_export(void *tree, int *idx, double *coefs, double *norms);
/* add(prec, out, FunctionTreeVector, maxIter, absPrec) (src/treebuilders/add.cpp:43-80):
 * out = sum_i c[i] trees[i], refined from out's end nodes by the wavelet test;
 * computed on the GPU, 1 <= n <= 8, out distinct from the inputs. */
int mrcpp_add(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec);
/* multiply(prec, out, FunctionTreeVector, maxIter, absPrec) (src/treebuilders/multiply.cpp:52-100):
 * out = prod_i c[i] trees[i] with WaveletAdaptor refinement, on the GPU, 1 <= n <= 8. */
int mrcpp_multiply(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec);
/* multiply(prec, out, FunctionTreeVector, maxIter, absPrec, useMaxNorms)
 * (src/treebuilders/multiply.cpp:104-119): use_max_norms != 0 refines with
 * MultiplicationAdaptor (src/treebuilders/MultiplicationAdaptor.h:46-66) from
 * the two operands' largest norms (n must be 2); 0 is mrcpp_multiply. */
int mrcpp_multiply_ex(double prec, void *out, int n, const double *c, void **trees, int max_iter, int abs_prec, int use_max_norms);
/* FunctionTree::rescale (src/trees/FunctionTree.cpp:651-668): end nodes times c,
 * then the bottom-up transform and norms (on the GPU for trees held in HBM) */
int mrcpp_tree_rescale(void *tree, double c);
/* apply(prec, CompFunction out, oper, CompFunction inp, metric, maxIter, absPrec)
 * (src/treebuilders/apply.h:42, src/treebuilders/apply.cpp:119-139) over real
 * components: out[4] and inp[ncomp] are tree handles (an out entry may be NULL
 * when no metric entry selects it); metric: 16 doubles [input][output],
 * row-major, NULL for the identity.  Pairs writing distinct outputs run as
 * one batch. */
int mrcpp_apply_components(double prec, void **out, void *oper, void **inp, int ncomp, const double *metric, int max_iter, int abs_prec);
/* FunctionTree::saveTree / loadTree (src/trees/FunctionTree.cpp:375-435): binary
 * "<file>.tree" of every real node in this library's layout (MRCPP's file is its
 * allocator's memory image); loading needs the MRA the tree was saved with. */
int mrcpp_tree_save(void *tree, const char *file);
int mrcpp_tree_load(void *tree, const char *file);
/* FunctionTree::saveTreeTXT / loadTreeTXT (src/trees/FunctionTree.cpp:127-373):
 * MADNESS-convention text (values of every end node's children at their
 * quadrature points, 14 significant digits); loading needs the world cube
 * [-L, L], L = 2^-rootScale, and the file's order. */
int mrcpp_tree_save_txt(void *tree, const char *file);
int mrcpp_tree_load_txt(void *tree, const char *file);
/* FunctionTree::evalf, integrate (src/trees/FunctionTree.cpp:438-546); dot (src/treebuilders/multiply.cpp:286-318) */
double mrcpp_tree_evalf(void *tree, double x, double y, double z);
double mrcpp_tree_integrate(void *tree);
double mrcpp_tree_dot(void *bra, void *ket);

/* PoissonOperator(mra, prec) (kind 0) / HelmholtzOperator(mra, mu, prec) (kind 1)
 * (src/operators/PoissonOperator.h:44, src/operators/HelmholtzOperator.h:44) */
int mrcpp_operator_create(void *mra, int kind, double mu, double prec, void **oper);
void mrcpp_operator_destroy(void *oper);
int mrcpp_operator_size(void *oper);
double mrcpp_operator_build_seconds(void *oper);
/* Where later operator constructions on this thread form their 2D operator
 * trees: 0 default (the host cores), 1 the host cores, 2 the GPU (one launch
 * per refinement level for all terms); the trees are bit-identical either
 * way.  No reference counterpart (CrossCorrelationCalculator runs on OpenMP threads). */
int mrcpp_set_operator_target(int target);
/* 1 if the operator's trees were formed on the GPU */
int mrcpp_operator_built_on_device(void *oper);
/* n operators on one MRA built together: opers[i] is Poisson when mu[i] <= 0,
 * else Helmholtz(mu[i]), at prec[i]; target 2 forms the 2D trees of all their
 * terms in one GPU launch per refinement level, 0/1 builds them one by one on
 * the host cores.  Each equals its single construction bit for bit.  No
 * reference counterpart (MRCPP builds each operator in its constructor,
 * src/operators/ConvolutionOperator.cpp:78-108). */
int mrcpp_operators_create(void *mra, int n, const double *mu, const double *prec, int target, void **opers);
/* seconds of the construction phases (out[5]): kernel trees and term set-up;
 * per refinement level the packing of the GPU's inputs, the GPU round trips
 * and the host's share (a host build counts its whole 2D tree construction
 * here); bottom-up transforms and translation caches */
int mrcpp_operator_build_phases(void *oper, double *out);
int mrcpp_operator_kernel(void *oper, double *alpha, double *coef);
/* band widths at an apply precision: out[(term * max_depth + depth) * 5 + c], c = T,C,B,A,max
 * (MWOperator::calcBandWidths, src/operators/MWOperator.cpp:78-108) */
int mrcpp_operator_bandwidths(void *oper, double prec, int *out, int max_depth);
/* one operator node component (OperatorTree::getNode(depth, l), component c) */
int mrcpp_operator_node(void *oper, int term, int depth, int l, int comp, double *mat, double *norm);
/* PoissonKernel / HelmholtzKernel expansion (src/operators/PoissonKernel.cpp:48-86,
 * src/operators/HelmholtzKernel.cpp:47-80); returns the number of terms */
int mrcpp_kernel_terms(int kind, double mu, double eps, double r_min, double r_max, double *alpha, double *coef, int cap);

/* Sweep precision of later mrcpp_apply calls on this thread: 0 FP64 (default,
 * matrix cores), 1 FP32, 2 BF16x3, 3 FP64 on the vector ALUs.  No reference
 * counterpart (the reference computes in FP64 only). */
int mrcpp_set_sweep_precision(int precision);
int mrcpp_get_sweep_precision(void);
/* Whether later mrcpp_apply calls on this thread form each breadth-first
 * step's next level and its pair screen on the device before the host's
 * split (single plain trees, steps of at most 2048 nodes; the host checks the
 * device's split against its own and takes it only when they agree bit for
 * bit).  1 or 0 (default); the trees are the same either way.  No reference
 * counterpart (TreeBuilder::build, src/treebuilders/TreeBuilder.cpp:48-80,
 * splits on the host). */
int mrcpp_set_level_ahead(int on);
int mrcpp_level_ahead(void);

/* Autotune the band sum's launch shape on the caller's workload: run(ctx) is
 * timed under each candidate (layout x minimum unit groups x workgroups per
 * CU; see mwg_set_band_layout, mwg_set_band_launch) and the fastest left set.
 * best (4 doubles): layout, unit groups, workgroups per CU, seconds per run;
 * table (table_cap rows of 4, may be NULL): every candidate.  Returns the
 * number of candidates timed, or -1.  No reference counterpart. */
int mrcpp_autotune_band(void (*run)(void *), void *ctx, int repeats, double *best, double *table, int table_cap);
/* apply(prec, out, oper, inp, max_iter, abs_prec) (src/treebuilders/apply.h:41).
 * stats (may be NULL, 16 doubles): applied tuples, candidate pairs, generated
 * nodes, output nodes, seconds, levels, t_screen, t_generate, t_band,
 * t_setup, t_transform, t_host, t_cleanup, steps formed ahead on the device
 * (their nodes and pair screen before the host's split) */
int mrcpp_apply(double prec, void *out, void *oper, void *inp, int max_iter, int abs_prec, double *stats);
/* apply(prec, out, oper, inp, precTrees, max_iter, abs_prec) (src/treebuilders/apply.h:43,
 * src/treebuilders/apply.cpp:214-251): prec_trees holds n_prec tree handles
 * (n_prec 0: every node's precision factor is 1); the output tree may not be
 * one of them.  stats as mrcpp_apply. */
int mrcpp_apply_prec(double prec, void *out, void *oper, void *inp, void *const *prec_trees, int n_prec, int max_iter, int abs_prec,
                     double *stats);
/* n applications at once (no reference counterpart; same result per tree as
 * mrcpp_apply): arrays of n arguments, max_iter / abs_prec may be NULL
 * (-1 / 0), stats (may be NULL) receives 16 doubles per task as above, the
 * seconds entry being the whole batch's wall time. */
int mrcpp_apply_batch(int n, const double *prec, void *const *out, void *const *oper, void *const *inp, const int *max_iter,
                      const int *abs_prec, double *stats);

/* Refinement predictor (no reference counterpart): a 12-64-64-1 perceptron
 * on the device predicting which output nodes the adaptive apply refines.
 * Parameters are MWG_MLP_PARAMS (5057) floats; see include/mwg.h. */
int mrcpp_predictor_create(void **pred);
int mrcpp_predictor_destroy(void *pred);
int mrcpp_predictor_num_params(void);
int mrcpp_predictor_set_params(void *pred, const float *params, int n);
int mrcpp_predictor_get_params(void *pred, float *params, int n);
/* refine a node when its predicted probability exceeds threshold */
int mrcpp_predictor_set_threshold(void *pred, double threshold);
/* prob[i] for features[n][12] (host arrays) */
int mrcpp_predictor_predict(void *pred, const float *features, int n, float *prob);
/* Adam on minibatches of the weighted cross-entropy, on the device; losses
 * (may be NULL) receives the mean loss after every epoch */
int mrcpp_predictor_train(void *pred, const float *features, const float *labels, int n, int epochs, int batch, double lr,
                          double pos_weight, unsigned long long seed, float *losses);
/* mrcpp_apply_batch with the predictor growing every output grid after the
 * first step (predictor may be NULL: plain).  centers: the source centres of
 * all tasks back to back (3 doubles each), ncenters[i] of task i (both may
 * be NULL).  collect != 0 gathers training samples (features, refined) of
 * every output node beyond the first step, read with mrcpp_samples_*.
 * stats: 16 doubles per task as mrcpp_apply, entries 13-15 = nodes
 * predicted, predicted nodes removed by the wavelet test, nodes computed in
 * fallback passes. */
int mrcpp_apply_batch_predicted(int n, const double *prec, void *const *out, void *const *oper, void *const *inp, void *predictor,
                                const double *centers, const int *ncenters, int collect, double *stats);
int mrcpp_samples_count(void);
/* copies the samples of the last collecting call; returns their number, -1 if cap is too small */
int mrcpp_samples_fetch(float *features, float *labels, int cap);
/* Where later projections (mrcpp_tree_project_gaussians) run on this thread:
 * 0 the GPU when present (default), 1 the host cores, 2 the GPU.  No reference
 * counterpart (ProjectionCalculator runs on OpenMP threads, src/treebuilders/ProjectionCalculator.cpp:34-51). */
int mrcpp_set_projection_target(int target);
/* 1 if the last projection on this thread ran on the GPU */
int mrcpp_last_projection_on_device(void);
/* nodes per breadth-first level of the last apply; returns the number of levels */
int mrcpp_last_apply_levels(int *level_nodes, int cap);

#ifdef __cplusplus
}
#endif

#endif /* MRCPP_AMD_H */
