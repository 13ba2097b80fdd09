/*
 * mwg.h -- C ABI of the MI355X (gfx950) band-sum and multiwavelet-transform
 * kernels behind MRCPP's operator application.
 *
 * This is the thin layer the host C++ (mrcpp_amd/csrc) calls in place of
 * MRCPP's OpenMP node loop.  Every entry point replaces one piece of the
 * reference's apply path; the citation above each declaration names it
 * (paths relative to the MRCPP source tree).  Plain C types only: device
 * pointers are void* / typed pointers obtained from mwg_malloc, sizes are
 * counts.  Every call returns MWG_OK or an error code; mwg_last_error()
 * describes the last failure of the calling thread.  Work is enqueued on the
 * library's stream; calls that return host data synchronise it.
 *
 * Threads: the device layer's scratch buffers, staging ring and stream are
 * per process and unlocked; callers serialise (the mrcpp_* ABI of
 * mrcpp_amd.h and the C++ apply() hold a process-wide lock for that).
 */
#ifndef MWG_H
#define MWG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MWG_OK = 0,
    MWG_ERR_ARG = 1,       /* invalid argument / shape */
    MWG_ERR_DEVICE = 2,    /* HIP runtime failure */
    MWG_ERR_NOMEM = 3,     /* device allocation failed */
    MWG_ERR_UNSUPPORTED = 4 /* order or depth outside the supported range */
} mwg_status;

/* Device selection and stream set-up.  Replaces nothing in MRCPP (the
 * reference has no device); must precede every other call. */
mwg_status mwg_init(int device);
const char *mwg_last_error(void);
int mwg_device_count(void);

/* Device memory helpers (bytes).  Freed blocks are kept in a per-size pool
 * and reused by later allocations; mwg_release_pool returns them.  mwg_free
 * does not wait: the library's work runs in order on the null stream, so
 * work enqueued after a free (the only work that can get the block again)
 * runs after the work enqueued before it.  Callers that touch device memory
 * from other streams must synchronise before mwg_free. */
mwg_status mwg_malloc(void **ptr, size_t bytes);
mwg_status mwg_free(void *ptr);
/* Page-locked host memory for large transfers: mwg_upload / mwg_download with
 * it on the host side run at the link's DMA rate.  mwg_host_free waits for the
 * device first; after process exit has begun it does nothing. */
mwg_status mwg_host_alloc(void **ptr, size_t bytes);
mwg_status mwg_host_free(void *ptr);
mwg_status mwg_upload(void *dst, const void *src, size_t bytes);
mwg_status mwg_download(void *dst, const void *src, size_t bytes);
/* n downloads (dst[k] <- src[k], bytes[k]) with one wait for the stream. */
mwg_status mwg_download_many(int n, void *const *dst, const void *const *src, const size_t *bytes);
/* Upload through a pinned staging ring and return without waiting: src may be
 * reused on return; the copy is ordered before later work of the library's
 * stream.  One host thread at a time. */
mwg_status mwg_upload_async(void *dst, const void *src, size_t bytes);
mwg_status mwg_copy(void *dst, const void *src, size_t bytes); /* device to device */
mwg_status mwg_zero(void *dst, size_t bytes);
/* Entries of a device key table changed on the host: upd holds n slots, then
 * n keys of two words each (hi, lo), then n values (64-bit each);
 * keys[2 slot], keys[2 slot + 1] = the key's words, vals[slot] = value.
 * Queued on the null stream (no wait). */
mwg_status mwg_hash_update(unsigned long long *keys, int *vals, const long long *upd, int n);
mwg_status mwg_synchronize(void);
mwg_status mwg_release_pool(void);

/* Per-kernel timing: while enabled, every launch is bracketed by HIP events
 * on its stream; mwg_profile_read sums the durations (ms) and launch counts
 * per kernel kind.  mwg_profile(0/1) clears the record. */
enum {
    MWG_KERNEL_SCREEN = 0,
    MWG_KERNEL_GENERATE = 1,
    MWG_KERNEL_BAND = 2,
    MWG_KERNEL_REDUCE = 3,
    MWG_KERNEL_RECONSTRUCT = 4,
    MWG_KERNEL_COMPRESS = 5,
    MWG_KERNEL_NORMS = 6,
    MWG_KERNEL_GROUPS = 7, /* exact group screen and work-unit construction */
    MWG_KERNEL_PROJECT = 8, /* projection of analytic functions onto tree nodes */
    MWG_KERNEL_ARITH = 9,   /* sums and products of trees */
    MWG_KERNEL_KINDS = 10
};
mwg_status mwg_profile(int enable);
/* As mwg_profile, bracketing only the launches of the kinds whose bit
 * (1 << MWG_KERNEL_*) is set in mask (0: off), so that timing one kernel
 * adds no events to the others. */
mwg_status mwg_profile_kinds(unsigned mask);
mwg_status mwg_profile_read(double *ms, long long *launches, int nkinds);

/* Band tables of a convolution operator at one apply precision: widths and
 * band-size factors (ConvolutionCalculator::initBandSizes / calcBandSizeFactor,
 * src/treebuilders/ConvolutionCalculator.cpp:105-139), operator component
 * matrices and norms indexed by (depth, term, translation, component)
 * (OperatorTree::getNode + OperatorNode::getComponentNorm,
 * src/trees/OperatorTree.h:53-57, src/trees/OperatorNode.cpp:56-80), and the
 * separable pair bound used to skip band nodes. All pointers are device. */
typedef struct {
    int M;                 /* separation rank (number of Gaussian terms) */
    int n_depth;           /* depths 0..n_depth-1 */
    int kp1;               /* k+1 */
    int wmax;              /* max band width over depths */
    const int *band_max;   /* [n_depth] */
    const int *widths;     /* [n_depth][M][5]: T,C,B,A,max */
    const int *band_size;  /* [n_depth][M][64]: gt*8+ft -> 64 x product over directions of (2 width + 1) of component 2 gbit + fbit
                              (a negative width drops its factor); the screen recomputes it from widths */
    const long long *mat_start; /* [n_depth][M] first matrix (dl=-w, comp 0) or -1 */
    const double *mats;    /* matrices (kp1*kp1 each, column-major: O(i_f, j_g)) */
    const double *mat_norm;/* one norm per matrix */
    const double *pair_bound; /* [n_depth][2*wmax+1] */
    const double *pair_bound_w; /* [n_depth][2*wmax+1], wavelet components only */
    const int *band_size_max; /* [n_depth][M] row maximum of band_size */
    const double *term_norm;  /* [matrix/4] max component norm per (depth, term, dl) */
    long long n_mat;          /* number of matrices in mats */
    const void *frag[3];      /* MFMA operand tables per sweep precision (MWG_PREC_*), NULL until built */
    const long long *screen_off; /* [n_depth + 1] first entry of depth d in screen_bound: (2 band_max[d] + 1)^3 entries, none if band_max[d] < 0 */
    const double *screen_bound;  /* [entry][8]: per depth, translation difference and f component, the largest
                                    per-term tuple bound over terms (NaN when no term reaches it) */
} mwg_band_tables;

/* Sweep precision of the band sum.  The norm screening, the operator
 * construction and the transforms stay FP64 in every mode. */
enum {
    MWG_PREC_FP64 = 0,   /* FP64 MFMA */
    MWG_PREC_FP32 = 1,   /* FP32 MFMA, band sum accumulated in FP64 */
    MWG_PREC_BF16X3 = 2  /* three-part split BF16 (hh + hm + mh + hl + lh + mm) MFMA, FP32 products and sums, FP64 band sum */
};

/* Size and construction of the MFMA operand table of one precision: every
 * (depth, term, translation) matrix pair of one f-bit, [O_(gbit 0) | O_(gbit 1)],
 * laid out in the lane order of the matrix instruction's B operand.  No
 * reference counterpart (layout of the operator blocks of
 * OperatorTree::getNode, src/trees/OperatorTree.h:53-57, for the device). */
size_t mwg_fragment_bytes(int kp1, long long n_mat, int precision);
mwg_status mwg_build_fragments(const mwg_band_tables *op, int precision, void *dst);

/* Screening table of an operator (op->screen_off set, op->screen_bound
 * ignored): for depth d, translation difference dl (|dl_i| <= band_max[d])
 * and f component ft, the largest over terms t with |dl_i| <= width_t of
 * bs_t * prod_t(ft, dl), the per-term norm bound of the tuple screen of
 * ConvolutionCalculator::applyOperator (src/treebuilders/
 * ConvolutionCalculator.cpp:297-329) without the band node's norm, written to
 * dst[8 * entries].  One table lookup then replaces the loop over terms. */
mwg_status mwg_build_screen_table(const mwg_band_tables *op, long long entries, double *dst);

/* Input function tree resident on the device: an open-addressing hash from a
 * packed node key to a node id, and per-node coefficient offset, 8 component
 * norms + square norm, and a generated-node flag.  Generated nodes hold the
 * scaling block only (FunctionNode::genChildren, src/trees/FunctionNode.cpp:293-328). */
typedef struct {
    int kp1;
    int root_scale;
    int corner[3];
    int nbox[3];
    const unsigned long long *hash_keys; /* [hash_cap][2]: node key words (hi, lo) of keys.h, hi = 0 for an empty slot */
    const int *hash_vals;                /* [hash_cap] */
    int hash_cap;                        /* power of two */
    double *coef;                        /* coefficient pool */
    const long long *coef_off;           /* [nodes] offset into coef (doubles) */
    double *norms;                       /* [nodes][9] */
    const unsigned char *is_gen;         /* [nodes] */
    double norm_max;                     /* at least every node's component norms and sqrt(square norm),
                                            generated nodes included (the screen skips a band position
                                            whose operator bound times it fails the threshold); +inf: no bound */
} mwg_ftree;

/* One run of output nodes of a level that share a tree, a scale and an
 * operator: nodes [first, first + count) of the level.  Trees of different
 * sizes, operators and precisions are batched by giving each its segments. */
typedef struct {
    int first, count;
    int scale;           /* scale of the nodes */
    int o_depth;         /* scale - operator root scale */
    int band_w;          /* band half-width at o_depth (band_max[o_depth] of op), -1 for an empty band */
    double g_thr;        /* screening threshold of the tree (<= 0 disables screening) */
    mwg_ftree f;         /* the tree's input */
    mwg_band_tables op;  /* its operator's tables */
    double *g_coef;      /* the tree's output coefficient pool */
    unsigned long long *counters; /* device [2] mwg_band_apply adds to: applied (f, ft, gt, term) tuples, candidate pairs */
} mwg_segment;

/* Screening-margin bins of a level (mwg_level::g_margin). */
#define MWG_MARGINS 8

/* One breadth-first level: count output nodes, ordered by segment.  segs is
 * HOST memory (the library copies it per call); the arrays are device
 * memory.  Output coefficients go to segs[s].g_coef + g_off[i] for the
 * segment s of node i; norms to g_norms + 9 i. */
typedef struct {
    int kp1;             /* k+1 of every tree in the level */
    int count;
    int nseg;
    const mwg_segment *segs; /* host [nseg], first/count partition [0, count) in order */
    const int *l;            /* [count][3] translations */
    const long long *g_off;  /* [count] */
    double *g_norms;         /* [count][9] */
    const double *node_thr;  /* [count] screening threshold of each node, or NULL for its segment's g_thr
                                (prec * precFunc(idx) * sqrt(|g|^2 / M) of apply with precTrees,
                                src/treebuilders/ConvolutionCalculator.cpp:241-246) */
    /* Optional screening margins (NULL: not computed; no reference
     * counterpart).  For node i and j < MWG_MARGINS: g_margin[MWG_MARGINS i + j] = sum over the
     * node's applied tuples whose test value |Ox||Oy||Oz| * bs * |f_ft| is at
     * most margin_ratio[j] times the node's threshold, of |Ox||Oy||Oz| |f_ft|
     * (Frobenius norms).  It bounds the 2-norm of what those tuples add to the
     * node, i.e. how far the node can move when it is recomputed with a
     * threshold up to margin_ratio[j] times higher. */
    double *g_margin;        /* [count][MWG_MARGINS] */
    double margin_ratio[MWG_MARGINS];
    /* Optional (pair screen only): the level's node count in device memory,
     * when the host knows only the upper bound `count` (a level formed on
     * the device by mwg_next_level); nodes at or above it are skipped. */
    const int *count_dev;
} mwg_level;

/* Candidate (output node, band node) pairs of a level.  pair_f >= 0 is the
 * band node id, pair_f < 0 encodes -(ancestor id + 1) when the band node does
 * not exist yet and must be generated below its nearest existing ancestor
 * (MWTree::getNode -> MWNode::retrieveNode, src/trees/MWNode.cpp:1097-1120). */
typedef struct {
    int capacity;
    int *count;          /* device int[4]: pairs written/needed, missing-node records, nonzero if a
                            segment's band_w is not its tables' band width, pairs whose band node is
                            missing */
    int *pair_g;         /* [capacity] output-node index in the level */
    int *pair_f;         /* [capacity] */
    int *pair_l;         /* [capacity][3] band node translation */
    int *g_first;        /* [level count] first pair of each output node */
    int *g_npairs;       /* [level count] */
    int *missing;        /* [missing_cap][6]: pair, code, node, translation: one record per distinct
                            missing band node of a segment (per missing pair where the node's offsets
                            exceed 17 bits or the segment index 12 bits), in no particular order
                            (count[1] of them) */
    int missing_cap;
    int *missing_pairs;  /* [capacity][2]: every pair whose band node is missing (count[3]): the pair
                            and its record, as a dedup slot (>= 0) or -(record + 1) */
    unsigned long long *dedup_keys; /* [dedup_cap], all zero before the screen; mwg_patch_pairs
                            returns the slots it used to zero */
    int *dedup_rec;      /* [dedup_cap] record of each claimed slot */
    int dedup_cap;       /* power of two, at least 2 capacity; 0: one record per missing pair */
} mwg_pairs;

/* Screen the band of every output node of a level and emit candidate pairs
 * (ConvolutionCalculator::makeOperBand/fillOperBand,
 * src/treebuilders/ConvolutionCalculator.cpp:142-222).  Two upper bounds of
 * the norm test of applyOperator (:297-329) discard band nodes: a product of
 * per-direction maxima over all terms, then a per-term bound over all
 * components; neither can reject a tuple the exact test would apply.  The
 * pairs of one output node are contiguous and in band order (x fastest);
 * *count receives the number of pairs (pairs beyond capacity are counted but
 * not written: grow the arrays and screen again).  With lvl->count_dev the
 * level's count is an upper bound and only nodes below *count_dev are
 * screened (their g_first / g_npairs written). */
mwg_status mwg_pair_screen(const mwg_level *lvl, mwg_pairs *out);

/* One upload for a level: `bytes` of the caller's own level data (returned at
 * *data_dev, 256-byte aligned) together with the level's segment table, the
 * segment of every output node and the pair screen's block table, so that
 * mwg_pair_screen and mwg_band_apply on this level (the same segments, same
 * count) upload nothing more.  Only lvl->segs, nseg and count are read; the
 * caller then points lvl->g_off, lvl->l and lvl->node_thr into *data_dev.
 * Queued on the null stream; *data_dev stays valid until the next call.  No
 * reference counterpart. */
mwg_status mwg_level_upload(const mwg_level *lvl, const void *data, size_t bytes, void **data_dev);

/* The split decisions of a finished breadth-first step of one tree at one
 * scale and the next step's nodes, formed on the device (TreeBuilder::build's
 * norm bookkeeping and WaveletAdaptor::splitNode,
 * src/treebuilders/TreeBuilder.cpp:48-80, src/treebuilders/WaveletAdaptor.h:
 * 45-56, src/utils/tree_utils.cpp:47-65), for a level whose nodes have no
 * children yet.  Every quantity is formed with the host's IEEE operations in
 * the host's order (no fused multiply-add; the step's wavelet norms summed in
 * node order), so the decisions equal the host's bit for bit:
 *   w_i = sum_{t=1..7} n_t^2 over the node's component norms (-1 if one is negative);
 *   w' = w_norm + sum_i w_i; tsn = s_norm < 0 || w' < 0 ? -1 : s_norm + w';
 *   thr = max(thr_floor, prec * (tsn > 0 && !abs_prec ? sqrt(tsn) : 1) * scale_fac);
 *   node i splits iff may_split && prec > 0 && sqrt(w_i) > thr;
 *   screening threshold of the next step = tsn > 0 ? prec * 1.0 * sqrt(tsn / n_terms) : tsn.
 * The children of the k-th splitting node (node order) get ids base_id + 8 k + c,
 * c the child index, translation 2 l + bit d of c.  One workgroup; at most
 * 4096 nodes (the sum over the level is sequential).
 * No reference counterpart as an entry point. */
typedef struct {
    int count;                 /* nodes of the finished level */
    int may_split;             /* scale + 2 <= the MRA's max scale, and not the last step (maxIter) */
    int abs_prec;
    double prec, s_norm, w_norm;
    double scale_fac;          /* 2^-(0.5 (scale + 1)), as the host forms it */
    double n_terms;            /* separation rank of the operator */
    double thr_floor;          /* 2 MachinePrec */
    long long base_id, ncoef;
    const double *g_norms;     /* [count][9] the finished level's component norms and square norm */
    const int *l;              /* [count][3] its translations */
    double *head;              /* out [4]: tsn, next screening threshold, w', number of children */
    int *split;                /* out [count] 0 / 1 */
    long long *next_off;       /* out [8 count]: child id * ncoef */
    int *next_l;               /* out [8 count][3] */
    double *next_thr;          /* out [8 count]: the next screening threshold */
    int *next_count;           /* out: number of children */
} mwg_next_level_args;
mwg_status mwg_next_level(const mwg_next_level_args *a);

/* For each of the count (= count[3] of the screen) pairs whose band node was
 * missing: pair_f = node_of_record[its record] (band nodes generated on the
 * host's plan of the records); the dedup slots used return to zero. */
mwg_status mwg_patch_pairs(mwg_pairs *pairs, const int *node_of_record, int count);

/* Generate scaling-only band nodes below existing parents: node ids[i] gets
 * the scaling block of child child_idx[i] of node parent[i], written at
 * coef_off[ids[i]], and its norms (MWNode::threadSafeGenChildren +
 * giveChildrenCoefs, src/trees/MWNode.cpp:312-335,410-418;
 * tree_utils::mw_transform, src/utils/tree_utils.cpp:113-216). */
mwg_status mwg_generate_nodes(const mwg_ftree *f, const double *filters, const int *ids, const int *parent, const int *child_idx,
                              int count);
/* As mwg_generate_nodes, and the new nodes are recorded in the tree's own
 * arrays first: coef_off[ids[i]] = offs[i] (the block is written there) and
 * is_gen[ids[i]] = 1, so a caller extending a tree needs no separate offset
 * and flag uploads.  coef_off_w and is_gen_w must be f->coef_off and
 * f->is_gen.  Nodes of one call must not be each other's parents (one call
 * per scale, coarse to fine). */
mwg_status mwg_generate_nodes_recorded(const mwg_ftree *f, const double *filters, const int *ids, const int *parent, const int *child_idx,
                                       const long long *offs, long long *coef_off_w, unsigned char *is_gen_w, int count);

/* The band nodes of several trees in one launch: node i of the call is
 * generated and recorded as by mwg_generate_nodes_recorded in tree
 * trees[tree_of[i]] (trees: a device array of ntrees entries, all of order
 * kp1, coef_off_w / is_gen_w the tree's own f.coef_off / f.is_gen).  Nodes of
 * one call must not be each other's parents (one call per scale, coarse to
 * fine).  No reference counterpart (one launch for a level batch of trees). */
typedef struct {
    mwg_ftree f;
    long long *coef_off_w;
    unsigned char *is_gen_w;
} mwg_gen_tree;
mwg_status mwg_generate_nodes_multi(int kp1, const mwg_gen_tree *trees, int ntrees, const double *filters, const int *tree_of, const int *ids,
                                    const int *parent, const int *child_idx, const long long *offs, int count);
/* Key-table slots of several trees in one launch: upd is a device array of n
 * records (table, slot, key hi, key lo, value), each written into
 * tables[table] (a device array of ntables).  No reference counterpart. */
typedef struct {
    unsigned long long *keys; /* [cap][2] */
    int *vals;                /* [cap] */
} mwg_hash_table;
mwg_status mwg_hash_update_multi(const mwg_hash_table *tables, int ntables, const long long *upd, int n);

/* Exact screening and the 3-sweep band sum of every candidate pair, then the
 * output coefficients of the level and their component norms
 * (ConvolutionCalculator::calcNode, applyOperComp, applyOperator and
 * tensorApplyOperComp, src/treebuilders/ConvolutionCalculator.cpp:224-382;
 * MWNode::calcNorms, src/trees/MWNode.cpp:609-655).  Every pair_f must be a
 * node id (missing band nodes generated first).  Adds the number of applied
 * (f, ft, gt, term) tuples to each segment's counters[0].  Summation order is fixed by the
 * pair order, so results are reproducible.  precision: MWG_PREC_*; every
 * segment's op.frag[precision] must be built.  pair_f of a pair refers to the
 * input tree of its output node's segment. */
mwg_status mwg_band_apply(const mwg_level *lvl, const mwg_pairs *pairs, int npairs, int precision);

/* Work distribution of the matrix-core band kernel at order 7 (kp1 = 8) for
 * MWG_PREC_FP64 and MWG_PREC_FP32.  MWG_BAND_ROWS: each wave of a workgroup
 * owns 16 rows of every sweep, so only the x -> y step crosses waves (one
 * workgroup barrier per group) and the z sweep is shared evenly whatever the
 * group's gt mask.  MWG_BAND_TILES: the tiles of each sweep are shared out
 * over the waves (two barriers per group), the layout of every other order
 * and precision.  MWG_BAND_AUTO (the default) is MWG_BAND_ROWS where it
 * applies.  The layouts give bit-identical results.  Process-wide; no
 * reference counterpart. */
enum { MWG_BAND_AUTO = 0, MWG_BAND_TILES = 1, MWG_BAND_ROWS = 2 };
mwg_status mwg_set_band_layout(int layout);
int mwg_band_layout(void);

/* Launch parameters of the band sum, process-wide: each work unit holds at
 * least unit_groups groups of an output node (default 32; 8..1024), and the
 * persistent band kernels run at most workgroups_per_cu workgroups per
 * compute unit (0, the default: as many as fit).  Results depend on
 * unit_groups only through the order in which a node's partial sums are
 * added (last bits); the defaults are what the parity tests hold.  No
 * reference counterpart (the launch shape of calcNodeVector's work). */
mwg_status mwg_set_band_launch(int unit_groups, int workgroups_per_cu);
void mwg_band_launch(int *unit_groups, int *workgroups_per_cu);

/* Add each parent's reconstructed s/d into the scaling block of its 8 children
 * (MWTree::mwTransform(TopDown, false) -> tree_utils::mw_transform,
 * src/trees/MWTree.cpp:194-216, src/utils/tree_utils.cpp:113-216).  Children
 * of parent i start at child_off[i] and are stride apart. */
mwg_status mwg_reconstruct_add(int kp1, const double *filters, double *coef, const long long *parent_off,
                               const long long *child_off, int count, long long stride);

/* Compress the children's scaling blocks into the parent's s/d
 * (MWTree::mwTransform(BottomUp) -> FunctionNode<3>::reCompress ->
 * tree_utils::mw_transform_back, src/trees/FunctionNode.cpp:391-409,
 * src/utils/tree_utils.cpp:229-301). */
mwg_status mwg_compress(int kp1, const double *filters, double *coef, const long long *parent_off,
                        const long long *child_off, int count, long long stride);

/* Component norms of count nodes at coef + off[i], written to norms[9 i]. */
mwg_status mwg_node_norms(int kp1, const double *coef, const long long *off, double *norms, int count);

/* Projection of a sum of Gaussians, sum_g coef_g exp(-sum_d alpha_gd (r_d - pos_gd)^2),
 * onto count nodes of one tree level (ProjectionCalculator::calcNode,
 * src/treebuilders/ProjectionCalculator.cpp:34-51: function values at the
 * Gauss-Legendre points of the node's 8 children, MWNode::getExpandedChildPts
 * src/trees/MWNode.cpp:854-925; value -> coefficient map, MWNode::cvTransform
 * :448-490; compression to the node's scaling and wavelet blocks,
 * MWNode::mwTransform; MWNode::calcNorms :609-616).  Node i has scale[i] and
 * translation l[3 i ..]; its 8 (k+1)^3 coefficients go to coef + off[i] and
 * its 8 component norms and square norm to norms[9 i].  All arrays are device
 * memory; gauss holds ngauss terms. */
typedef struct {
    double coef;
    double alpha[3];
    double pos[3];
} mwg_gauss;
typedef struct {
    int kp1;
    double sf[3];            /* world box scaling factors */
    const double *nodes;     /* [kp1] quadrature nodes on [0, 1] */
    const double *vcmap;     /* [kp1 * kp1] value -> coefficient map, column-major */
    const double *filters;   /* mwg_filters_upload */
    const mwg_gauss *gauss;
    int ngauss;
} mwg_projection;
mwg_status mwg_project_nodes(const mwg_projection *pj, const int *scale, const int *l, const long long *off, double *coef, double *norms,
                             int count);

/* count operator-tree nodes of one refinement level, any mix of separated
 * terms (CrossCorrelationCalculator::calcNode / applyCcc,
 * src/treebuilders/CrossCorrelationCalculator.cpp:36-96; MWNode::mwTransform
 * compression; OperatorNode::calcComponentNorm, src/trees/OperatorNode.cpp:56-80).
 * kernel_blocks: scaling blocks (2 kp1 doubles each) of the 1D kernel nodes at
 * the children's scale; node i at translation (l0, l1) with D = 2 (l1 - l0)
 * reads the four consecutive blocks from slot[i] on, at translations D - 2 ..
 * D + 1 (child c, x bit c & 1, y bit c >> 1, uses l1' - l0' - 1 and
 * l1' - l0').  fac[i] = sqrt(sf) 2^-(n+1)/2, thr[i] the norm threshold
 * max(MachinePrec, prec / (8 2^depth)).  left / right: the cross-correlation
 * tables (kp1^2 rows of 2 kp1, row-major).  Writes 4 kp1^2 coefficients per
 * node to coef and 4 component norms + square norm to norms[5 i].  Results
 * are bit-identical to the host formation (no fused multiply-add, the host's
 * summation order). */
mwg_status mwg_operator_nodes(int kp1, const double *left, const double *right, const double *filters, const double *kernel_blocks,
                              const int *slot, const double *fac, const double *thr, double *coef, double *norms, int count);

/* Tree arithmetic over output nodes of any scales (AdditionCalculator /
 * MultiplicationCalculator::calcNode, src/treebuilders/AdditionCalculator.h:45-70,
 * src/treebuilders/MultiplicationCalculator.h:45-83).  Mode 0: the node's
 * coefficients are sum_i c[i] f_i; mode 1: every operand node is reconstructed
 * to its children and mapped to values at their quadrature points, the values
 * multiplied (each times c[i]), mapped back and compressed.  in_off[k nin + i]
 * >= 0: operand i's node k (8 blocks) at that offset in pool[i]; < 0: a
 * generated node, its scaling block at -in_off - 1 (wavelet blocks zero).
 * scale[k] is the node's scale, out_off[k] its offset in out; norms[9 k] the 8
 * component norms and the square norm.  No fused multiply-add: the host's
 * order of operations. */
#define MWG_MAX_OPERANDS 8
typedef struct {
    int kp1;
    int mode; /* 0 sum, 1 product */
    int nin;  /* operands, 1 .. MWG_MAX_OPERANDS */
    double sf[3];
    const double *pool[MWG_MAX_OPERANDS];
    double c[MWG_MAX_OPERANDS];
    const double *filters; /* mwg_filters_upload */
    const double *vcmap;   /* [kp1 * kp1] value -> coefficient map, column-major */
    const double *cvmap;   /* [kp1 * kp1] coefficient -> value map, column-major */
} mwg_combine;
mwg_status mwg_combine_nodes(const mwg_combine *cb, const long long *in_off, const int *scale, const long long *out_off, double *out,
                             double *norms, int count);

/* Inner products of count node pairs (dot, src/treebuilders/multiply.cpp:286-318):
 * partial[2 k] = sum of the scaling blocks' products when root[k] is set
 * (else 0), partial[2 k + 1] = sum over the 7 wavelet blocks; node k of a at
 * a + a_off[k], of b at b + b_off[k] (8 blocks each).  Each node's sums run in
 * a fixed order (deterministic). */
mwg_status mwg_dot_nodes(int kp1, const double *a, const long long *a_off, const double *b, const long long *b_off, const int *root,
                         double *partial, int count);

/* Filter bank upload: 8 matrices (compression H0t,H1t,G0t,G1t then
 * reconstruction H0,G0,H1,G1), kp1*kp1 column-major each, from host memory;
 * returns a device pointer (MWFilter::getSubFilter, src/core/MWFilter.cpp:100-133). */
mwg_status mwg_filters_upload(int kp1, const double *host_filters, double **dev_filters);

/* Refinement predictor (no reference counterpart): a 12 -> 64 -> 64 -> 1
 * ReLU perceptron in FP32 that predicts, per output node, whether the
 * adaptive apply refines it.  params: MWG_MLP_PARAMS floats laid out
 * W1[12][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.  features:
 * [n][12] floats; labels: [n] floats in {0, 1}.  All device memory. */
#define MWG_MLP_FEATURES 12
#define MWG_MLP_HIDDEN 64
#define MWG_MLP_PARAMS (MWG_MLP_FEATURES * 64 + 64 + 64 * 64 + 64 + 64 + 1)
/* prob[i] = sigmoid(logit of sample i) */
mwg_status mwg_mlp_forward(const float *params, const float *features, int n, float *prob);
/* grad[0..PARAMS) = gradient of the mean loss
 * pos_weight * y * softplus(-z) + (1 - y) * softplus(z); grad[PARAMS] = the
 * mean loss.  grad holds MWG_MLP_PARAMS + 1 floats.  Deterministic. */
mwg_status mwg_mlp_gradient(const float *params, const float *features, const float *labels, int n, float pos_weight, float *grad);
/* mwg_mlp_gradient without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (a training step captured in a
 * HIP graph, loops that forbid hipMalloc).  Same arguments and the same
 * result bit for bit.  Its scratch (512 per-wave partial vectors, 10,358,784
 * bytes) is a file-scope device array that each GPU holds from mwg_init until
 * the process ends, whatever n is (no input size exceeds it).  Every call on a
 * GPU shares that GPU's array and writes all of it, so nothing carries over
 * between calls; calls that overlap on one GPU (several host threads or
 * streams) race on it and must be serialised by the caller.  Needs mwg_init
 * for the current device. */
mwg_status mwg_mlp_gradient_noalloc(const float *params, const float *features, const float *labels, int n, float pos_weight, float *grad);
/* One Adam step (step counts from 1) on params with first/second moments m, v. */
mwg_status mwg_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif /* MWG_H */
