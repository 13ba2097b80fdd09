// Multiwavelet transforms on the device: generated band nodes, top-down
// reconstruction, bottom-up compression and component norms.  One 256-thread
// workgroup per node; the 8 blocks of a node are staged in LDS.
#include "device_common.h"

namespace mwgdev {

namespace {

// The generation, top-down and bottom-up passes are latency-bound for the
// few nodes of a launch: their workgroups have kPassThreads threads, each forming a 2 x 2
// register tile of a pass (the same per-output sum, over i in increasing
// order, as the 4 x 4 tiles of mw_pass, so the same results).
constexpr int kPassThreads = 1024;
template <int N1> struct SmallTile {
    static constexpr int N2 = N1 * N1, J = 2, R = 2, RB = N2 / R, JG = N1 / J, TILES = 8 * JG * RB;
    static_assert(N1 % J == 0 && N2 % R == 0, "small pass tile shape");
};
template <int N1>
__device__ __forceinline__ void small_tile(const double *in, const double *filt, int dir_base, int dim, int t, int &gt, int &j0, int &rb, double (&s0)[2][2],
                                           double (&s1)[2][2]) {
    using P = SmallTile<N1>;
    constexpr int N2 = P::N2, N3 = N2 * N1, J = P::J, R = P::R, RB = P::RB;
    gt = t / (P::JG * RB);
    j0 = ((t / RB) % P::JG) * J;
    rb = t % RB;
    const int bit = 1 << dim;
    const int ft0 = gt & ~bit, ft1 = gt | bit;
    const double *F0 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 0) * N2 + N1 * j0;
    const double *F1 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 1) * N2 + N1 * j0;
#pragma unroll
    for (int p = 0; p < J; p++)
#pragma unroll
        for (int q = 0; q < R; q++) {
            s0[p][q] = 0.0;
            s1[p][q] = 0.0;
        }
#pragma unroll 2
    for (int i = 0; i < N1; i++) {
        double x0[R], x1[R], f0[J], f1[J];
#pragma unroll
        for (int q = 0; q < R; q++) {
            x0[q] = in[pad<N1>(ft0 * N3 + N1 * (rb + q * RB) + i)];
            x1[q] = in[pad<N1>(ft1 * N3 + N1 * (rb + q * RB) + i)];
        }
#pragma unroll
        for (int p = 0; p < J; p++) {
            f0[p] = F0[N1 * p + i];
            f1[p] = F1[N1 * p + i];
        }
#pragma unroll
        for (int p = 0; p < J; p++)
#pragma unroll
            for (int q = 0; q < R; q++) {
                s0[p][q] += x0[q] * f0[p];
                s1[p][q] += x1[q] * f1[p];
            }
    }
}
template <int N1> __device__ __forceinline__ void small_pass(double *out, const double *in, const double *filt, int dir_base, int dim) {
    using P = SmallTile<N1>;
    constexpr int N2 = P::N2, N3 = N2 * N1;
    for (int t = threadIdx.x; t < P::TILES; t += kPassThreads) {
        double s0[2][2], s1[2][2];
        int gt, j0, rb;
        small_tile<N1>(in, filt, dir_base, dim, t, gt, j0, rb, s0, s1);
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++) out[pad<N1>(gt * N3 + (j0 + p) * N2 + rb + q * P::RB)] = s0[p][q] + s1[p][q];
    }
}
template <int N1>
__device__ __forceinline__ void small_pass_add_children(double *coef, long long child_off, long long stride, const double *in, const double *filt) {
    using P = SmallTile<N1>;
    constexpr int N2 = P::N2;
    for (int t = threadIdx.x; t < P::TILES; t += kPassThreads) {
        double s0[2][2], s1[2][2];
        int gt, j0, rb;
        small_tile<N1>(in, filt, 4, 2, t, gt, j0, rb, s0, s1);
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int q = 0; q < 2; q++) {
                double *dst = coef + child_off + (long long)gt * stride + (j0 + p) * N2 + rb + q * P::RB;
                double v = *dst;
                v += s0[p][q];
                v += s1[p][q];
                *dst = v;
            }
    }
}

// The scaling block of child c of node pid (staged filters sfl; a, b, red:
// LDS of the workgroup), written at offs[k] (recorded in the tree) or at the
// node's own offset, and its norms.
template <int N1>
__device__ __forceinline__ void generate_node(const mwg_ftree &f, const double *sfl, int k, int id, int pid, int c, const long long *offs,
                                              long long *coef_off_w, unsigned char *is_gen_w, double *a, double *b, double *red) {
    constexpr int N3 = N1 * N1 * N1;
    const double *src = f.coef + f.coef_off[pid];
    const int nblk = f.is_gen[pid] ? 1 : 8; // generated parents carry the scaling block only
    for (int o = threadIdx.x; o < 8 * N3; o += kPassThreads) a[pad<N1>(o)] = (o < nblk * N3) ? src[o] : 0.0;
    __syncthreads();
    small_pass<N1>(b, a, sfl, 4, 0);
    __syncthreads();
    small_pass<N1>(a, b, sfl, 4, 1);
    __syncthreads();
    small_pass<N1>(b, a, sfl, 4, 2);
    __syncthreads();
    double *dst = f.coef + (offs ? offs[k] : f.coef_off[id]);
    if (offs && threadIdx.x == 0) {
        coef_off_w[id] = offs[k];
        is_gen_w[id] = 1;
    }
    for (int e = threadIdx.x; e < N3; e += kPassThreads) dst[e] = b[pad<N1>(c * N3 + e)];
    // the norm summed by the first kThreads threads in their order (the
    // order of the band screen's other norms)
    if (threadIdx.x < kThreads) {
        double ss = 0.0;
        for (int e = threadIdx.x; e < N3; e += kThreads) {
            const double v = b[pad<N1>(c * N3 + e)];
            ss += v * v;
        }
        for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < kThreads / 64; w++) v += red[w];
        const double n0 = sqrt(v);
        double *nr = f.norms + (size_t)id * 9;
        nr[0] = n0;
        for (int t = 1; t < 8; t++) nr[t] = 0.0;
        nr[8] = n0 * n0;
    }
}

// Scaling block of one child of an existing node (MWNode::threadSafeGenChildren
// + giveChildrenCoefs, src/trees/MWNode.cpp:312-335,410-418).
// With offs given, the node is also recorded in the mirror: coef_off_w[id] =
// offs[k] (where its block goes) and is_gen_w[id] = 1.
template <int N1>
__global__ __launch_bounds__(kPassThreads) void k_generate(mwg_ftree f, const double *filt, const int *ids, const int *parent, const int *child,
                                                           const long long *offs, long long *coef_off_w, unsigned char *is_gen_w) {
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ double sfl[8 * N1 * N1];
    for (int i = threadIdx.x; i < 8 * N1 * N1; i += kPassThreads) sfl[i] = filt[i];
    const int k = blockIdx.x;
    generate_node<N1>(f, sfl, k, ids[k], parent[k], child[k], offs, coef_off_w, is_gen_w, a, b, red);
}

// The same for nodes of several trees: node k in trees[tree_of[k]].
template <int N1>
__global__ __launch_bounds__(kPassThreads) void k_generate_multi(const mwg_gen_tree *trees, const double *filt, const int *tree_of, const int *ids,
                                                                 const int *parent, const int *child, const long long *offs) {
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ double sfl[8 * N1 * N1];
    for (int i = threadIdx.x; i < 8 * N1 * N1; i += kPassThreads) sfl[i] = filt[i];
    const int k = blockIdx.x;
    const mwg_gen_tree &t = trees[tree_of[k]];
    generate_node<N1>(t.f, sfl, k, ids[k], parent[k], child[k], offs, t.coef_off_w, t.is_gen_w, a, b, red);
}

template <int N1>
__global__ __launch_bounds__(kPassThreads) void k_reconstruct_add(const double *filt, double *coef, const long long *poff, const long long *coff,
                                                                  long long stride) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double sfl[8 * N1 * N1];
    for (int i = threadIdx.x; i < 8 * N1 * N1; i += kPassThreads) sfl[i] = filt[i];
    const int k = blockIdx.x;
    const double *src = coef + poff[k];
    for (int o = threadIdx.x; o < 8 * N3; o += kPassThreads) a[pad<N1>(o)] = src[o];
    __syncthreads();
    small_pass<N1>(b, a, sfl, 4, 0);
    __syncthreads();
    small_pass<N1>(a, b, sfl, 4, 1);
    __syncthreads();
    small_pass_add_children<N1>(coef, coff[k], stride, a, sfl);
}

template <int N1>
__global__ __launch_bounds__(kPassThreads) void k_compress(const double *filt, double *coef, const long long *poff, const long long *coff, long long stride) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double sfl[4 * N1 * N1];
    for (int i = threadIdx.x; i < 4 * N1 * N1; i += kPassThreads) sfl[i] = filt[i];
    const int k = blockIdx.x;
    for (int o = threadIdx.x; o < 8 * N3; o += kPassThreads) {
        const int t = o / N3, e = o % N3;
        a[pad<N1>(o)] = coef[coff[k] + (long long)t * stride + e];
    }
    __syncthreads();
    small_pass<N1>(b, a, sfl, 0, 0);
    __syncthreads();
    small_pass<N1>(a, b, sfl, 0, 1);
    __syncthreads();
    small_pass<N1>(b, a, sfl, 0, 2);
    __syncthreads();
    double *dst = coef + poff[k];
    for (int o = threadIdx.x; o < 8 * N3; o += kPassThreads) dst[o] = b[pad<N1>(o)];
}

template <int N1> __global__ __launch_bounds__(kThreads) void k_node_norms(const double *coef, const long long *off, double *norms) {
    __shared__ double red[8 * (kThreads / 64)];
    const int k = blockIdx.x;
    block_norms<N1>(coef + off[k], norms + (size_t)k * 9, red);
}

template <int N1> struct GenerateL {
    static mwg_status run(const mwg_ftree *f, const double *filt, const int *ids, const int *par, const int *ch, int count, const long long *offs,
                          long long *coef_off_w, unsigned char *is_gen_w) {
        ProfileScope prof_(MWG_KERNEL_GENERATE);
        hipLaunchKernelGGL(k_generate<N1>, dim3(count), dim3(kPassThreads), 0, 0, *f, filt, ids, par, ch, offs, coef_off_w, is_gen_w);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct GenerateMultiL {
    static mwg_status run(const mwg_gen_tree *trees, const double *filt, const int *tree_of, const int *ids, const int *par, const int *ch,
                          const long long *offs, int count) {
        ProfileScope prof_(MWG_KERNEL_GENERATE);
        hipLaunchKernelGGL(k_generate_multi<N1>, dim3(count), dim3(kPassThreads), 0, 0, trees, filt, tree_of, ids, par, ch, offs);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct ReconL {
    static mwg_status run(const double *filt, double *coef, const long long *po, const long long *co, int count, long long stride) {
        ProfileScope prof_(MWG_KERNEL_RECONSTRUCT);
        hipLaunchKernelGGL(k_reconstruct_add<N1>, dim3(count), dim3(kPassThreads), 0, 0, filt, coef, po, co, stride);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct CompressL {
    static mwg_status run(const double *filt, double *coef, const long long *po, const long long *co, int count, long long stride) {
        ProfileScope prof_(MWG_KERNEL_COMPRESS);
        hipLaunchKernelGGL(k_compress<N1>, dim3(count), dim3(kPassThreads), 0, 0, filt, coef, po, co, stride);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct NormsL {
    static mwg_status run(const double *coef, const long long *off, double *norms, int count) {
        ProfileScope prof_(MWG_KERNEL_NORMS);
        hipLaunchKernelGGL(k_node_norms<N1>, dim3(count), dim3(kThreads), 0, 0, coef, off, norms);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" {

mwg_status mwg_generate_nodes(const mwg_ftree *f, const double *filters, const int *ids, const int *parent, const int *child_idx, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!f || !filters || !ids || !parent || !child_idx) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<GenerateL>(f->kp1, f, filters, ids, parent, child_idx, count, (const long long *)nullptr, (long long *)nullptr,
                               (unsigned char *)nullptr);
}

mwg_status mwg_generate_nodes_recorded(const mwg_ftree *f, const double *filters, const int *ids, const int *parent, const int *child_idx,
                                       const long long *offs, long long *coef_off_w, unsigned char *is_gen_w, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!f || !filters || !ids || !parent || !child_idx || !offs || !coef_off_w || !is_gen_w) return fail(MWG_ERR_ARG, "null argument");
    if (coef_off_w != f->coef_off || is_gen_w != f->is_gen) return fail(MWG_ERR_ARG, "recorded offsets and flags must be the tree's own arrays");
    return dispatch<GenerateL>(f->kp1, f, filters, ids, parent, child_idx, count, offs, coef_off_w, is_gen_w);
}

mwg_status mwg_generate_nodes_multi(int kp1, const mwg_gen_tree *trees, int ntrees, const double *filters, const int *tree_of, const int *ids,
                                    const int *parent, const int *child_idx, const long long *offs, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!trees || ntrees <= 0 || !filters || !tree_of || !ids || !parent || !child_idx || !offs) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<GenerateMultiL>(kp1, trees, filters, tree_of, ids, parent, child_idx, offs, count);
}

mwg_status mwg_reconstruct_add(int kp1, const double *filters, double *coef, const long long *parent_off, const long long *child_off, int count,
                               long long stride) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!filters || !coef || !parent_off || !child_off) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<ReconL>(kp1, filters, coef, parent_off, child_off, count, stride);
}

mwg_status mwg_compress(int kp1, const double *filters, double *coef, const long long *parent_off, const long long *child_off, int count,
                        long long stride) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!filters || !coef || !parent_off || !child_off) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<CompressL>(kp1, filters, coef, parent_off, child_off, count, stride);
}

mwg_status mwg_node_norms(int kp1, const double *coef, const long long *off, double *norms, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!coef || !off || !norms) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<NormsL>(kp1, coef, off, norms, count);
}

} // extern "C"
