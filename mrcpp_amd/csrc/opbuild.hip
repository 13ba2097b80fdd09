// Operator-tree nodes of one refinement level, for every separated term of an
// operator at once (CrossCorrelationCalculator::calcNode / applyCcc,
// src/treebuilders/CrossCorrelationCalculator.cpp:36-96, then the node's
// compression, MWNode::mwTransform, and OperatorNode::calcComponentNorm,
// src/trees/OperatorNode.cpp:56-80).
//
// Every product and sum is formed in the host's order with no fused
// multiply-add, so the nodes are bit-identical to the host construction (the
// operator norms decide the screening of the band sum).
#pragma clang fp contract(off)

#include "device_common.h"

namespace mwgdev {

namespace {

constexpr int kOpThreads = 256;

// One workgroup per node.  blocks: 1D kernel scaling blocks (2K values each,
// K = k + 1); the node's four consecutive blocks from slot[k] hold the
// translations D - 2 .. D + 1, D = 2 (l1 - l0).  Child c (x bit c & 1, y bit
// c >> 1) correlates the blocks at l1' - l0' - 1 and l1' - l0', where
// l1' - l0' = D + (c >> 1) - (c & 1): offsets 1 + (c >> 1) - (c & 1) and one more.
template <int N1>
__global__ __launch_bounds__(kOpThreads) void k_cross_correlate(const double *left, const double *right, const double *filt, const double *blocks,
                                                                 const int *slot, const double *fac, const double *thr, double *coef,
                                                                 double *norms) {
    constexpr int KD = N1 * N1, K2 = 2 * N1;
    __shared__ double a[4 * KD], b[4 * KD];
    __shared__ double s1d[8 * K2];
    const int k = blockIdx.x;
    const size_t s0 = (size_t)slot[k];
    for (int i = threadIdx.x; i < 8 * K2; i += kOpThreads) {
        const int c = i / (2 * K2), which = (i / K2) & 1;
        const int off = 1 + ((c >> 1) & 1) - (c & 1) + which;
        s1d[i] = blocks[(s0 + off) * K2 + i % K2];
    }
    __syncthreads();
    const double f = fac[k];
    for (int o = threadIdx.x; o < 4 * KD; o += kOpThreads) {
        const int c = o / KD, r = o % KD;
        const double *L = left + (size_t)r * K2, *R = right + (size_t)r * K2;
        const double *sa = s1d + (2 * c) * K2, *sb = s1d + (2 * c + 1) * K2;
        double xa = 0.0, xb = 0.0;
        for (int p = 0; p < K2; p++) {
            xa += L[p] * sa[p];
            xb += R[p] * sb[p];
        }
        a[o] = f * (xa + xb);
    }
    __syncthreads();
    // compression, direction 0 then 1: out[gt] = in[ft0]^T F(2 b + 0) + in[ft1]^T F(2 b + 1)
    for (int dim = 0; dim < 2; dim++) {
        const double *src = dim == 0 ? a : b;
        double *dst = dim == 0 ? b : a;
        const int bit = 1 << dim;
        for (int o = threadIdx.x; o < 4 * KD; o += kOpThreads) {
            const int gt = o / KD, e = o % KD;
            const int r = e % N1, j = e / N1;
            const int ft0 = gt & ~bit, ft1 = gt | bit;
            const double *F0 = filt + (size_t)(2 * ((gt >> dim) & 1) + 0) * KD + N1 * j;
            const double *F1 = filt + (size_t)(2 * ((gt >> dim) & 1) + 1) * KD + N1 * j;
            const double *x0 = src + (size_t)ft0 * KD + N1 * r;
            const double *x1 = src + (size_t)ft1 * KD + N1 * r;
            double s0 = 0.0, s1 = 0.0;
            for (int i = 0; i < N1; i++) {
                s0 += x0[i] * F0[i];
                s1 += x1[i] * F1[i];
            }
            dst[o] = s0 + s1;
        }
        __syncthreads();
    }
    double *dst = coef + (size_t)k * 4 * KD;
    for (int o = threadIdx.x; o < 4 * KD; o += kOpThreads) dst[o] = a[o];
    // component norms: Frobenius norm if it and sqrt(|.|_1 |.|_inf) pass the threshold, else 0
    if (threadIdx.x < 4) {
        const int t = threadIdx.x;
        const double *m = a + (size_t)t * KD;
        double fro = 0.0;
        for (int i = 0; i < KD; i++) fro += m[i] * m[i];
        fro = sqrt(fro);
        const double th = thr[k];
        double v = 0.0;
        if (fro > th) {
            double rmax = 0.0, cmax = 0.0;
            for (int r = 0; r < N1; r++) {
                double s = 0.0;
                for (int c = 0; c < N1; c++) s += fabs(m[r + N1 * c]);
                rmax = fmax(rmax, s);
            }
            for (int c = 0; c < N1; c++) {
                double s = 0.0;
                for (int r = 0; r < N1; r++) s += fabs(m[r + N1 * c]);
                cmax = fmax(cmax, s);
            }
            if (sqrt(rmax * cmax) > th && fro > th) v = fro;
        }
        norms[(size_t)k * 5 + t] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int t = 0; t < 4; t++) {
            const double v = norms[(size_t)k * 5 + t];
            tot += v * v;
        }
        norms[(size_t)k * 5 + 4] = tot;
    }
}

template <int N1> struct CrossL {
    static mwg_status run(const double *left, const double *right, const double *filt, const double *blocks, const int *slot, const double *fac,
                          const double *thr, double *coef, double *norms, int count) {
        hipLaunchKernelGGL(k_cross_correlate<N1>, dim3(count), dim3(kOpThreads), 0, 0, left, right, filt, blocks, slot, fac, thr, coef, norms);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_operator_nodes(int kp1, const double *left, const double *right, const double *filters, const double *kernel_blocks,
                                         const int *slot, const double *fac, const double *thr, double *coef, double *norms, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count < 0) return fail(MWG_ERR_ARG, "negative node count");
    if (count == 0) return MWG_OK;
    if (!left || !right || !filters || !kernel_blocks || !slot || !fac || !thr || !coef || !norms) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<CrossL>(kp1, left, right, filters, kernel_blocks, slot, fac, thr, coef, norms, count);
}
