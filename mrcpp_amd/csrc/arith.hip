// Sums and products of trees over one batch of output nodes
// (AdditionCalculator::calcNode, src/treebuilders/AdditionCalculator.h:45-70;
// MultiplicationCalculator::calcNode, src/treebuilders/MultiplicationCalculator.h:45-83).
//
// One 256-thread workgroup per output node.  A sum reads the operands' nodes
// and writes their linear combination.  A product takes each operand's node
// through its reconstruction to the 8 children's scaling blocks and the
// coefficient -> value maps (MWNode::mwTransform(Reconstruction),
// MWNode::cvTransform(Forward), src/trees/MWNode.cpp:448-594) in LDS,
// multiplies the values into registers, then maps back and compresses.  The
// file is compiled without fused multiply-adds and every sum runs in the
// host's order, so the results equal the host formation.
#pragma clang fp contract(off)

#include "device_common.h"

namespace mwgdev {

namespace {

// out(r, j) = sum_i in(i, r) M(i, j) for each of the 8 blocks (the host's
// math_utils::apply_filter view, in N1 x N1^2 column-major), padded LDS layout.
template <int N1> __device__ inline void map_pass(double *out, const double *in, const double *M) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const int t = o / N3, e = o % N3;
        const int r = e % N2, j = e / N2;
        const double *x = in + pad<N1>(t * N3 + N1 * r);
        const double *m = M + N1 * j;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N1; i++) s += x[i] * m[i];
        out[pad<N1>(o)] = s;
    }
}

// operand i's node k into LDS (a generated node: scaling block, zero wavelets)
template <int N1> __device__ inline void load_operand(double *dst, const double *pool, long long off) {
    constexpr int N3 = N1 * N1 * N1;
    const bool gen = off < 0;
    const long long base = gen ? -off - 1 : off;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[pad<N1>(o)] = (gen && o >= N3) ? 0.0 : pool[base + o];
}

template <int N1>
__global__ __launch_bounds__(kThreads) void k_combine(mwg_combine cb, const long long *in_off, const int *scale, const long long *out_off,
                                                      double *out, double *norms) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1, PT = (8 * N3 + kThreads - 1) / kThreads;
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ double cvm[N2], vcm[N2];
    __shared__ double sfl[8 * N2];
    const int k = blockIdx.x;
    const long long *offs = in_off + (size_t)k * cb.nin;
    double *dst = out + out_off[k];
    if (cb.mode == 0) {
        // coefs_o = 0; coefs_o += c_i coefs_i, operand by operand
        for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
            double s = 0.0;
            for (int i = 0; i < cb.nin; i++) {
                const long long off = offs[i];
                const bool gen = off < 0;
                const double v = gen ? (o < N3 ? cb.pool[i][-off - 1 + o] : 0.0) : cb.pool[i][off + o];
                s += cb.c[i] * v;
            }
            a[pad<N1>(o)] = s;
            dst[o] = s;
        }
        __syncthreads();
        block_norms<N1, true>(a, norms + (size_t)k * 9, red);
        return;
    }
    for (int i = threadIdx.x; i < N2; i += kThreads) {
        cvm[i] = cb.cvmap[i];
        vcm[i] = cb.vcmap[i];
    }
    stage_filters<N1>(sfl, cb.filters, 8);
    const int n = scale[k];
    double sfp = cb.sf[0] * cb.sf[1] * cb.sf[2];
    if (sfp <= kMachineZero) sfp = 1.0;
    const double two_fac = ldexp(1.0, 3 * (n + 1)) / sfp;
    const double fwd = sqrt(two_fac), bwd = sqrt(1.0 / two_fac);
    double val[PT];
#pragma unroll
    for (int p = 0; p < PT; p++) val[p] = 1.0;
    for (int i = 0; i < cb.nin; i++) {
        __syncthreads(); // the previous operand's values are consumed
        load_operand<N1>(a, cb.pool[i], offs[i]);
        __syncthreads();
        mw_pass<N1>(b, a, sfl, 4, 0);
        __syncthreads();
        mw_pass<N1>(a, b, sfl, 4, 1);
        __syncthreads();
        mw_pass<N1>(b, a, sfl, 4, 2);
        __syncthreads();
        map_pass<N1>(a, b, cvm);
        __syncthreads();
        map_pass<N1>(b, a, cvm);
        __syncthreads();
        map_pass<N1>(a, b, cvm);
        __syncthreads();
        const double ci = cb.c[i];
#pragma unroll
        for (int p = 0; p < PT; p++) {
            const int o = threadIdx.x + kThreads * p;
            if (o < 8 * N3) val[p] = val[p] * (ci * (fwd * a[pad<N1>(o)]));
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int o = threadIdx.x + kThreads * p;
        if (o < 8 * N3) a[pad<N1>(o)] = val[p];
    }
    __syncthreads();
    // values -> scaling coefficients of the children, then compression
    map_pass<N1>(b, a, vcm);
    __syncthreads();
    map_pass<N1>(a, b, vcm);
    __syncthreads();
    map_pass<N1>(b, a, vcm);
    __syncthreads();
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) b[pad<N1>(o)] = bwd * b[pad<N1>(o)];
    __syncthreads();
    mw_pass<N1>(a, b, sfl, 0, 0);
    __syncthreads();
    mw_pass<N1>(b, a, sfl, 0, 1);
    __syncthreads();
    mw_pass<N1>(a, b, sfl, 0, 2);
    __syncthreads();
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[o] = a[pad<N1>(o)];
    block_norms<N1, true>(a, norms + (size_t)k * 9, red);
}

template <int N1> struct CombineL {
    static mwg_status run(const mwg_combine *cb, const long long *in_off, const int *scale, const long long *out_off, double *out, double *norms,
                          int count) {
        ProfileScope prof_(MWG_KERNEL_ARITH);
        hipLaunchKernelGGL(k_combine<N1>, dim3(count), dim3(kThreads), 0, 0, *cb, in_off, scale, out_off, out, norms);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};

// One workgroup per node pair: the scaling and wavelet parts of the pair's
// inner product, each thread over a fixed stride, then a fixed-order
// reduction over the workgroup.
template <int N1>
__global__ __launch_bounds__(kThreads) void k_dot_nodes(const double *a, const long long *a_off, const double *b, const long long *b_off,
                                                        const int *root, double *partial) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double red[2 * kThreads];
    const int k = blockIdx.x;
    const double *x = a + a_off[k], *y = b + b_off[k];
    double s = 0.0, w = 0.0;
    if (root[k])
        for (int o = threadIdx.x; o < N3; o += kThreads) s += x[o] * y[o];
    for (int o = N3 + threadIdx.x; o < 8 * N3; o += kThreads) w += x[o] * y[o];
    red[threadIdx.x] = s;
    red[kThreads + threadIdx.x] = w;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[threadIdx.x] += red[threadIdx.x + h];
            red[kThreads + threadIdx.x] += red[kThreads + threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * k] = red[0];
        partial[2 * k + 1] = red[kThreads];
    }
}

template <int N1> struct DotL {
    static mwg_status run(const double *a, const long long *a_off, const double *b, const long long *b_off, const int *root, double *partial,
                          int count) {
        ProfileScope prof_(MWG_KERNEL_ARITH);
        hipLaunchKernelGGL(k_dot_nodes<N1>, dim3(count), dim3(kThreads), 0, 0, a, a_off, b, b_off, root, partial);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_dot_nodes(int kp1, const double *a, const long long *a_off, const double *b, const long long *b_off, const int *root,
                                    double *partial, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count < 0) return fail(MWG_ERR_ARG, "negative node count");
    if (count == 0) return MWG_OK;
    if (!a || !a_off || !b || !b_off || !root || !partial) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<DotL>(kp1, a, a_off, b, b_off, root, partial, count);
}

extern "C" mwg_status mwg_combine_nodes(const mwg_combine *cb, const long long *in_off, const int *scale, const long long *out_off, double *out,
                                        double *norms, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count < 0) return fail(MWG_ERR_ARG, "negative node count");
    if (count == 0) return MWG_OK;
    if (!cb || !in_off || !scale || !out_off || !out || !norms) return fail(MWG_ERR_ARG, "null argument");
    if (cb->mode < 0 || cb->mode > 1) return fail(MWG_ERR_ARG, "combine mode: 0 sum, 1 product");
    if (cb->nin < 1 || cb->nin > MWG_MAX_OPERANDS) return fail(MWG_ERR_ARG, "operand count outside 1..8");
    for (int i = 0; i < cb->nin; i++)
        if (!cb->pool[i]) return fail(MWG_ERR_ARG, "null operand pool");
    if (!cb->filters || (cb->mode == 1 && (!cb->vcmap || !cb->cvmap))) return fail(MWG_ERR_ARG, "incomplete combine description");
    return dispatch<CombineL>(cb->kp1, cb, in_off, scale, out_off, out, norms, count);
}
