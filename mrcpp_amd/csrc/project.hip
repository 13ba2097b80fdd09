// Projection of a sum of Gaussians onto the nodes of one tree level
// (ProjectionCalculator::calcNode, src/treebuilders/ProjectionCalculator.cpp:34-51).
//
// One 256-thread workgroup per node.  The function is sampled at the
// Gauss-Legendre points of the node's 8 children (MWNode::getExpandedChildPts,
// src/trees/MWNode.cpp:854-925): point (t, i0, i1, i2) sits at
// sf_d 2^-(n+1) (x_{i_d} + 2 l_d + bit_d(t)).  A Gaussian is separable, so per
// node and term the 3 x 2(k+1) one-dimensional factors exp(-alpha_d q^2) are
// formed once in LDS and every point takes coef * e_x e_y e_z: 6(k+1) exp
// evaluations per term instead of 8(k+1)^3.  Then the value -> coefficient
// map in each direction (MWNode::cvTransform, :448-490), the scaling factor
// sqrt(sf_x sf_y sf_z / 2^(3(n+1))), the compression to the node's scaling and
// wavelet blocks (MWNode::mwTransform(Compression)) and the component norms.
#include "device_common.h"

namespace mwgdev {

namespace {

constexpr int kTermsPerRound = 16; // Gaussian terms whose 1D factors share LDS at a time (LDS <= 80 KB at k = 7: two workgroups per CU)

// out(r, j) = sum_i in(i, r) V(i, j) for each of the 8 blocks (in viewed
// N1 x N1^2, column-major; core::filter_pass), padded LDS layout.
template <int N1> __device__ inline void vc_pass(double *out, const double *in, const double *V) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const int t = o / N3, e = o % N3;
        const int r = e % N2, j = e / N2;
        const double *x = in + pad<N1>(t * N3 + N1 * r);
        const double *v = V + N1 * j;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N1; i++) s += x[i] * v[i];
        out[pad<N1>(o)] = s;
    }
}

template <int N1>
__global__ __launch_bounds__(kThreads) void k_project(mwg_projection pj, const int *scale, const int *lv, const long long *off, double *coef,
                                                      double *norms) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1, PT = (8 * N3 + kThreads - 1) / kThreads;
    __shared__ double a[Lds<N1>::size], b[Lds<N1>::size];
    __shared__ double fac1d[kTermsPerRound][3][2 * N1]; // exp(-alpha_d q^2) at the 2 N1 points of each direction
    __shared__ double gcoef[kTermsPerRound];
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ double sx[N1], svc[N2];
    const int k = blockIdx.x;
    const int n = scale[k];
    const int l[3] = {lv[3 * k], lv[3 * k + 1], lv[3 * k + 2]};
    const double h = ldexp(1.0, -(n + 1));
    for (int i = threadIdx.x; i < N1; i += kThreads) sx[i] = pj.nodes[i];
    for (int i = threadIdx.x; i < N2; i += kThreads) svc[i] = pj.vcmap[i];
    __syncthreads();

    // function values: each thread owns points o = tid + 256 p
    double val[PT];
#pragma unroll
    for (int p = 0; p < PT; p++) val[p] = 0.0;
    for (int g0 = 0; g0 < pj.ngauss; g0 += kTermsPerRound) {
        const int ng = min(kTermsPerRound, pj.ngauss - g0);
        __syncthreads(); // previous round's factors consumed
        for (int q = threadIdx.x; q < ng * 3 * 2 * N1; q += kThreads) {
            const int gi = q / (6 * N1), rest = q % (6 * N1), d = rest / (2 * N1), hi = rest % (2 * N1);
            const int half = hi / N1, i = hi % N1;
            const mwg_gauss &G = pj.gauss[g0 + gi];
            const double base = sx[i] + 2.0 * (double)l[d];
            const double r = pj.sf[d] * (h * (half ? base + 1.0 : base));
            const double dq = r - G.pos[d];
            fac1d[gi][d][hi] = exp(-(G.alpha[d] * dq * dq));
        }
        for (int gi = threadIdx.x; gi < ng; gi += kThreads) gcoef[gi] = pj.gauss[g0 + gi].coef;
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PT; p++) {
            const int o = threadIdx.x + kThreads * p;
            if (o >= 8 * N3) break;
            const int t = o / N3, e = o % N3;
            const int ix = (t & 1) * N1 + e % N1, iy = ((t >> 1) & 1) * N1 + (e / N1) % N1, iz = ((t >> 2) & 1) * N1 + e / N2;
            double v = val[p];
            for (int gi = 0; gi < ng; gi++) v += gcoef[gi] * (fac1d[gi][0][ix] * fac1d[gi][1][iy] * fac1d[gi][2][iz]);
            val[p] = v;
        }
    }
#pragma unroll
    for (int p = 0; p < PT; p++) {
        const int o = threadIdx.x + kThreads * p;
        if (o < 8 * N3) a[pad<N1>(o)] = val[p];
    }
    __syncthreads();
    // values -> coefficients, one direction per pass
    vc_pass<N1>(b, a, svc);
    __syncthreads();
    vc_pass<N1>(a, b, svc);
    __syncthreads();
    vc_pass<N1>(b, a, svc);
    __syncthreads();
    double sfp = pj.sf[0] * pj.sf[1] * pj.sf[2];
    if (sfp <= kMachineZero) sfp = 1.0;
    const double fac = sqrt(1.0 / (ldexp(1.0, 3 * (n + 1)) / sfp));
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) b[pad<N1>(o)] *= fac;
    __syncthreads();
    // children's scaling blocks -> the node's scaling and wavelet blocks
    mw_pass<N1>(a, b, pj.filters, 0, 0); // filters read through the cache
    __syncthreads();
    mw_pass<N1>(b, a, pj.filters, 0, 1);
    __syncthreads();
    mw_pass<N1>(a, b, pj.filters, 0, 2);
    __syncthreads();
    double *dst = coef + off[k];
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[o] = a[pad<N1>(o)];
    block_norms<N1, true>(a, norms + (size_t)k * 9, red);
}

template <int N1> struct ProjectL {
    static mwg_status run(const mwg_projection *pj, const int *scale, const int *l, const long long *off, double *coef, double *norms, int count) {
        ProfileScope prof_(MWG_KERNEL_PROJECT);
        hipLaunchKernelGGL(k_project<N1>, dim3(count), dim3(kThreads), 0, 0, *pj, scale, l, off, coef, norms);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};

} // namespace

} // namespace mwgdev

using namespace mwgdev;

extern "C" mwg_status mwg_project_nodes(const mwg_projection *pj, const int *scale, const int *l, const long long *off, double *coef, double *norms,
                                        int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count < 0) return fail(MWG_ERR_ARG, "negative node count");
    if (count == 0) return MWG_OK;
    if (!pj || !scale || !l || !off || !coef || !norms) return fail(MWG_ERR_ARG, "null argument");
    if (!pj->nodes || !pj->vcmap || !pj->filters || (pj->ngauss > 0 && !pj->gauss) || pj->ngauss < 0)
        return fail(MWG_ERR_ARG, "incomplete projection description");
    return dispatch<ProjectL>(pj->kp1, pj, scale, l, off, coef, norms, count);
}
