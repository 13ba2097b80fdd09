// gfx950 kernels for MRCPP's operator application and the mwg_* C ABI.
//
// Launch shapes: one 256-thread workgroup (4 wave64) per output node, per
// candidate (output node, band node) pair or per parent node; coefficient
// blocks of (k+1)^3 doubles are staged in LDS and every thread owns a fixed
// set of output coefficients across the whole pair.  Layout of a node: 8
// blocks (block t = wavelet bits t, x->bit 0), element i0 + (k+1) i1 + (k+1)^2 i2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mwg.h"
#include "keys.h"

namespace {

thread_local std::string g_err;
bool g_ready = false;

mwg_status fail(mwg_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define MWG_HIP(call)                                                                                                   \
    do {                                                                                                                \
        hipError_t e_ = (call);                                                                                         \
        if (e_ != hipSuccess) return fail(MWG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define MWG_LAUNCHED()                                                                                                  \
    do {                                                                                                                \
        hipError_t e_ = hipGetLastError();                                                                              \
        if (e_ != hipSuccess) return fail(MWG_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_));     \
    } while (0)

// Optional per-kernel timing with HIP events on the launch stream.
struct TimedLaunch {
    int kind;
    hipEvent_t start, stop;
};
bool g_profile = false;
std::vector<TimedLaunch> &timed_launches() {
    static std::vector<TimedLaunch> v;
    return v;
}
struct ProfileScope {
    int kind;
    hipEvent_t start{}, stop{};
    explicit ProfileScope(int k)
            : kind(k) {
        if (!g_profile) return;
        (void)hipEventCreate(&start);
        (void)hipEventCreate(&stop);
        (void)hipEventRecord(start, 0);
    }
    ~ProfileScope() {
        if (!g_profile) return;
        (void)hipEventRecord(stop, 0);
        timed_launches().push_back({kind, start, stop});
    }
};

constexpr int kThreads = 256;
constexpr int kMaxTerms = 512;
constexpr double kMachineZero = 1.0e-14;
// relative slack of the separable pair bound against rounding
constexpr double kBoundSlack = 1.0 + 1.0e-9;

__device__ inline int f_lookup(const mwg_ftree &f, int depth, int x, int y, int z) {
    const int ox = x - f.corner[0] * (1 << depth);
    const int oy = y - f.corner[1] * (1 << depth);
    const int oz = z - f.corner[2] * (1 << depth);
    const uint64_t key = mwgkey::pack(depth, ox, oy, oz);
    const int log2cap = 31 - __clz(f.hash_cap);
    const uint32_t mask = (uint32_t)f.hash_cap - 1;
    uint32_t s = mwgkey::slot(key, log2cap);
    for (;;) {
        const unsigned long long k = f.hash_keys[s];
        if (k == key) return f.hash_vals[s];
        if (k == 0ull) return -1;
        s = (s + 1) & mask;
    }
}

// Block-wide exclusive scan of one int per thread; returns the total.
__device__ inline int block_exclusive_scan(int v, int *tmp, int &excl) {
    const int tid = threadIdx.x;
    tmp[tid] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        int add = (tid >= off) ? tmp[tid - off] : 0;
        __syncthreads();
        tmp[tid] += add;
        __syncthreads();
    }
    const int incl = tmp[tid];
    const int total = tmp[kThreads - 1];
    excl = incl - v;
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------
// Pair screen: one workgroup per output node of the level.
// ---------------------------------------------------------------------------
__device__ inline bool screen_position(const mwg_level &lvl, const mwg_ftree &f, const mwg_band_tables &op, const int *gl, int p,
                                       const int *s, const int *n, int *out_f, int *out_l) {
    const int d = lvl.o_depth;
    const int x = s[0] + p % n[0];
    const int y = s[1] + (p / n[0]) % n[1];
    const int z = s[2] + p / (n[0] * n[1]);
    const double *pb = op.pair_bound + (size_t)d * (2 * op.wmax + 1) + op.wmax;
    double a = pb[x - gl[0]] * pb[y - gl[1]] * pb[z - gl[2]];
    if (d > 0) {
        // below the root scale every tuple has a wavelet component (C, B or A)
        // in at least one direction
        const double *pw = op.pair_bound_w + (size_t)d * (2 * op.wmax + 1) + op.wmax;
        const double mx = pb[x - gl[0]], my = pb[y - gl[1]], mz = pb[z - gl[2]];
        a = fmax(pw[x - gl[0]] * my * mz, fmax(mx * pw[y - gl[1]] * mz, mx * my * pw[z - gl[2]]));
    }
    // a zero bound cannot pass a non-negative threshold: skip the lookup
    if (a == 0.0 && lvl.g_thr >= 0.0) return false;
    int id = f_lookup(f, d, x, y, z);
    double maxf;
    double fnv[8]; // per-component norms of the band node (or their bound)
    int code;
    if (id >= 0) {
        const double *nr = f.norms + (size_t)id * 9;
        maxf = 0.0;
        for (int t = 0; t < 8; t++) {
            fnv[t] = nr[t];
            maxf = fmax(maxf, nr[t]);
        }
        code = id;
    } else {
        int dd = d, xx = x, yy = y, zz = z, anc = -1;
        while (anc < 0 && dd > 0) {
            dd--;
            xx >>= 1;
            yy >>= 1;
            zz >>= 1;
            anc = f_lookup(f, dd, xx, yy, zz);
        }
        if (anc < 0) return false; // outside the world (not reachable for clamped boxes)
        const double *nr = f.norms + (size_t)anc * 9;
        // a generated node holds a scaling block only, whose norm is bounded
        // by its nearest existing ancestor's (orthogonal reconstruction)
        maxf = f.is_gen[anc] ? nr[0] : sqrt(nr[8]);
        fnv[0] = maxf;
        for (int t = 1; t < 8; t++) fnv[t] = 0.0;
        code = -(anc + 1);
    }
    if (!(a * maxf * kBoundSlack > lvl.g_thr)) return false;
    // Per-term bound, resolved per f component ft: the direction-d factor of
    // a tuple is the norm of component 2*gbit_d + ftbit_d, so over all gt it is
    // at most max(n_d(ftbit_d), n_d(2+ftbit_d)); at depth > 0 the (ft,gt) =
    // (0,0) tuple is skipped, so some direction carries the gt bit.
    const int M = op.M;
    const int dl3[3] = {x - gl[0], y - gl[1], z - gl[2]};
    const int maxdl = max(abs(dl3[0]), max(abs(dl3[1]), abs(dl3[2])));
    const int *wd = op.widths + (size_t)d * M * 5;
    const long long *ms = op.mat_start + (size_t)d * M;
    const int *bsm = op.band_size_max + (size_t)d * M;
    bool any = false;
    for (int i = 0; i < M && !any; i++) {
        const int *wi = wd + i * 5;
        const int w = wi[4];
        if (maxdl > w) continue;
        const double *mn = op.mat_norm + ms[i];
        double n[3][4];
        for (int dim = 0; dim < 3; dim++)
            for (int c = 0; c < 4; c++) n[dim][c] = (abs(dl3[dim]) <= wi[c]) ? mn[(dl3[dim] + w) * 4 + c] : 0.0;
        const double bs = (double)bsm[i];
        for (int ft = 0; ft < 8 && !any; ft++) {
            if (fnv[ft] < kMachineZero) continue;
            double g0[3], g1[3], gm[3];
            for (int dim = 0; dim < 3; dim++) {
                const int fb = (ft >> dim) & 1;
                g0[dim] = n[dim][fb];
                g1[dim] = n[dim][2 + fb];
                gm[dim] = fmax(g0[dim], g1[dim]);
            }
            double prod;
            if (d > 0 && ft == 0)
                prod = fmax(g1[0] * gm[1] * gm[2], fmax(gm[0] * g1[1] * gm[2], gm[0] * gm[1] * g1[2]));
            else
                prod = gm[0] * gm[1] * gm[2];
            any = bs * prod * fnv[ft] * kBoundSlack > lvl.g_thr;
        }
    }
    if (!any) return false;
    *out_f = code;
    out_l[0] = x;
    out_l[1] = y;
    out_l[2] = z;
    return true;
}

__global__ __launch_bounds__(kThreads) void k_pair_screen(mwg_level lvl, mwg_ftree f, mwg_band_tables op, mwg_pairs out) {
    __shared__ int tmp[kThreads];
    __shared__ int base;
    const int g = blockIdx.x;
    const int d = lvl.o_depth;
    const int W = op.band_max[d];
    int gl[3], s[3], n[3];
    for (int i = 0; i < 3; i++) {
        gl[i] = lvl.l[3 * g + i];
        const int c = f.corner[i] * (1 << d);
        const int nb = f.nbox[i] * (1 << d);
        s[i] = max(gl[i] - W, c);
        const int e = min(gl[i] + W, c + nb - 1);
        n[i] = max(e - s[i] + 1, 0);
    }
    const int P = n[0] * n[1] * n[2];
    int cnt = 0, fid, fl[3];
    for (int p = threadIdx.x; p < P; p += kThreads)
        if (screen_position(lvl, f, op, gl, p, s, n, &fid, fl)) cnt++;
    int excl;
    const int total = block_exclusive_scan(cnt, tmp, excl);
    if (threadIdx.x == 0) {
        base = atomicAdd(out.count, total);
        out.g_first[g] = base;
        out.g_npairs[g] = total;
    }
    __syncthreads();
    int w = base + excl;
    for (int p = threadIdx.x; p < P; p += kThreads)
        if (screen_position(lvl, f, op, gl, p, s, n, &fid, fl)) {
            if (w < out.capacity) {
                out.pair_g[w] = g;
                out.pair_f[w] = fid;
                out.pair_l[3 * w + 0] = fl[0];
                out.pair_l[3 * w + 1] = fl[1];
                out.pair_l[3 * w + 2] = fl[2];
            }
            w++;
        }
}

// ---------------------------------------------------------------------------
// One separable 3D multiwavelet pass set over 8 blocks held in LDS.
// dir_base 0: compression sub-filters, 4: reconstruction sub-filters.
// pass dim: out[gt] = sum over the two ft sharing gt's other bits, ascending
// ft, of in[ft]^T F(2*bit(gt)+bit(ft)), in viewed (N1 x N2) column-major.
// ---------------------------------------------------------------------------
template <int N1> __device__ inline void mw_pass(double *out, const double *in, const double *filt, int dir_base, int dim) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    const int bit = 1 << dim;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const int gt = o / N3, e = o % N3;
        const int r = e % N2, j = e / N2;
        const int ft0 = gt & ~bit, ft1 = gt | bit;
        const double *F0 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 0) * N2 + N1 * j;
        const double *F1 = filt + (size_t)(dir_base + 2 * ((gt >> dim) & 1) + 1) * N2 + N1 * j;
        const double *x0 = in + (size_t)ft0 * N3 + N1 * r;
        const double *x1 = in + (size_t)ft1 * N3 + N1 * r;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int i = 0; i < N1; i++) {
            s0 += x0[i] * F0[i];
            s1 += x1[i] * F1[i];
        }
        out[o] = s0 + s1;
    }
}

// Last reconstruction pass accumulating into the children, each of the two
// contributions added separately (apply_filter with accumulate).
template <int N1>
__device__ inline void mw_pass_add_children(double *coef, long long child_off, long long stride, const double *in, const double *filt) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    const int dim = 2, bit = 4;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const int gt = o / N3, e = o % N3;
        const int r = e % N2, j = e / N2;
        const int ft0 = gt & ~bit, ft1 = gt | bit;
        const double *F0 = filt + (size_t)(4 + 2 * ((gt >> dim) & 1) + 0) * N2 + N1 * j;
        const double *F1 = filt + (size_t)(4 + 2 * ((gt >> dim) & 1) + 1) * N2 + N1 * j;
        const double *x0 = in + (size_t)ft0 * N3 + N1 * r;
        const double *x1 = in + (size_t)ft1 * N3 + N1 * r;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int i = 0; i < N1; i++) {
            s0 += x0[i] * F0[i];
            s1 += x1[i] * F1[i];
        }
        double *dst = coef + child_off + (long long)gt * stride + e;
        double v = *dst;
        v += s0;
        v += s1;
        *dst = v;
    }
}

// Component norms of 8 blocks (LDS or global), written to nrm[0..8].
template <int N1> __device__ inline void block_norms(const double *c, double *nrm, double *red) {
    constexpr int N3 = N1 * N1 * N1;
    double ss[8];
#pragma unroll
    for (int t = 0; t < 8; t++) ss[t] = 0.0;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const double v = c[o];
        ss[o / N3] += v * v;
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        double v = ss[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        ss[t] = v;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int t = 0; t < 8; t++) red[wave * 8 + t] = ss[t];
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int t = 0; t < 8; t++) {
            double v = 0.0;
            for (int w = 0; w < kThreads / 64; w++) v += red[w * 8 + t];
            const double nv = sqrt(v);
            nrm[t] = nv;
            tot += nv * nv;
        }
        nrm[8] = tot;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Generated band nodes: scaling block of one child of an existing node.
// ---------------------------------------------------------------------------
template <int N1>
__global__ __launch_bounds__(kThreads) void k_generate(mwg_ftree f, const double *filt, const int *ids, const int *parent, const int *child) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double a[8 * N3], b[8 * N3];
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ double nrm[9];
    const int k = blockIdx.x;
    const int pid = parent[k], id = ids[k], c = child[k];
    const double *src = f.coef + f.coef_off[pid];
    const int nblk = f.is_gen[pid] ? 1 : 8; // generated parents carry the scaling block only
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) a[o] = (o < nblk * N3) ? src[o] : 0.0;
    __syncthreads();
    mw_pass<N1>(b, a, filt, 4, 0);
    __syncthreads();
    mw_pass<N1>(a, b, filt, 4, 1);
    __syncthreads();
    mw_pass<N1>(b, a, filt, 4, 2);
    __syncthreads();
    double *dst = f.coef + f.coef_off[id];
    double ss = 0.0;
    for (int e = threadIdx.x; e < N3; e += kThreads) {
        const double v = b[c * N3 + e];
        dst[e] = v;
        ss += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
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

// ---------------------------------------------------------------------------
// Band sum for one candidate pair: exact screening of every (ft, term, gt)
// and the three 1D sweeps, x and y sweeps shared between the gt that need
// the same intermediate.
// ---------------------------------------------------------------------------
template <int N1>
__global__ __launch_bounds__(kThreads) void k_band_apply(mwg_level lvl, mwg_ftree f, mwg_band_tables op, const int *pair_g, const int *pair_f,
                                                        const int *pair_l, int *pair_work, double *partial, unsigned long long *n_applied) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    constexpr int R = (8 * N3 + kThreads - 1) / kThreads;
    __shared__ double sF[N3];
    __shared__ double sT1[2 * N3];
    __shared__ double sT2[4 * N3];
    __shared__ double sO[6 * N2];
    __shared__ unsigned char smask[8 * kMaxTerms];
    __shared__ int sgroup[8 * kMaxTerms];
    __shared__ int tmp[kThreads];
    __shared__ int ngroups;
    __shared__ unsigned int napp;

    const int p = blockIdx.x;
    const int g = pair_g[p];
    const int fid = pair_f[p];
    const int d = lvl.o_depth;
    const int M = op.M;
    int dl[3], maxdl = 0;
    for (int i = 0; i < 3; i++) {
        dl[i] = pair_l[3 * p + i] - lvl.l[3 * g + i];
        maxdl = max(maxdl, abs(dl[i]));
    }
    const double *fn = f.norms + (size_t)fid * 9;
    const int *wd = op.widths + (size_t)d * M * 5;
    const int *bs = op.band_size + (size_t)d * M * 64;
    const long long *ms = op.mat_start + (size_t)d * M;
    if (threadIdx.x == 0) napp = 0;
    __syncthreads();

    // --- exact screening: item q = ft * M + i -------------------------------
    unsigned int local_app = 0;
    for (int q = threadIdx.x; q < 8 * M; q += kThreads) {
        const int ft = q / M, i = q - ft * M;
        unsigned int m = 0;
        const double fnorm = fn[ft];
        const int *w = wd + i * 5;
        if (!(fnorm < kMachineZero) && maxdl <= w[4]) {
            for (int gt = 0; gt < 8; gt++) {
                if (!(d == 0 || gt != 0 || ft != 0)) continue;
                double onorm = 1.0;
                bool inside = true;
                for (int dim = 0; dim < 3; dim++) {
                    const int c = 2 * ((gt >> dim) & 1) + ((ft >> dim) & 1);
                    if (abs(dl[dim]) > w[c]) {
                        inside = false;
                        break;
                    }
                    onorm *= op.mat_norm[ms[i] + (long long)(dl[dim] + w[4]) * 4 + c];
                }
                if (!inside) continue;
                const double fthr = bs[i * 64 + gt * 8 + ft] * fnorm;
                if (onorm * fthr > lvl.g_thr) m |= 1u << gt;
            }
        }
        smask[q] = (unsigned char)m;
        local_app += __popc(m);
    }
    atomicAdd(&napp, local_app);
    __syncthreads();
    // compact groups in q order (ft-major), contiguous chunks per thread
    const int nq = 8 * M;
    const int chunk = (nq + kThreads - 1) / kThreads;
    const int q0 = threadIdx.x * chunk, q1 = min(nq, q0 + chunk);
    int cnt = 0;
    for (int q = q0; q < q1; q++) cnt += smask[q] != 0;
    int excl;
    const int total = block_exclusive_scan(cnt, tmp, excl);
    for (int q = q0; q < q1; q++)
        if (smask[q]) sgroup[excl++] = q;
    if (threadIdx.x == 0) {
        ngroups = total;
        if (napp) atomicAdd(n_applied, (unsigned long long)napp);
    }
    __syncthreads();

    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0;

    const double *fsrc = f.coef + f.coef_off[fid];
    int curft = -1;
    const int G = ngroups;
    for (int k = 0; k < G; k++) {
        const int q = sgroup[k];
        const int ft = q / M, i = q - ft * M;
        const unsigned int m = smask[q];
        if (ft != curft) {
            for (int e = threadIdx.x; e < N3; e += kThreads) sF[e] = fsrc[(size_t)ft * N3 + e];
            curft = ft;
        }
        // operator blocks: x (gx=0,1), y (gy=0,1), z (gz=0,1)
        const int w4 = wd[i * 5 + 4];
        for (int e = threadIdx.x; e < 6 * N2; e += kThreads) {
            const int which = e / N2, el = e % N2;
            const int dim = which >> 1, gb = which & 1;
            const int c = 2 * gb + ((ft >> dim) & 1);
            sO[e] = op.mats[(size_t)(ms[i] + (long long)(dl[dim] + w4) * 4 + c) * N2 + el];
        }
        __syncthreads();
        // x sweep: T1[gx](r + N2 j0) = sum_i0 F(i0 + N1 r) Ox[gx](i0, j0)
        unsigned int xneed = 0, yneed = 0;
        for (int gt = 0; gt < 8; gt++)
            if (m >> gt & 1) {
                xneed |= 1u << (gt & 1);
                yneed |= 1u << (gt & 3);
            }
        for (int e = threadIdx.x; e < 2 * N3; e += kThreads) {
            const int gx = e / N3;
            if (!(xneed >> gx & 1)) continue;
            const int o = e % N3, r = o % N2, j = o / N2;
            const double *x = sF + N1 * r;
            const double *O = sO + gx * N2 + N1 * j;
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < N1; a++) s += x[a] * O[a];
            sT1[e] = s;
        }
        __syncthreads();
        // y sweep: T2[gx + 2 gy](r + N2 j1) = sum_i1 T1[gx](i1 + N1 r) Oy[gy](i1, j1)
        for (int e = threadIdx.x; e < 4 * N3; e += kThreads) {
            const int cb = e / N3;
            if (!(yneed >> cb & 1)) continue;
            const int gx = cb & 1, gy = cb >> 1;
            const int o = e % N3, r = o % N2, j = o / N2;
            const double *x = sT1 + gx * N3 + N1 * r;
            const double *O = sO + (2 + gy) * N2 + N1 * j;
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < N1; a++) s += x[a] * O[a];
            sT2[e] = s;
        }
        __syncthreads();
        // z sweep into the owned outputs: G[gt](r + N2 j2) += sum_i2 T2[gt&3](i2 + N1 r) Oz[gz](i2, j2)
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const int o = threadIdx.x + rr * kThreads;
            if (o < 8 * N3) {
                const int gt = o / N3;
                if (m >> gt & 1) {
                    const int e = o % N3, r = e % N2, j = e / N2;
                    const double *x = sT2 + (gt & 3) * N3 + N1 * r;
                    const double *O = sO + (4 + (gt >> 2)) * N2 + N1 * j;
                    double s = 0.0;
#pragma unroll
                    for (int a = 0; a < N1; a++) s += x[a] * O[a];
                    acc[rr] += s;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) pair_work[p] = G > 0;
    if (G == 0) return; // nothing applied: the reduction skips this pair
    double *dst = partial + (size_t)p * 8 * N3;
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
        const int o = threadIdx.x + rr * kThreads;
        if (o < 8 * N3) dst[o] = acc[rr];
    }
}

// ---------------------------------------------------------------------------
// Band sum, one wave64 per candidate pair, for (k+1)^2 <= 64.
//
// Lane t screens term c0 + t for every (ft, gt) and keeps one gt mask per ft;
// groups (ft, term) are then taken in (term chunk, ft, term) order through a
// wave ballot, so the whole wave works on one group at a time with uniform
// control flow and wave-uniform operator blocks.  In every sweep a lane owns
// one line of the block: it reads N1 values, does N1^2 FMAs and writes N1
// values.  LDS arrays are [contracted index][line] with a row pitch of 68
// doubles (64 + 4), so the writes of one sweep (stride 68 across one lane
// index, 1 across the other) and the reads of the next (stride 1) hit 32
// distinct double banks in each half wave.
// ---------------------------------------------------------------------------
template <int N1>
__global__ __launch_bounds__(64) void k_band_apply_wave(mwg_level lvl, mwg_ftree f, mwg_band_tables op, const int *pair_g, const int *pair_f,
                                                       const int *pair_l, int *pair_work, double *partial, unsigned long long *n_applied) {
    constexpr int N2 = N1 * N1, N3 = N2 * N1;
    constexpr int P = 68;
    static_assert(N2 <= 64, "one line per lane");
    __shared__ double sF[N1 * P];
    __shared__ double sT1[N1 * P];
    __shared__ double sT2[N1 * P];

    const int lane = threadIdx.x;
    const int p = blockIdx.x;
    const int g = pair_g[p];
    const int fid = pair_f[p];
    const int d = lvl.o_depth;
    const int M = op.M;
    int dl[3], maxdl = 0;
    for (int i = 0; i < 3; i++) {
        dl[i] = pair_l[3 * p + i] - lvl.l[3 * g + i];
        maxdl = max(maxdl, abs(dl[i]));
    }
    double fn[8];
    for (int t = 0; t < 8; t++) fn[t] = f.norms[(size_t)fid * 9 + t];
    const double *fsrc = f.coef + f.coef_off[fid];
    const int *wd = op.widths + (size_t)d * M * 5;
    const int *bsz = op.band_size + (size_t)d * M * 64;
    const long long *ms = op.mat_start + (size_t)d * M;

    double acc[8][N1];
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int j = 0; j < N1; j++) acc[t][j] = 0.0;

    unsigned int napp = 0;
    bool any_work = false;
    int curft = -1;
    for (int c0 = 0; c0 < M; c0 += 64) {
        // ---- exact screening of term c0 + lane (applyOperComp / applyOperator) ----
        const int i = c0 + lane;
        unsigned int m8[8];
#pragma unroll
        for (int ft = 0; ft < 8; ft++) m8[ft] = 0;
        if (i < M && maxdl <= wd[i * 5 + 4]) {
            const int *w = wd + i * 5;
            const int w4 = w[4];
            double nrm[3][4];
#pragma unroll
            for (int dim = 0; dim < 3; dim++)
#pragma unroll
                for (int c = 0; c < 4; c++)
                    nrm[dim][c] = (abs(dl[dim]) <= w[c]) ? op.mat_norm[ms[i] + (long long)(dl[dim] + w4) * 4 + c] : -1.0;
#pragma unroll
            for (int ft = 0; ft < 8; ft++) {
                if (fn[ft] < kMachineZero) continue;
                unsigned int m = 0;
#pragma unroll
                for (int gt = 0; gt < 8; gt++) {
                    if (!(d == 0 || gt != 0 || ft != 0)) continue;
                    const double nx = nrm[0][2 * (gt & 1) + (ft & 1)];
                    const double ny = nrm[1][2 * ((gt >> 1) & 1) + ((ft >> 1) & 1)];
                    const double nz = nrm[2][2 * ((gt >> 2) & 1) + ((ft >> 2) & 1)];
                    if (nx < 0.0 || ny < 0.0 || nz < 0.0) continue;
                    double onorm = 1.0;
                    onorm *= nx;
                    onorm *= ny;
                    onorm *= nz;
                    const double fthr = bsz[i * 64 + gt * 8 + ft] * fn[ft];
                    if (onorm * fthr > lvl.g_thr) m |= 1u << gt;
                }
                m8[ft] = m;
                napp += __popc(m);
            }
        }
        // gt masks packed one byte per ft, so no register array is indexed at run time
        const unsigned int mlo = m8[0] | (m8[1] << 8) | (m8[2] << 16) | (m8[3] << 24);
        const unsigned int mhi = m8[4] | (m8[5] << 8) | (m8[6] << 16) | (m8[7] << 24);
        // ---- groups of this term chunk, ft-major ----
        for (int ft = 0; ft < 8; ft++) {
            const unsigned int mine = ((ft < 4 ? mlo : mhi) >> (8 * (ft & 3))) & 0xffu;
            unsigned long long bal = __ballot(mine != 0);
            while (bal) {
                const int src = __ffsll((long long)bal) - 1;
                bal &= bal - 1;
                const int ti = __builtin_amdgcn_readfirstlane(c0 + src);
                const unsigned int gm = (unsigned int)__builtin_amdgcn_readlane((int)mine, src);
                any_work = true;
                if (ft != curft) {
                    __syncthreads();
                    for (int e = lane; e < N3; e += 64) sF[(e % N1) * P + e / N1] = fsrc[(size_t)ft * N3 + e];
                    curft = ft;
                    __syncthreads();
                }
                const int w4 = __builtin_amdgcn_readfirstlane(wd[ti * 5 + 4]);
                const long long mbase = ms[ti];
                unsigned int xneed = 0, yneed = 0;
#pragma unroll
                for (int gt = 0; gt < 8; gt++)
                    if (gm >> gt & 1) {
                        xneed |= 1u << (gt & 1);
                        yneed |= 1u << (gt & 3);
                    }
#pragma unroll
                for (int gx = 0; gx < 2; gx++) {
                    if (!(xneed >> gx & 1)) continue;
                    // x sweep: T1[i1][i2 + N1 j0] = sum_i0 F[i0][i1 + N1 i2] Ox(i0, j0)
                    const double *Ox = op.mats + (size_t)(mbase + (long long)(dl[0] + w4) * 4 + 2 * gx + (ft & 1)) * N2;
                    if (lane < N2) {
                        double x[N1];
#pragma unroll
                        for (int a = 0; a < N1; a++) x[a] = sF[a * P + lane];
                        const int i1 = lane % N1, i2 = lane / N1;
#pragma unroll
                        for (int j = 0; j < N1; j++) {
                            double s = 0.0;
#pragma unroll
                            for (int a = 0; a < N1; a++) s += x[a] * Ox[a + N1 * j];
                            sT1[i1 * P + i2 + N1 * j] = s;
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int gy = 0; gy < 2; gy++) {
                        if (!(yneed >> (gx + 2 * gy) & 1)) continue;
                        // y sweep: T2[i2][j0 + N1 j1] = sum_i1 T1[i1][i2 + N1 j0] Oy(i1, j1)
                        const double *Oy = op.mats + (size_t)(mbase + (long long)(dl[1] + w4) * 4 + 2 * gy + ((ft >> 1) & 1)) * N2;
                        if (lane < N2) {
                            double y[N1];
#pragma unroll
                            for (int a = 0; a < N1; a++) y[a] = sT1[a * P + lane];
                            const int i2 = lane % N1, j0 = lane / N1;
#pragma unroll
                            for (int j = 0; j < N1; j++) {
                                double s = 0.0;
#pragma unroll
                                for (int a = 0; a < N1; a++) s += y[a] * Oy[a + N1 * j];
                                sT2[i2 * P + j0 + N1 * j] = s;
                            }
                        }
                        __syncthreads();
                        // z sweep into the owned line (j0, j1): G[gt][j0 + N1 j1 + N2 j2] += sum_i2 T2[i2][line] Oz(i2, j2)
                        if (lane < N2) {
                            double z[N1];
#pragma unroll
                            for (int a = 0; a < N1; a++) z[a] = sT2[a * P + lane];
#pragma unroll
                            for (int gz = 0; gz < 2; gz++) {
                                const int gt = gx + 2 * gy + 4 * gz;
                                if (!(gm >> gt & 1)) continue;
                                const double *Oz = op.mats + (size_t)(mbase + (long long)(dl[2] + w4) * 4 + 2 * gz + ((ft >> 2) & 1)) * N2;
#pragma unroll
                                for (int j = 0; j < N1; j++) {
                                    double s = 0.0;
#pragma unroll
                                    for (int a = 0; a < N1; a++) s += z[a] * Oz[a + N1 * j];
                                    acc[gt][j] += s;
                                }
                            }
                        }
                        __syncthreads();
                    }
                }
            }
        }
    }
    // applied-tuple count of this pair
    for (int off = 32; off > 0; off >>= 1) napp += __shfl_xor(napp, off);
    if (lane == 0) {
        if (napp) atomicAdd(n_applied, (unsigned long long)napp);
        pair_work[p] = any_work ? 1 : 0;
    }
    if (!any_work) return;
    double *dst = partial + (size_t)p * 8 * N3;
    if (lane < N2) {
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int j = 0; j < N1; j++) dst[t * N3 + lane + N2 * j] = acc[t][j];
    }
}

// ---------------------------------------------------------------------------
// Per output node: sum the pair partials in pair order, store, norms.
// ---------------------------------------------------------------------------
template <int N1> __global__ __launch_bounds__(kThreads) void k_reduce(mwg_level lvl, mwg_pairs pr, const double *partial) {
    constexpr int N3 = N1 * N1 * N1;
    constexpr int R = (8 * N3 + kThreads - 1) / kThreads;
    constexpr int kChunk = 4 * kThreads;
    __shared__ double red[8 * (kThreads / 64)];
    __shared__ int list[kChunk];
    __shared__ int tmp[kThreads];
    const int g = blockIdx.x;
    const int first = pr.g_first[g], n = pr.g_npairs[g];
    double *dst = lvl.g_coef + lvl.g_off[g];
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0;
    // pairs with work, in pair order, a chunk at a time
    for (int c0 = 0; c0 < n; c0 += kChunk) {
        const int m = min(kChunk, n - c0);
        const int q0 = threadIdx.x * 4, q1 = min(m, q0 + 4);
        int cnt = 0;
        for (int q = q0; q < q1; q++) cnt += pr.pair_work[first + c0 + q] != 0;
        int excl;
        const int tot = block_exclusive_scan(cnt, tmp, excl);
        for (int q = q0; q < q1; q++)
            if (pr.pair_work[first + c0 + q]) list[excl++] = first + c0 + q;
        __syncthreads();
        for (int k = 0; k < tot; k++) {
            const double *src = partial + (size_t)list[k] * 8 * N3;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int o = threadIdx.x + r * kThreads;
                if (o < 8 * N3) acc[r] += src[o];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int o = threadIdx.x + r * kThreads;
        if (o < 8 * N3) dst[o] = acc[r];
    }
    __syncthreads();
    block_norms<N1>(dst, lvl.g_norms + (size_t)g * 9, red);
}

template <int N1>
__global__ __launch_bounds__(kThreads) void k_reconstruct_add(const double *filt, double *coef, const long long *poff, const long long *coff, long long stride) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double a[8 * N3], b[8 * N3];
    const int k = blockIdx.x;
    const double *src = coef + poff[k];
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) a[o] = src[o];
    __syncthreads();
    mw_pass<N1>(b, a, filt, 4, 0);
    __syncthreads();
    mw_pass<N1>(a, b, filt, 4, 1);
    __syncthreads();
    mw_pass_add_children<N1>(coef, coff[k], stride, a, filt);
}

template <int N1>
__global__ __launch_bounds__(kThreads) void k_compress(const double *filt, double *coef, const long long *poff, const long long *coff, long long stride) {
    constexpr int N3 = N1 * N1 * N1;
    __shared__ double a[8 * N3], b[8 * N3];
    const int k = blockIdx.x;
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) {
        const int t = o / N3, e = o % N3;
        a[o] = coef[coff[k] + (long long)t * stride + e];
    }
    __syncthreads();
    mw_pass<N1>(b, a, filt, 0, 0);
    __syncthreads();
    mw_pass<N1>(a, b, filt, 0, 1);
    __syncthreads();
    mw_pass<N1>(b, a, filt, 0, 2);
    __syncthreads();
    double *dst = coef + poff[k];
    for (int o = threadIdx.x; o < 8 * N3; o += kThreads) dst[o] = b[o];
}

template <int N1> __global__ __launch_bounds__(kThreads) void k_node_norms(const double *coef, const long long *off, double *norms) {
    __shared__ double red[8 * (kThreads / 64)];
    const int k = blockIdx.x;
    block_norms<N1>(coef + off[k], norms + (size_t)k * 9, red);
}

template <template <int> class Launcher, typename... Args> mwg_status dispatch(int kp1, Args... args) {
    switch (kp1) {
    case 4:
        return Launcher<4>::run(args...);
    case 6:
        return Launcher<6>::run(args...);
    case 8:
        return Launcher<8>::run(args...);
    case 10:
        return Launcher<10>::run(args...);
    default:
        return fail(MWG_ERR_UNSUPPORTED, "unsupported order: k+1 = " + std::to_string(kp1));
    }
}

template <int N1> struct GenerateL {
    static mwg_status run(const mwg_ftree *f, const double *filt, const int *ids, const int *par, const int *ch, int count) {
        ProfileScope prof_(MWG_KERNEL_GENERATE);
        hipLaunchKernelGGL(k_generate<N1>, dim3(count), dim3(kThreads), 0, 0, *f, filt, ids, par, ch);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct BandApplyL {
    static mwg_status run(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, const mwg_pairs *pr, int npairs,
                          double *partial, unsigned long long *napp) {
        ProfileScope prof_(MWG_KERNEL_BAND);
        if constexpr (N1 * N1 <= 64) {
            hipLaunchKernelGGL(k_band_apply_wave<N1>, dim3(npairs), dim3(64), 0, 0, *lvl, *f, *op, pr->pair_g, pr->pair_f, pr->pair_l,
                               pr->pair_work, partial, napp);
            MWG_LAUNCHED();
            return MWG_OK;
        }
        hipLaunchKernelGGL(k_band_apply<N1>, dim3(npairs), dim3(kThreads), 0, 0, *lvl, *f, *op, pr->pair_g, pr->pair_f, pr->pair_l, pr->pair_work, partial,
                           napp);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct ReduceL {
    static mwg_status run(const mwg_level *lvl, const mwg_pairs *pr, const double *partial) {
        ProfileScope prof_(MWG_KERNEL_REDUCE);
        hipLaunchKernelGGL(k_reduce<N1>, dim3(lvl->count), dim3(kThreads), 0, 0, *lvl, *pr, partial);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct ReconL {
    static mwg_status run(const double *filt, double *coef, const long long *po, const long long *co, int count, long long stride) {
        ProfileScope prof_(MWG_KERNEL_RECONSTRUCT);
        hipLaunchKernelGGL(k_reconstruct_add<N1>, dim3(count), dim3(kThreads), 0, 0, filt, coef, po, co, stride);
        MWG_LAUNCHED();
        return MWG_OK;
    }
};
template <int N1> struct CompressL {
    static mwg_status run(const double *filt, double *coef, const long long *po, const long long *co, int count, long long stride) {
        ProfileScope prof_(MWG_KERNEL_COMPRESS);
        hipLaunchKernelGGL(k_compress<N1>, dim3(count), dim3(kThreads), 0, 0, filt, coef, po, co, stride);
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

mwg_status need_init() { return g_ready ? MWG_OK : fail(MWG_ERR_ARG, "mwg_init has not been called"); }

// Device blocks cached by exact size for reuse (freed blocks are not returned
// to the runtime until mwg_release_pool).
std::mutex &pool_mutex() {
    static std::mutex m;
    return m;
}
std::map<size_t, std::vector<void *>> &free_blocks() {
    static std::map<size_t, std::vector<void *>> m;
    return m;
}
std::map<void *, size_t> &block_size() {
    static std::map<void *, size_t> m;
    return m;
}

} // namespace

extern "C" {

const char *mwg_last_error(void) { return g_err.c_str(); }

int mwg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

mwg_status mwg_init(int device) {
    int n = 0;
    MWG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(MWG_ERR_ARG, "no such device");
    MWG_HIP(hipSetDevice(device));
    g_ready = true;
    return MWG_OK;
}

mwg_status mwg_malloc(void **ptr, size_t bytes) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!ptr) return fail(MWG_ERR_ARG, "null pointer");
    if (!bytes) bytes = 1;
    {
        std::lock_guard<std::mutex> lk(pool_mutex());
        auto &fl = free_blocks()[bytes];
        if (!fl.empty()) {
            *ptr = fl.back();
            fl.pop_back();
            block_size()[*ptr] = bytes;
            return MWG_OK;
        }
    }
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        // return cached blocks to the runtime and retry once
        mwg_release_pool();
        if (hipMalloc(ptr, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MWG_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> lk(pool_mutex());
    block_size()[*ptr] = bytes;
    return MWG_OK;
}

mwg_status mwg_free(void *ptr) {
    if (!ptr) return MWG_OK;
    std::lock_guard<std::mutex> lk(pool_mutex());
    auto it = block_size().find(ptr);
    if (it == block_size().end()) return fail(MWG_ERR_ARG, "mwg_free of an unknown pointer");
    free_blocks()[it->second].push_back(ptr);
    block_size().erase(it);
    return MWG_OK;
}

mwg_status mwg_profile(int enable) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    MWG_HIP(hipDeviceSynchronize());
    for (auto &t : timed_launches()) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    timed_launches().clear();
    g_profile = enable != 0;
    return MWG_OK;
}

mwg_status mwg_profile_read(double *ms, long long *launches, int nkinds) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!ms || !launches) return fail(MWG_ERR_ARG, "null argument");
    MWG_HIP(hipDeviceSynchronize());
    for (int k = 0; k < nkinds; k++) {
        ms[k] = 0.0;
        launches[k] = 0;
    }
    for (auto &t : timed_launches()) {
        if (t.kind < 0 || t.kind >= nkinds) continue;
        float e = 0.0f;
        MWG_HIP(hipEventElapsedTime(&e, t.start, t.stop));
        ms[t.kind] += e;
        launches[t.kind]++;
    }
    return MWG_OK;
}

mwg_status mwg_release_pool(void) {
    std::lock_guard<std::mutex> lk(pool_mutex());
    for (auto &kv : free_blocks())
        for (void *p : kv.second) MWG_HIP(hipFree(p));
    free_blocks().clear();
    return MWG_OK;
}

mwg_status mwg_upload(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MWG_OK;
}

mwg_status mwg_download(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MWG_OK;
}

mwg_status mwg_copy(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MWG_OK;
}

mwg_status mwg_zero(void *dst, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipMemset(dst, 0, bytes));
    return MWG_OK;
}

mwg_status mwg_synchronize(void) {
    MWG_HIP(hipDeviceSynchronize());
    return MWG_OK;
}

mwg_status mwg_filters_upload(int kp1, const double *host_filters, double **dev_filters) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!host_filters || !dev_filters || kp1 < 1) return fail(MWG_ERR_ARG, "bad filter upload arguments");
    const size_t bytes = sizeof(double) * 8 * (size_t)kp1 * kp1;
    void *p = nullptr;
    if (mwg_malloc(&p, bytes) != MWG_OK) return MWG_ERR_NOMEM;
    MWG_HIP(hipMemcpy(p, host_filters, bytes, hipMemcpyHostToDevice));
    *dev_filters = static_cast<double *>(p);
    return MWG_OK;
}

mwg_status mwg_pair_screen(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, mwg_pairs *out) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!lvl || !f || !op || !out || !out->count) return fail(MWG_ERR_ARG, "null argument");
    if (lvl->o_depth < 0 || lvl->o_depth >= op->n_depth) return fail(MWG_ERR_ARG, "level depth outside the band tables");
    if (lvl->o_depth > mwgkey::kMaxDepth) return fail(MWG_ERR_UNSUPPORTED, "level deeper than the device key range");
    MWG_HIP(hipMemset(out->count, 0, sizeof(int)));
    if (lvl->count == 0) return MWG_OK;
    ProfileScope prof_(MWG_KERNEL_SCREEN);
    hipLaunchKernelGGL(k_pair_screen, dim3(lvl->count), dim3(kThreads), 0, 0, *lvl, *f, *op, *out);
    MWG_LAUNCHED();
    return MWG_OK;
}

mwg_status mwg_generate_nodes(const mwg_ftree *f, const double *filters, const int *ids, const int *parent, const int *child_idx, int count) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (count <= 0) return MWG_OK;
    if (!f || !filters || !ids || !parent || !child_idx) return fail(MWG_ERR_ARG, "null argument");
    return dispatch<GenerateL>(f->kp1, f, filters, ids, parent, child_idx, count);
}

mwg_status mwg_band_apply(const mwg_level *lvl, const mwg_ftree *f, const mwg_band_tables *op, const mwg_pairs *pairs, int npairs, double *partial,
                          unsigned long long *n_applied, int precision) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (npairs <= 0) return MWG_OK;
    if (!lvl || !f || !op || !pairs || !partial || !n_applied) return fail(MWG_ERR_ARG, "null argument");
    if (precision != 0) return fail(MWG_ERR_UNSUPPORTED, "only FP64 sweeps are available");
    if (op->M > kMaxTerms) return fail(MWG_ERR_UNSUPPORTED, "separation rank above 512");
    if (op->kp1 != f->kp1) return fail(MWG_ERR_ARG, "operator and tree orders differ");
    return dispatch<BandApplyL>(f->kp1, lvl, f, op, pairs, npairs, partial, n_applied);
}

mwg_status mwg_reduce_level(const mwg_level *lvl, const mwg_pairs *pairs, const double *partial) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!lvl || !pairs) return fail(MWG_ERR_ARG, "null argument");
    if (lvl->count == 0) return MWG_OK;
    return dispatch<ReduceL>(lvl->kp1, lvl, pairs, partial);
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
