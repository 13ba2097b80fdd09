// Device exclusive prefix sums used to place candidate pairs and work units
// in a deterministic order: 1024 items per 256-thread block, block totals
// scanned recursively, block offsets added back.
#include "device_common.h"

namespace mwgdev {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4 * kScanThreads;

// Items scanned: n, or with a device count n_dev the first min(n, *n_dev * mult).
__device__ inline int scan_length(int n, const int *n_dev, int mult) { return n_dev ? min(n, *n_dev * mult) : n; }

// out[i] = exclusive scan of in within the block's 1024 items; bsum[block] = block total.
template <typename T> __global__ __launch_bounds__(kScanThreads) void k_scan_block(const T *in, T *out, T *bsum, int n, const int *n_dev, int mult) {
    __shared__ T tmp[kScanThreads / 64];
    n = scan_length(n, n_dev, mult);
    const int base = blockIdx.x * kScanItems + 4 * threadIdx.x;
    T v[4], s = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = (base + i < n) ? in[base + i] : T(0);
        s += v[i];
    }
    T excl;
    const T total = block_exclusive_scan<T, kScanThreads>(s, tmp, excl);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (base + i < n) out[base + i] = excl;
        excl += v[i];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// out[i] += boff[block]; out[n] = boff[nblocks].
template <typename T> __global__ __launch_bounds__(kScanThreads) void k_scan_add(T *out, const T *boff, int n, int nblocks, const int *n_dev, int mult) {
    n = scan_length(n, n_dev, mult);
    const T add = boff[blockIdx.x];
    const int base = blockIdx.x * kScanItems;
    for (int i = threadIdx.x; i < kScanItems; i += kScanThreads)
        if (base + i < n) out[base + i] += add;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = boff[nblocks];
}

// The whole scan in one workgroup for n <= kSingleMax: thread t owns the
// contiguous items [t c, t c + c); two passes over its items (sum, then
// write), so in and out may alias.
constexpr int kSingleThreads = 1024;
constexpr int kSingleMax = 64 * kSingleThreads;
template <typename T> __global__ __launch_bounds__(kSingleThreads) void k_scan_single(const T *in, T *out, int n, const int *n_dev, int mult) {
    __shared__ T tmp[kSingleThreads / 64];
    n = scan_length(n, n_dev, mult);
    const int c = (n + kSingleThreads - 1) / kSingleThreads;
    const int a = min(n, (int)threadIdx.x * c), b = min(n, a + c);
    T s = 0;
    for (int i = a; i < b; i++) s += in[i];
    T excl;
    const T total = block_exclusive_scan<T, kSingleThreads>(s, tmp, excl);
    for (int i = a; i < b; i++) {
        const T v = in[i];
        out[i] = excl;
        excl += v;
    }
    if (threadIdx.x == 0) out[n] = total;
}

template <typename T> Buffer &scan_temp(int level) {
    static Buffer b[8];
    return b[level];
}

// The total goes to out[n], or out[min(n, *n_dev * mult)] with a device count.
template <typename T> mwg_status scan_rec(const T *in, T *out, int n, int level, const int *n_dev, int mult) {
    const int nb = (n + kScanItems - 1) / kScanItems;
    if (nb > 1 && n <= kSingleMax) {
        hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(kSingleThreads), 0, 0, in, out, n, n_dev, mult);
        MWG_LAUNCHED();
        return MWG_OK;
    }
    if (nb <= 1) {
        if (n_dev) { // the total's place is known on the device only
            hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(kSingleThreads), 0, 0, in, out, n, n_dev, mult);
        } else {
            hipLaunchKernelGGL(k_scan_block<T>, dim3(1), dim3(kScanThreads), 0, 0, in, out, out + n, n, nullptr, 0);
        }
        MWG_LAUNCHED();
        return MWG_OK;
    }
    if (level >= 8) return fail(MWG_ERR_ARG, "scan too long");
    Buffer &tb = scan_temp<T>(level);
    if (tb.ensure(sizeof(T) * ((size_t)nb + 1)) != MWG_OK) return MWG_ERR_NOMEM;
    T *bs = tb.as<T>();
    // blocks past the device count have zero totals
    hipLaunchKernelGGL(k_scan_block<T>, dim3(nb), dim3(kScanThreads), 0, 0, in, out, bs, n, n_dev, mult);
    MWG_LAUNCHED();
    const mwg_status s = scan_rec<T>(bs, bs, nb, level + 1, nullptr, 0);
    if (s != MWG_OK) return s;
    hipLaunchKernelGGL(k_scan_add<T>, dim3(nb), dim3(kScanThreads), 0, 0, out, bs, n, nb, n_dev, mult);
    MWG_LAUNCHED();
    return MWG_OK;
}

} // namespace

mwg_status scan_exclusive(const int *in, int *out, int n, const int *n_dev, int mult) {
    if (n < 0 || (n_dev && mult <= 0)) return fail(MWG_ERR_ARG, "bad scan length");
    return scan_rec<int>(in, out, n, 0, n_dev, mult);
}

mwg_status scan_exclusive(const long long *in, long long *out, int n) {
    if (n < 0) return fail(MWG_ERR_ARG, "negative scan length");
    return scan_rec<long long>(in, out, n, 0, nullptr, 0);
}

} // namespace mwgdev
