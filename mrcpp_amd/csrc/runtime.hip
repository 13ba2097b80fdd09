// Device runtime of the mwg_* C ABI: device selection, the pooled allocator,
// transfers and per-kernel timing with HIP events.
#include <cstring>
#include <map>
#include <mutex>

#include "device_common.h"

#include <cstdlib>

namespace mwgdev {

namespace {

thread_local std::string g_err;
bool g_ready = false;
unsigned g_profile = 0; // kinds (bit 1 << kind) whose launches are bracketed by events

struct TimedLaunch {
    int kind;
    hipEvent_t start, stop;
};
std::vector<TimedLaunch> &timed_launches() {
    static std::vector<TimedLaunch> v;
    return v;
}
// Events are created once and reused, so timing adds no creation cost to the
// measured launches.
std::vector<hipEvent_t> &event_pool() {
    static std::vector<hipEvent_t> v;
    return v;
}
hipEvent_t take_event() {
    auto &pool = event_pool();
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e{};
    (void)hipEventCreate(&e);
    return e;
}
void recycle_events() {
    for (auto &t : timed_launches()) {
        event_pool().push_back(t.start);
        event_pool().push_back(t.stop);
    }
    timed_launches().clear();
}

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

// Pinned staging ring of asynchronous uploads.  Space is reused after the
// null stream has drained (any synchronising call, or when the ring is full).
constexpr size_t kStageBytes = 64ull << 20;
constexpr size_t kStageMaxCopy = 16ull << 20;
unsigned char *g_stage = nullptr;
bool g_exiting = false; // set by an exit handler
// pinned landing buffer of downloads up to kStageMaxCopy: the copy is queued
// on the stream behind the kernels (a pageable destination makes the runtime
// drain the stream before it starts the copy)
unsigned char *g_landing = nullptr;
size_t g_stage_used = 0;
// Downloads up to kWriteBackMax bytes: a kernel stores them straight into
// the mapped landing buffer (no copy engine round trip); larger ones use the
// copy engine.
constexpr size_t kWriteBackMax = 1ull << 20;
unsigned char *g_landing_dev = nullptr; // device address of g_landing, or null

} // namespace

__global__ __launch_bounds__(256) void k_write_back(unsigned *dst, const unsigned *src, size_t words) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// Up to kGatherMax ranges in one launch (blockIdx.y: the range).
constexpr int kGatherMax = 4;
struct GatherRanges {
    unsigned *dst[kGatherMax];
    const unsigned *src[kGatherMax];
    size_t words[kGatherMax];
};
__global__ __launch_bounds__(256) void k_write_back_many(GatherRanges r) {
    const int k = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < r.words[k]; i += (size_t)gridDim.x * 256) r.dst[k][i] = r.src[k][i];
}

mwg_status fail(mwg_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

mwg_status need_init() { return g_ready ? MWG_OK : fail(MWG_ERR_ARG, "mwg_init has not been called"); }

ProfileScope::ProfileScope(int k)
        : kind(k) {
    if (!(g_profile >> k & 1u)) return;
    start = take_event();
    stop = take_event();
    (void)hipEventRecord(start, 0);
}

ProfileScope::~ProfileScope() {
    if (!(g_profile >> kind & 1u)) return;
    (void)hipEventRecord(stop, 0);
    timed_launches().push_back({kind, start, stop});
}

mwg_status Buffer::ensure(size_t need) {
    if (need <= bytes) return MWG_OK;
    if (p) {
        mwg_free(p);
        p = nullptr;
        bytes = 0;
    }
    size_t b = need + need / 4;
    if (mwg_malloc(&p, b) != MWG_OK) return MWG_ERR_NOMEM;
    bytes = b;
    return MWG_OK;
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    }
    return cus;
}

__global__ __launch_bounds__(256) void k_hash_update(unsigned long long *keys, int *vals, const long long *upd, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long slot = upd[i];
    keys[2 * slot] = (unsigned long long)upd[n + 2 * (size_t)i];
    keys[2 * slot + 1] = (unsigned long long)upd[n + 2 * (size_t)i + 1];
    vals[slot] = (int)upd[3 * (size_t)n + i];
}

__global__ __launch_bounds__(256) void k_hash_update_multi(const mwg_hash_table *tables, const long long *upd, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long *r = upd + 5 * (size_t)i;
    const mwg_hash_table t = tables[r[0]];
    const long long slot = r[1];
    t.keys[2 * slot] = (unsigned long long)r[2];
    t.keys[2 * slot + 1] = (unsigned long long)r[3];
    t.vals[slot] = (int)r[4];
}

} // namespace mwgdev

using namespace mwgdev;

extern "C" {

mwg_status mwg_hash_update(unsigned long long *keys, int *vals, const long long *upd, int n) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (n < 0) return fail(MWG_ERR_ARG, "negative entry count");
    if (n == 0) return MWG_OK;
    if (!keys || !vals || !upd) return fail(MWG_ERR_ARG, "null argument");
    hipLaunchKernelGGL(k_hash_update, dim3((n + 255) / 256), dim3(256), 0, 0, keys, vals, upd, n);
    MWG_LAUNCHED();
    return MWG_OK;
}

mwg_status mwg_hash_update_multi(const mwg_hash_table *tables, int ntables, const long long *upd, int n) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (n < 0 || ntables < 0) return fail(MWG_ERR_ARG, "negative count");
    if (n == 0) return MWG_OK;
    if (!tables || !upd || ntables == 0) return fail(MWG_ERR_ARG, "null argument");
    hipLaunchKernelGGL(k_hash_update_multi, dim3((n + 255) / 256), dim3(256), 0, 0, tables, upd, n);
    MWG_LAUNCHED();
    return MWG_OK;
}

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
    const mwg_status st = mlp_scratch_setup(device);
    if (st != MWG_OK) return st;
    // registered after the runtime's own exit handlers, so it runs before
    // them: device memory held by static objects is not freed at exit
    static const bool registered = std::atexit([] { g_exiting = true; }) == 0;
    (void)registered;
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
    if (g_exiting) return MWG_OK; // process exit: the runtime may be gone, the memory goes with the process
    // No wait: every kernel and copy of the library runs on the null stream
    // in order, and a block handed out again is only used by work enqueued
    // after this free, which the stream runs after the work that still uses
    // the block (synchronous host copies are ordered on the same stream).
    std::lock_guard<std::mutex> lk(pool_mutex());
    auto it = block_size().find(ptr);
    if (it == block_size().end()) return fail(MWG_ERR_ARG, "mwg_free of an unknown pointer");
    free_blocks()[it->second].push_back(ptr);
    block_size().erase(it);
    return MWG_OK;
}

mwg_status mwg_profile(int enable) { return mwg_profile_kinds(enable ? ~0u : 0u); }

mwg_status mwg_profile_kinds(unsigned mask) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    MWG_HIP(hipDeviceSynchronize());
    recycle_events();
    g_profile = mask;
    if (g_profile) { // enough events for a few thousand launches, each recorded once so first use costs nothing
        while (event_pool().size() < 8192) {
            hipEvent_t e{};
            MWG_HIP(hipEventCreate(&e));
            MWG_HIP(hipEventRecord(e, 0));
            event_pool().push_back(e);
        }
        MWG_HIP(hipDeviceSynchronize());
    }
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

mwg_status mwg_host_alloc(void **ptr, size_t bytes) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!ptr) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return MWG_OK;
}

mwg_status mwg_host_free(void *ptr) {
    if (!ptr || g_exiting) return MWG_OK;
    MWG_HIP(hipDeviceSynchronize());
    g_stage_used = 0;
    MWG_HIP(hipHostFree(ptr));
    return MWG_OK;
}

mwg_status mwg_upload(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    MWG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MWG_OK;
}

mwg_status mwg_upload_async(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    if (bytes > kStageMaxCopy) return mwg_upload(dst, src, bytes);
    if (!g_stage) MWG_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_stage), kStageBytes, hipHostMallocDefault));
    const size_t at = (g_stage_used + 255) & ~size_t(255);
    if (at + bytes > kStageBytes) {
        MWG_HIP(hipStreamSynchronize(0));
        g_stage_used = 0;
        return mwg_upload_async(dst, src, bytes);
    }
    std::memcpy(g_stage + at, src, bytes);
    MWG_HIP(hipMemcpyAsync(dst, g_stage + at, bytes, hipMemcpyHostToDevice, 0));
    g_stage_used = at + bytes;
    return MWG_OK;
}

// The pinned landing buffer of downloads and its device address (when the
// runtime maps it).
static mwg_status ensure_landing() {
    if (g_landing) return MWG_OK;
    MWG_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_landing), kStageMaxCopy, hipHostMallocMapped | hipHostMallocCoherent));
    void *dev = nullptr;
    if (hipHostGetDevicePointer(&dev, g_landing, 0) == hipSuccess) g_landing_dev = static_cast<unsigned char *>(dev);
    else (void)hipGetLastError();
    return MWG_OK;
}

mwg_status mwg_download(void *dst, const void *src, size_t bytes) {
    if (!bytes) return MWG_OK;
    if (!dst || !src) return fail(MWG_ERR_ARG, "null pointer");
    if (bytes <= kStageMaxCopy) {
        if (mwg_status s = ensure_landing(); s != MWG_OK) return s;
        const bool words = bytes % 4 == 0 && reinterpret_cast<uintptr_t>(src) % 4 == 0;
        if (g_landing_dev && words && bytes <= kWriteBackMax) {
            const size_t n = bytes / 4;
            const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
            hipLaunchKernelGGL(k_write_back, dim3(blocks), dim3(256), 0, 0, reinterpret_cast<unsigned *>(g_landing_dev),
                               static_cast<const unsigned *>(src), n);
            MWG_LAUNCHED();
        } else {
            MWG_HIP(hipMemcpyAsync(g_landing, src, bytes, hipMemcpyDeviceToHost, 0));
        }
        MWG_HIP(hipStreamSynchronize(0));
        std::memcpy(dst, g_landing, bytes);
    } else {
        MWG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
    g_stage_used = 0; // the null stream has drained
    return MWG_OK;
}

mwg_status mwg_download_many(int n, void *const *dst, const void *const *src, const size_t *bytes) {
    if (n < 0 || (n > 0 && (!dst || !src || !bytes))) return fail(MWG_ERR_ARG, "null argument");
    size_t total = 0;
    bool words = n <= kGatherMax;
    for (int k = 0; k < n; k++) {
        if (bytes[k] && (!dst[k] || !src[k])) return fail(MWG_ERR_ARG, "null pointer");
        total += (bytes[k] + 255) & ~size_t(255);
        words = words && bytes[k] % 4 == 0 && reinterpret_cast<uintptr_t>(src[k]) % 4 == 0;
    }
    if (mwg_status s = ensure_landing(); s != MWG_OK) return s;
    if (n == 1 || total > kWriteBackMax || !words || !g_landing_dev) { // one download each
        for (int k = 0; k < n; k++)
            if (mwg_status s = mwg_download(dst[k], src[k], bytes[k]); s != MWG_OK) return s;
        return MWG_OK;
    }
    GatherRanges r{};
    size_t at = 0, most = 0;
    for (int k = 0; k < n; k++) {
        r.dst[k] = reinterpret_cast<unsigned *>(g_landing_dev + at);
        r.src[k] = static_cast<const unsigned *>(src[k]);
        r.words[k] = bytes[k] / 4;
        most = std::max(most, r.words[k]);
        at += (bytes[k] + 255) & ~size_t(255);
    }
    if (most) {
        const unsigned blocks = (unsigned)std::min<size_t>((most + 255) / 256, 1024);
        hipLaunchKernelGGL(k_write_back_many, dim3(blocks, (unsigned)n), dim3(256), 0, 0, r);
        MWG_LAUNCHED();
    }
    MWG_HIP(hipStreamSynchronize(0));
    at = 0;
    for (int k = 0; k < n; k++) {
        if (bytes[k]) std::memcpy(dst[k], g_landing + at, bytes[k]);
        at += (bytes[k] + 255) & ~size_t(255);
    }
    g_stage_used = 0; // the null stream has drained
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
    g_stage_used = 0;
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

} // extern "C"
