// Host-side owners of device memory for the MI355X path: growable device
// arrays, operator band tables and the device mirror of an input tree.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mwg.h"
#include "keys.h"
#include "operators.h"
#include "trees.h"

namespace mrcpp {
namespace device {

void ensure_init();
// Process-wide lock of the device path.  The device layer keeps its scratch
// buffers, staging ring and stream per process, so calls that use it run one
// at a time: every mrcpp_* C ABI entry point and apply_batch() hold this
// (recursive) lock for their whole call.
std::recursive_mutex &api_mutex();
// whether a GPU is present and the device layer initialised
bool available();
void check(mwg_status s, const char *what);

template <typename T> class DeviceArray {
public:
    DeviceArray() = default;
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept
            : ptr_(o.ptr_)
            , cap_(o.cap_) {
        o.ptr_ = nullptr;
        o.cap_ = 0;
    }

    T *get() const { return ptr_; }
    size_t capacity() const { return cap_; }
    // grow to at least n elements; keeps the first `keep` elements (at most
    // the current capacity)
    void reserve(size_t n, size_t keep = 0) {
        if (n <= cap_) return;
        keep = std::min(keep, cap_);
        size_t nc = cap_ ? cap_ : 1024;
        while (nc < n) nc *= 2;
        void *p = nullptr;
        check(mwg_malloc(&p, nc * sizeof(T)), "device allocation");
        if (keep && ptr_) check(mwg_copy(p, ptr_, keep * sizeof(T)), "device copy");
        release();
        ptr_ = static_cast<T *>(p);
        cap_ = nc;
    }
    // enqueued without waiting; src may be reused on return
    void upload(const T *src, size_t n, size_t at = 0) {
        if (n) check(mwg_upload_async(ptr_ + at, src, n * sizeof(T)), "upload");
    }
    void download(T *dst, size_t n, size_t at = 0) const {
        if (n) check(mwg_download(dst, ptr_ + at, n * sizeof(T)), "download");
    }
    void zero(size_t n, size_t at = 0) {
        if (n) check(mwg_zero(ptr_ + at, n * sizeof(T)), "memset");
    }
    void release() {
        if (ptr_) mwg_free(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
    }

private:
    T *ptr_ = nullptr;
    size_t cap_ = 0;
};

// Page-locked host buffer, grown on demand (contents are not kept).
template <typename T> class PinnedArray {
public:
    PinnedArray() = default;
    ~PinnedArray() { release(); }
    PinnedArray(const PinnedArray &) = delete;
    PinnedArray &operator=(const PinnedArray &) = delete;

    T *data() const { return ptr_; }
    size_t capacity() const { return cap_; }
    void reserve(size_t n) {
        if (n <= cap_) return;
        size_t nc = cap_ ? cap_ : 4096;
        while (nc < n) nc *= 2;
        release();
        void *p = nullptr;
        check(mwg_host_alloc(&p, nc * sizeof(T)), "pinned host allocation");
        ptr_ = static_cast<T *>(p);
        cap_ = nc;
    }
    void release() {
        if (ptr_) mwg_host_free(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
    }

private:
    T *ptr_ = nullptr;
    size_t cap_ = 0;
};

} // namespace device

namespace detail {
struct DeviceBandTables {
    device::DeviceArray<int> bandMax, widths, bandSize;
    device::DeviceArray<long long> matStart;
    device::DeviceArray<double> mats, matNorm, pairBound, pairBoundW, termNorm;
    device::DeviceArray<int> bandSizeMax;
    device::DeviceArray<long long> screenOff;
    device::DeviceArray<double> screenBound; // largest per-term tuple bound per (depth, dl, ft)
    device::DeviceArray<unsigned char> frag[3]; // MFMA operand tables, built on first use
    mwg_band_tables view{};
    // view with the operand table of this sweep precision built
    const mwg_band_tables &viewFor(int precision);
};
} // namespace detail

namespace device {

detail::DeviceBandTables *create_band_tables(const detail::BandTables &t);
void destroy_band_tables(detail::DeviceBandTables *t);

// Device filter banks per order: 4 compression then 4 reconstruction blocks.
const double *filters(int order);

// Host mirror of the device open-addressing hash.
class KeyTable {
public:
    void reset(int min_entries);
    int log2cap() const { return log2cap_; }
    int capacity() const { return 1 << log2cap_; }
    // insert, or return the existing id
    int insert(const mwgkey::Key &key, int id);
    int find(const mwgkey::Key &key) const;
    int size() const { return count_; }
    // two words (hi, lo) per slot
    const std::vector<unsigned long long> &keys() const { return keys_; }
    const std::vector<int> &vals() const { return vals_; }
    // slots filled since markUploaded(); fullUpload() after a reset or a rehash
    bool fullUpload() const { return full_; }
    // Remember the current entries; rollback() empties every slot filled
    // since (queued as changed slots for the next upload) and returns true,
    // or returns false when a rehash moved the entries in between.  Entries
    // present at mark() were all inserted before the removed ones, so their
    // probe chains never pass through a removed slot.
    void mark() {
        sinceMark_.clear();
        markValid_ = true;
    }
    bool rollback();
    const std::vector<int> &dirty() const { return dirty_; }
    void markUploaded() {
        full_ = false;
        dirty_.clear();
    }

private:
    int log2cap_ = 0, count_ = 0;
    std::vector<unsigned long long> keys_;
    std::vector<int> vals_;
    bool full_ = true;
    std::vector<int> dirty_;
    std::vector<int> sinceMark_;
    bool markValid_ = false;
};

// Device mirror of a 3D function tree used as apply input: real nodes keep
// their host ids, generated band nodes are appended after them.
struct TreeMirror {
    int kp1 = 0, ncoef = 0;
    int nReal = 0, nGen = 0;
    long long genBase = 0; // coefficient offset of the generated-node region
    KeyTable table;
    DeviceArray<unsigned long long> keys;
    DeviceArray<int> vals;
    DeviceArray<long long> keyUpdates; // staging of incremental table uploads
    DeviceArray<double> coef, norms;
    DeviceArray<long long> coefOff;
    DeviceArray<unsigned char> isGen;
    std::vector<long long> hostOff;    // mirrors coefOff
    std::vector<int> genParent;        // host bookkeeping for generated nodes
    std::vector<int> genScale;
    std::vector<std::array<int, 3>> genL;
    // at least every node's norms (mwg_ftree::norm_max): the real nodes'
    // largest, with room for the rounding of generated nodes, which the
    // orthogonal two-scale transform bounds by their real ancestors
    double normMax = HUGE_VAL;
    bool valid = false;
    mwg_ftree view(const MultiResolutionAnalysis<3> &mra) const;
    void uploadTable();
};

mwgkey::Key key_of(const MultiResolutionAnalysis<3> &mra, int scale, const std::array<int, 3> &l);

// A generated node planned into a mirror: its id, parent id, child index in
// the parent and scale.
struct GenNode {
    int id, parent, child, scale;
};

// Plan node (n, l) of a mirror: its id if present (real or generated), else
// the chain of generated nodes from its nearest present ancestor, appended
// to `made` (MWTree::getNode -> MWNode::retrieveNode).  Host only; the new
// nodes get storage and coefficients in generate_mirror_nodes.
int plan_mirror_node(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, int n, const std::array<int, 3> &l, std::vector<GenNode> &made);

// Planned generated nodes of several mirrors (a mirror may appear more than
// once) and integers of the caller, sent to the device in one upload: per
// entry its key-table updates and, per new node, its coefficient offset, id,
// parent and child index.  run() then launches, entry by entry, the table
// update and one recording generation per scale, coarse to fine
// (mwg_hash_update, mwg_generate_nodes_recorded).
class MirrorBatch {
public:
    // Reserves the mirror's device storage for the nodes of `made` (new ids
    // following the mirror's previous nodes) and queues them; sorts `made`
    // by scale.
    void add(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const double *filters, std::vector<GenNode> &made);
    // Integers sent with the same upload; returns the index of the first.
    size_t addInts(const int *p, size_t n);
    // One upload, then the launches; returns the device address of the
    // integers added with addInts (valid until `staging` is reused).
    const int *run(DeviceArray<long long> &staging);
    bool empty() const { return entries_.empty() && ints_.empty(); }

private:
    const int *runTogether(DeviceArray<long long> &staging);
    struct Entry {
        TreeMirror *fm;
        const MultiResolutionAnalysis<3> *mra;
        const double *filters;
        size_t upd, nupd;              // table updates: words_[upd ..), 4 nupd words
        size_t offs, ids, n;           // offsets words_[offs ..), ids / parents / children ints_[ids ..)
        std::vector<std::pair<size_t, size_t>> runs; // (first, count) of each scale
        std::vector<int> runScale;                    // the scale of each run
    };
    std::vector<Entry> entries_;
    std::vector<long long> words_;
    std::vector<int> ints_;
};

// One mirror's planned nodes through a MirrorBatch of their own.
void generate_mirror_nodes(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const double *filters, std::vector<GenNode> &made,
                           DeviceArray<long long> &staging);

// Forget a mirror's generated nodes (FunctionTree::deleteGenerated).
void drop_generated(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const detail::NodeStore<3> &f);
// The same in two parts: the host table and lists (no device call; mirrors
// may be dropped concurrently), then the device table (returns at once when
// nothing was dropped).
bool drop_generated_host(TreeMirror &fm, const MultiResolutionAnalysis<3> &mra, const detail::NodeStore<3> &f);
void drop_generated_device(TreeMirror &fm);

} // namespace device
} // namespace mrcpp
