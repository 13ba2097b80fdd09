// Multiresolution analysis, world box and the flat node store used by every tree
// on the host (1D kernel trees, 2D operator trees, 3D functions).
//
// Nodes live in index-addressed arrays (metadata vectors plus chunked
// coefficient storage) so that a 3D tree maps onto device buffers one to one:
// a node id is also its coefficient slot.  The 2^D children of a node always
// occupy consecutive ids, which the MW transforms rely on.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <unordered_map>

#include <sys/mman.h>
#include <vector>

#include "core.h"

namespace mrcpp {

template <int D> using Coord = std::array<double, D>;

// World box (src/trees/BoundingBox.cpp:47-80, 232-239): root scale n, corner
// translation, number of root boxes per direction and scaling factor.
template <int D> class BoundingBox {
public:
    BoundingBox() = default;
    // {-L, L} or {0, L}: root scale chosen so that the unit length is in [1,2)
    explicit BoundingBox(std::array<int, 2> box) {
        if (box[1] < 1) throw Error("Invalid upper bound");
        if (!(box[0] == 0 || box[0] == -box[1])) throw Error("Invalid lower bound");
        int n = 0;
        double size = 1.0 * box[1];
        while (size >= 2.0) {
            size /= 2.0;
            n--;
        }
        scale_ = n;
        corner_.fill(box[0] == 0 ? 0 : -1);
        nbox_.fill(box[0] == 0 ? 1 : 2);
        sf_.fill(size);
    }
    // root scale n, corner translation, root boxes per direction (one where
    // not positive) and scaling factors (all-zero: 1.0 each), as
    // src/trees/BoundingBox.cpp:100-106,211-217,247-253 default them
    explicit BoundingBox(int n, const std::array<int, D> &corner = {}, const std::array<int, D> &nb = {}, const std::array<double, D> &sf = {})
            : scale_(n)
            , corner_(corner)
            , nbox_(nb)
            , sf_(sf) {
        for (int &x : nbox_) x = x > 0 ? x : 1;
        if (sf_ == std::array<double, D>{}) sf_.fill(1.0);
        for (double x : sf_)
            if (!(x > 0.0)) throw Error("Non-positive scaling factor");
    }

    int getScale() const { return scale_; }
    int size(int d) const { return nbox_[d]; }
    int size() const {
        int t = 1;
        for (int d = 0; d < D; d++) t *= nbox_[d];
        return t;
    }
    const std::array<int, D> &getCornerIndex() const { return corner_; }
    const std::array<int, D> &getNBoxes() const { return nbox_; }
    double getScalingFactor(int d) const { return sf_[d]; }
    const std::array<double, D> &getScalingFactors() const { return sf_; }
    double getUnitLength(int d) const { return sf_[d] * std::pow(2.0, -scale_); }
    double getLowerBound(int d) const { return corner_[d] * getUnitLength(d); }
    double getUpperBound(int d) const { return getLowerBound(d) + getUnitLength(d) * nbox_[d]; }

    // root number -> translation, x fastest (BoundingBox::getNodeIndex)
    std::array<int, D> rootTranslation(int r) const {
        std::array<int, D> l;
        for (int d = 0; d < D; d++) {
            l[d] = corner_[d] + r % nbox_[d];
            r /= nbox_[d];
        }
        return l;
    }
    // translation at scale n -> root number, or -1 outside (BoundingBox::getBoxIndex)
    int rootIndex(int n, const std::array<int, D> &l) const {
        int rel = n - scale_;
        if (rel < 0) return -1;
        int r = 0, stride = 1;
        for (int d = 0; d < D; d++) {
            int t = (l[d] >> rel) - corner_[d];
            if (t < 0 || t >= nbox_[d]) return -1;
            r += stride * t;
            stride *= nbox_[d];
        }
        return r;
    }
    bool operator==(const BoundingBox &o) const { return scale_ == o.scale_ && corner_ == o.corner_ && nbox_ == o.nbox_ && sf_ == o.sf_; }

private:
    int scale_ = 0;
    std::array<int, D> corner_{}, nbox_{};
    std::array<double, D> sf_{};
};

// Scaling basis descriptors (src/core/ScalingBasis.h:36-67)
class ScalingBasis {
public:
    ScalingBasis(int order, int type)
            : order_(order)
            , type_(type) {
        if (order < 0 || order > MaxOrder) throw Error("Invalid scaling order");
    }
    int getScalingOrder() const { return order_; }
    int getScalingType() const { return type_; }

private:
    int order_, type_;
};

class LegendreBasis : public ScalingBasis {
public:
    explicit LegendreBasis(int order)
            : ScalingBasis(order, Legendre) {}
};

// src/trees/MultiResolutionAnalysis.h:49-82 (Legendre bases only)
template <int D> class MultiResolutionAnalysis {
public:
    MultiResolutionAnalysis(const BoundingBox<D> &bb, const ScalingBasis &sb, int depth = MaxDepth)
            : box_(bb)
            , order_(sb.getScalingOrder())
            , depth_(depth) {
        if (sb.getScalingType() != Legendre) throw Error("only the Legendre scaling basis is supported");
        if (depth_ > MaxDepth) throw Error("Beyond MaxDepth");
        basis_ = &core::legendre(order_);
        filter_ = &core::filters(order_);
    }
    MultiResolutionAnalysis(std::array<int, 2> bb, int order, int depth = MaxDepth)
            : MultiResolutionAnalysis(BoundingBox<D>(bb), LegendreBasis(order), depth) {}

    int getOrder() const { return order_; }
    int getKp1() const { return order_ + 1; }
    int getKp1_d() const { return core::ipow(order_ + 1, D); }
    int getMaxDepth() const { return depth_; }
    int getMaxScale() const { return box_.getScale() + depth_; }
    int getRootScale() const { return box_.getScale(); }
    const BoundingBox<D> &getWorldBox() const { return box_; }
    const core::LegendreTables &basis() const { return *basis_; }
    const core::FilterBank &filter() const { return *filter_; }
    // MultiResolutionAnalysis::calcMinDistance / calcMaxDistance
    double calcMinDistance(double eps) const { return std::sqrt(eps * std::pow(2.0, -getMaxScale())); }
    double calcMaxDistance() const {
        double s = 0.0;
        for (int d = 0; d < D; d++) {
            double x = box_.getUpperBound(d) - box_.getLowerBound(d);
            s += x * x;
        }
        return std::sqrt(s);
    }
    bool operator==(const MultiResolutionAnalysis &o) const { return box_ == o.box_ && order_ == o.order_ && depth_ == o.depth_; }
    bool operator!=(const MultiResolutionAnalysis &o) const { return !(*this == o); }

private:
    BoundingBox<D> box_;
    int order_, depth_;
    const core::LegendreTables *basis_ = nullptr;
    const core::FilterBank *filter_ = nullptr;
};

namespace detail {

template <int D> struct KeyHash {
    size_t operator()(const std::pair<int, std::array<int, D>> &k) const {
        uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)k.first;
        for (int d = 0; d < D; d++) {
            h ^= (uint64_t)(uint32_t)k.second[d] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        }
        return (size_t)h;
    }
};

// Flat node store.  Kinds: a function tree (children generated below leaves
// are scaling-only "generated" nodes) or an operator tree (children generated
// on demand are ordinary nodes, src/trees/OperatorNode.cpp:144-147).
template <int D> class NodeStore {
public:
    static constexpr int TD = 1 << D;
    using Index = std::array<int, D>;

    NodeStore(const MultiResolutionAnalysis<D> &mra, bool operator_tree = false, double norm_prec = 0.0)
            : mra_(&mra)
            , oper_(operator_tree)
            , norm_prec_(norm_prec) {
        kp1_ = mra.getKp1();
        kp1_d_ = mra.getKp1_d();
        ncoef_ = TD * kp1_d_;
        const auto &box = mra.getWorldBox();
        for (int r = 0; r < box.size(); r++) roots_.push_back(add_node(box.getScale(), box.rootTranslation(r), -1, false));
    }

    const MultiResolutionAnalysis<D> &mra() const { return *mra_; }
    bool isOperator() const { return oper_; }
    double normPrec() const { return norm_prec_; }
    int kp1() const { return kp1_; }
    int kp1_d() const { return kp1_d_; }
    int ncoef() const { return ncoef_; }
    int size() const { return (int)n_.size(); }
    // Room for n nodes without reallocation: the per-node arrays and the
    // index (kept at most a quarter full up to n nodes).
    void reserve(int n) {
        if (n <= 0) return;
        n_.reserve(n);
        l_.reserve(n);
        parent_.reserve(n);
        child0_.reserve(n);
        gen_.reserve(n);
        has_coefs_.reserve(n);
        norm_.reserve((size_t)n * TD);
        sq_.reserve(n);
        size_t want = 1024;
        while (want < 4 * (size_t)n) want *= 2;
        if (want > slots_.size()) {
            std::vector<int> old(want, -1);
            old.swap(slots_);
            indexed_ = 0;
            for (int v : old)
                if (v >= 0) place(v);
        }
    }
    const std::vector<int> &roots() const { return roots_; }
    int rootScale() const { return mra_->getRootScale(); }

    int scale(int u) const { return n_[u]; }
    int depth(int u) const { return n_[u] - rootScale(); }
    const Index &transl(int u) const { return l_[u]; }
    int parent(int u) const { return parent_[u]; }
    int child0(int u) const { return child0_[u]; }
    bool isGen(int u) const { return gen_[u] != 0; }
    bool hasChildren(int u) const { return child0_[u] >= 0; }
    bool hasRealChildren(int u) const { return child0_[u] >= 0 && !gen_[child0_[u]]; }
    bool isEnd(int u) const { return !gen_[u] && !hasRealChildren(u); }
    bool hasCoefs(int u) const { return has_coefs_[u] != 0; }
    void setHasCoefs(int u, bool v) { has_coefs_[u] = v; }
    double *coefs(int u) { return chunks_[u >> kChunkBits].get() + (size_t)(u & kChunkMask) * ncoef_; }
    const double *coefs(int u) const { return chunks_[u >> kChunkBits].get() + (size_t)(u & kChunkMask) * ncoef_; }
    double *norms(int u) { return &norm_[(size_t)u * TD]; }
    const double *norms(int u) const { return &norm_[(size_t)u * TD]; }
    double compNorm(int u, int t) const { return norm_[(size_t)u * TD + t]; }
    double squareNorm(int u) const { return sq_[u]; }
    void setSquareNorm(int u, double v) { sq_[u] = v; }
    // MWNode::getScalingNorm / getWaveletNorm (src/trees/MWNode.cpp:619-640)
    double scalingNorm(int u) const {
        double s = compNorm(u, 0);
        return s >= 0.0 ? s * s : -1.0;
    }
    double waveletNorm(int u) const {
        double w = 0.0;
        for (int t = 1; t < TD; t++) {
            double v = compNorm(u, t);
            if (v >= 0.0)
                w += v * v;
            else
                w = -1.0;
        }
        return w;
    }
    void clearNorms(int u) {
        for (int t = 0; t < TD; t++) norm_[(size_t)u * TD + t] = -1.0;
        sq_[u] = -1.0;
    }
    void zeroCoefs(int u) {
        std::memset(coefs(u), 0, sizeof(double) * (isGen(u) ? kp1_d_ : ncoef_));
        for (int t = 0; t < TD; t++) norm_[(size_t)u * TD + t] = 0.0;
        sq_[u] = 0.0;
        has_coefs_[u] = 1;
    }
    std::vector<int> &nodesAtDepth() { return nodes_at_depth_; }
    int getDepth() const { return (int)nodes_at_depth_.size(); }

    double treeSquareNorm = -1.0;

    int find(int n, const Index &l) const {
        if (slots_.empty()) return -1;
        const size_t mask = slots_.size() - 1;
        for (size_t s = slot_of(n, l);; s = (s + 1) & mask) {
            const int u = slots_[s];
            if (u < 0) return -1;
            if (n_[u] == n && l_[u] == l) return u;
        }
    }
    int findReal(int n, const Index &l) const {
        int u = find(n, l);
        return (u >= 0 && !isGen(u)) ? u : -1;
    }

    // New block of 2^D children (consecutive ids); returns the first id.
    int createChildren(int u, bool gen) {
        if (hasChildren(u)) throw Error("Node already has children");
        int first = -1;
        for (int c = 0; c < TD; c++) {
            Index l;
            for (int d = 0; d < D; d++) l[d] = 2 * l_[u][d] + ((c >> d) & 1);
            int v = add_node(n_[u] + 1, l, u, gen);
            if (c == 0) first = v;
        }
        child0_[u] = first;
        return first;
    }

    // Remove the subtree below u (its nodes leave the index and the depth
    // counts; their slots stay allocated and unreachable).
    void detachChildren(int u) {
        if (!hasChildren(u)) return;
        std::vector<int> stack = {u};
        while (!stack.empty()) {
            const int v = stack.back();
            stack.pop_back();
            if (!hasChildren(v)) continue;
            for (int c = 0; c < TD; c++) {
                const int w = child0_[v] + c;
                unindex(w);
                if (!gen_[w]) nodes_at_depth_[n_[w] - rootScale()]--;
                stack.push_back(w);
            }
            if (v != u) child0_[v] = -1;
        }
        child0_[u] = -1;
    }

    // Drop every generated child block (FunctionTree::deleteGenerated).  Ids of
    // generated nodes are recycled only by truncation, so generated nodes are
    // expected at the tail (they are created after all real nodes of a build).
    void deleteGenerated() {
        int keep = size();
        for (int u = 0; u < size(); u++)
            if (gen_[u]) {
                keep = std::min(keep, u);
            }
        for (int u = 0; u < keep; u++)
            if (child0_[u] >= keep) child0_[u] = -1;
        for (int u = keep; u < size(); u++) unindex(u);
        n_.resize(keep);
        l_.resize(keep);
        parent_.resize(keep);
        child0_.resize(keep);
        gen_.resize(keep);
        has_coefs_.resize(keep);
        norm_.resize((size_t)keep * TD);
        sq_.resize(keep);
    }

private:
    static constexpr int kChunkBits = 11; // 2048 nodes per host coefficient chunk (virtual until touched)
    static constexpr int kChunkMask = (1 << kChunkBits) - 1;

    int add_node(int n, const Index &l, int parent, bool gen) {
        int u = size();
        if ((u >> kChunkBits) >= (int)chunks_.size()) {
            // anonymous pages are zero and mapped lazily: trees whose
            // coefficients live on the device never touch this memory (a
            // calloc from a reused heap block would clear it eagerly)
            const size_t bytes = (size_t)(1 << kChunkBits) * ncoef_ * sizeof(double);
            void *p = ::mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
            if (p == MAP_FAILED) throw Error("host allocation failed");
            chunks_.emplace_back(static_cast<double *>(p), MapDeleter{bytes});
        }
        n_.push_back(n);
        l_.push_back(l);
        parent_.push_back(parent);
        child0_.push_back(-1);
        gen_.push_back(gen ? 1 : 0);
        has_coefs_.push_back(0);
        norm_.insert(norm_.end(), TD, -1.0);
        sq_.push_back(-1.0);
        index(u);
        if (!gen) {
            int dep = n - rootScale();
            if (dep >= (int)nodes_at_depth_.size()) nodes_at_depth_.resize(dep + 1, 0);
            nodes_at_depth_[dep]++;
        }
        return u;
    }

    const MultiResolutionAnalysis<D> *mra_;
    bool oper_;
    double norm_prec_;
    int kp1_ = 0, kp1_d_ = 0, ncoef_ = 0;
    std::vector<int> roots_;
    std::vector<int> n_;
    std::vector<Index> l_;
    std::vector<int> parent_, child0_;
    std::vector<uint8_t> gen_, has_coefs_;
    std::vector<double> norm_, sq_;
    struct MapDeleter {
        size_t bytes = 0;
        void operator()(double *p) const { ::munmap(p, bytes); }
    };
    std::vector<std::unique_ptr<double, MapDeleter>> chunks_;
    // Open-addressing index (n, l) -> node id with linear probing: a slot holds
    // a node id (-1 empty) and the key is that node's own (n, l).
    std::vector<int> slots_;
    size_t indexed_ = 0;
    size_t slot_of(int n, const Index &l) const { return KeyHash<D>{}({n, l}) & (slots_.size() - 1); }
    void index(int u) {
        if (2 * (indexed_ + 1) > slots_.size()) {
            std::vector<int> old(std::max<size_t>(1024, 2 * slots_.size()), -1);
            old.swap(slots_);
            indexed_ = 0;
            for (int v : old)
                if (v >= 0) place(v);
        }
        place(u);
    }
    void place(int u) {
        const size_t mask = slots_.size() - 1;
        size_t s = slot_of(n_[u], l_[u]);
        while (slots_[s] >= 0) {
            if (n_[slots_[s]] == n_[u] && l_[slots_[s]] == l_[u]) { // same index: the newer node replaces it
                slots_[s] = u;
                return;
            }
            s = (s + 1) & mask;
        }
        slots_[s] = u;
        indexed_++;
    }
    // Remove node u's entry (if the index still maps its (n, l) to u), closing
    // the gap by shifting later entries of the probe run back.
    void unindex(int u) {
        if (slots_.empty()) return;
        const size_t mask = slots_.size() - 1;
        size_t s = slot_of(n_[u], l_[u]);
        while (slots_[s] != u) {
            if (slots_[s] < 0) return;
            s = (s + 1) & mask;
        }
        slots_[s] = -1;
        indexed_--;
        for (size_t j = (s + 1) & mask; slots_[j] >= 0; j = (j + 1) & mask) {
            const size_t h = slot_of(n_[slots_[j]], l_[slots_[j]]);
            // entry j may move to the hole s unless its home h lies cyclically in (s, j]
            const bool stays = (s < j) ? (h > s && h <= j) : (h > s || h <= j);
            if (!stays) {
                slots_[s] = slots_[j];
                slots_[j] = -1;
                s = j;
            }
        }
    }
    std::vector<int> nodes_at_depth_;
};

// ---------------------------------------------------------------------------
// Node-level operations
// ---------------------------------------------------------------------------

// Component norms.  Function trees: plain 2-norms, zero wavelet norms for
// generated nodes (MWNode::calcComponentNorm, src/trees/MWNode.cpp:644-655).
// Operator trees: the vector / sqrt(|.|_1 |.|_inf) / Frobenius cascade against
// max(MachinePrec, prec/(8*2^depth)) (src/trees/OperatorNode.cpp:56-80).
template <int D> void compute_norms(NodeStore<D> &s, int u) {
    const int kd = s.kp1_d();
    double *cn = s.norms(u);
    double total = 0.0;
    for (int t = 0; t < NodeStore<D>::TD; t++) {
        double v = 0.0;
        if (s.isOperator()) {
            const double *m = s.coefs(u) + (size_t)t * kd;
            const int n = s.kp1();
            double fro = 0.0;
            for (int i = 0; i < kd; i++) fro += m[i] * m[i];
            fro = std::sqrt(fro);
            const double thr = std::max(MachinePrec, s.normPrec() / (8.0 * (1 << s.depth(u))));
            if (fro > thr) {
                double rmax = 0.0, cmax = 0.0;
                for (int r = 0; r < n; r++) {
                    double a = 0.0;
                    for (int c = 0; c < n; c++) a += std::abs(m[r + n * c]);
                    rmax = std::max(rmax, a);
                }
                for (int c = 0; c < n; c++) {
                    double a = 0.0;
                    for (int r = 0; r < n; r++) a += std::abs(m[r + n * c]);
                    cmax = std::max(cmax, a);
                }
                if (std::sqrt(rmax * cmax) > thr && fro > thr) v = fro;
            }
        } else if (!(s.isGen(u) && t != 0)) {
            const double *b = s.coefs(u) + (size_t)t * kd;
            for (int i = 0; i < kd; i++) v += b[i] * b[i];
            v = std::sqrt(v);
        }
        cn[t] = v;
        total += v * v;
    }
    s.setSquareNorm(u, total);
}

// Node-local multiwavelet transform between the 2^D child scaling blocks held in
// the node's own coefficients and the s/d blocks (MWNode::mwTransform,
// src/trees/MWNode.cpp:557-594).
template <int D> void local_transform(const NodeStore<D> &s, double *coefs, int dir) {
    constexpr int TD = NodeStore<D>::TD;
    const int kd = s.kp1_d(), m = kd / s.kp1();
    const auto &fb = s.mra().filter();
    std::vector<double> buf((size_t)TD * kd);
    double *src = coefs, *dst = buf.data();
    for (int dim = 0; dim < D; dim++) {
        const int bit = 1 << dim;
        for (int gt = 0; gt < TD; gt++) {
            bool acc = false;
            for (int ft = 0; ft < TD; ft++) {
                if ((gt | bit) != (ft | bit)) continue;
                const int f = 2 * ((gt >> dim) & 1) + ((ft >> dim) & 1);
                core::filter_pass(dst + (size_t)gt * kd, src + (size_t)ft * kd, fb.sub(f, dir), m, acc);
                acc = true;
            }
        }
        std::swap(src, dst);
    }
    if (D % 2 == 1) std::memcpy(coefs, src, sizeof(double) * TD * kd);
}

template <int D> void local_transform(NodeStore<D> &s, int u, int dir) { local_transform(s, s.coefs(u), dir); }

// Parent s/d (scaling only for a generated parent) -> scaling blocks of the
// 2^D children, overwriting or accumulating (tree_utils::mw_transform,
// src/utils/tree_utils.cpp:113-216).
template <int D> void reconstruct_into_children(NodeStore<D> &s, int u, bool overwrite) {
    constexpr int TD = NodeStore<D>::TD;
    const int kd = s.kp1_d(), m = kd / s.kp1();
    const auto &fb = s.mra().filter();
    const bool scaling_only = s.isGen(u);
    const int c0 = s.child0(u);
    std::vector<double> a((size_t)TD * kd, 0.0), b((size_t)TD * kd, 0.0);
    const double *src = s.coefs(u);
    double *bufs[2] = {a.data(), b.data()};
    for (int dim = 0; dim < D; dim++) {
        const int bit = 1 << dim;
        const int nin = scaling_only ? (1 << dim) : TD; // inputs available at this pass
        const bool last = (dim == D - 1);
        double *dst = bufs[dim & 1];
        for (int gt = 0; gt < TD; gt++) {
            bool acc = false;
            for (int ft = 0; ft < nin; ft++) {
                if ((gt | bit) != (ft | bit)) continue;
                const int f = 2 * ((gt >> dim) & 1) + ((ft >> dim) & 1);
                double *out = dst + (size_t)gt * kd;
                if (last && D == 3) {
                    out = s.coefs(c0 + gt);
                    core::filter_pass(out, src + (size_t)ft * kd, fb.sub(f, Reconstruction), m, acc || !overwrite);
                } else {
                    core::filter_pass(out, src + (size_t)ft * kd, fb.sub(f, Reconstruction), m, acc);
                }
                acc = true;
            }
        }
        src = dst;
    }
    if (D < 3) {
        for (int c = 0; c < TD; c++) {
            double *o = s.coefs(c0 + c);
            const double *x = src + (size_t)c * kd;
            for (int i = 0; i < kd; i++) o[i] = overwrite ? x[i] : o[i] + x[i];
        }
    }
}

// Scaling blocks of the 2^D children of a node from coefficients held outside
// any store: all 2^D blocks of in, or its scaling block only (a generated
// node), written to out[TD][kd] (tree_utils::mw_transform,
// src/utils/tree_utils.cpp:113-216, with the passes of reconstruct_into_children).
template <int D> void children_scaling_blocks(const NodeStore<D> &s, const double *in, bool scaling_only, double *out) {
    constexpr int TD = NodeStore<D>::TD;
    const int kd = s.kp1_d(), m = kd / s.kp1();
    const auto &fb = s.mra().filter();
    std::vector<double> a((size_t)TD * kd, 0.0), b((size_t)TD * kd, 0.0);
    const double *src = in;
    double *bufs[2] = {a.data(), b.data()};
    for (int dim = 0; dim < D; dim++) {
        const int bit = 1 << dim;
        const int nin = scaling_only ? (1 << dim) : TD;
        double *dst = (dim == D - 1) ? out : bufs[dim & 1];
        for (int gt = 0; gt < TD; gt++) {
            bool acc = false;
            for (int ft = 0; ft < nin; ft++) {
                if ((gt | bit) != (ft | bit)) continue;
                const int f = 2 * ((gt >> dim) & 1) + ((ft >> dim) & 1);
                core::filter_pass(dst + (size_t)gt * kd, src + (size_t)ft * kd, fb.sub(f, Reconstruction), m, acc);
                acc = true;
            }
        }
        src = dst;
    }
}

// MWNode::giveChildrenCoefs (src/trees/MWNode.cpp:312-335)
template <int D> void give_children_coefs(NodeStore<D> &s, int u, bool overwrite) {
    const int c0 = s.child0(u);
    if (overwrite)
        for (int c = 0; c < NodeStore<D>::TD; c++) s.zeroCoefs(c0 + c);
    reconstruct_into_children(s, u, overwrite);
    for (int c = 0; c < NodeStore<D>::TD; c++) {
        s.setHasCoefs(c0 + c, true);
        compute_norms(s, c0 + c);
    }
}

// Children scaling blocks -> parent s/d (MWNode::reCompress and
// FunctionNode<3>::reCompress via tree_utils::mw_transform_back).
template <int D> void recompress(NodeStore<D> &s, int u) {
    const int kd = s.kp1_d(), c0 = s.child0(u);
    double *c = s.coefs(u);
    for (int t = 0; t < NodeStore<D>::TD; t++) std::memcpy(c + (size_t)t * kd, s.coefs(c0 + t), sizeof(double) * kd);
    local_transform(s, u, Compression);
    s.setHasCoefs(u, true);
    compute_norms(s, u);
}

// Node lookup that always succeeds, generating missing nodes from their
// ancestors (MWTree::getNode -> MWNode::retrieveNode).  Serial.
template <int D> int get_or_generate(NodeStore<D> &s, int n, const typename NodeStore<D>::Index &l) {
    int u = s.find(n, l);
    if (u >= 0) return u;
    int r = s.mra().getWorldBox().rootIndex(n, l);
    if (r < 0) throw Error("node outside the world box");
    u = s.roots()[r];
    while (s.scale(u) < n) {
        if (!s.hasChildren(u)) {
            bool gen = !s.isOperator();
            s.createChildren(u, gen);
            give_children_coefs(s, u, true);
        }
        const int rel = n - s.scale(u) - 1;
        int c = 0;
        for (int d = 0; d < D; d++) c |= ((l[d] >> rel) & 1) << d;
        u = s.child0(u) + c;
    }
    return u;
}

// Depth-first listing of real nodes, roots in box order and children in index
// order (tree_utils::make_node_table uses a Hilbert child order; only the
// summation order of tree norms depends on it).
template <int D> std::vector<int> real_nodes(const NodeStore<D> &s) {
    std::vector<int> out, stack;
    for (auto it = s.roots().rbegin(); it != s.roots().rend(); ++it) stack.push_back(*it);
    while (!stack.empty()) {
        int u = stack.back();
        stack.pop_back();
        out.push_back(u);
        if (s.hasRealChildren(u))
            for (int c = NodeStore<D>::TD - 1; c >= 0; c--) stack.push_back(s.child0(u) + c);
    }
    return out;
}

template <int D> std::vector<std::vector<int>> real_nodes_by_depth(const NodeStore<D> &s) {
    std::vector<std::vector<int>> out;
    for (int u : real_nodes(s)) {
        int d = s.depth(u);
        if (d >= (int)out.size()) out.resize(d + 1);
        out[d].push_back(u);
    }
    return out;
}

template <int D> std::vector<int> end_nodes(const NodeStore<D> &s) {
    std::vector<int> out;
    for (int u : real_nodes(s))
        if (s.isEnd(u)) out.push_back(u);
    return out;
}

// MWTree::mwTransformUp (src/trees/MWTree.cpp:166-181)
template <int D> void transform_up(NodeStore<D> &s) {
    auto tab = real_nodes_by_depth(s);
    for (int d = (int)tab.size() - 2; d >= 0; d--) {
        const auto &v = tab[d];
        const int nn = (int)v.size();
#pragma omp parallel for schedule(guided) if (!s.isOperator() && nn > 16)
        for (int i = 0; i < nn; i++)
            if (s.hasRealChildren(v[i])) recompress(s, v[i]);
    }
}

// MWTree::calcSquareNorm (src/trees/MWTree.cpp:109-118)
template <int D> double calc_square_norm(NodeStore<D> &s) {
    double t = 0.0;
    for (int u : end_nodes(s)) t += s.squareNorm(u);
    s.treeSquareNorm = t;
    return t;
}

// ---------------------------------------------------------------------------
// Breadth-first tree builder (src/treebuilders/TreeBuilder.cpp:38-87)
// ---------------------------------------------------------------------------
template <int D> struct BuildHooks {
    std::function<std::vector<int>(NodeStore<D> &)> initial;
    std::function<void(NodeStore<D> &, const std::vector<int> &)> compute; // whole level
    std::function<bool(const NodeStore<D> &, int)> split;
};

template <int D> std::vector<int> build_tree(NodeStore<D> &s, const BuildHooks<D> &h, int maxIter) {
    std::vector<int> work = h.initial(s), levels;
    double sNorm = 0.0, wNorm = 0.0;
    const int maxScale = s.mra().getMaxScale();
    for (int iter = 0; !work.empty(); iter++) {
        levels.push_back((int)work.size());
        h.compute(s, work);
        if (iter == 0) {
            sNorm = 0.0;
            for (int u : work)
                if (s.depth(u) >= 0) sNorm += s.scalingNorm(u);
        }
        // the step's wavelet norms summed first, then added once
        // (wNorm += calcWaveletNorm(*workVec), src/treebuilders/TreeBuilder.cpp:58)
        double stepW = 0.0;
        for (int u : work)
            if (s.depth(u) >= 0) stepW += s.waveletNorm(u);
        wNorm += stepW;
        s.treeSquareNorm = (sNorm < 0.0 || wNorm < 0.0) ? -1.0 : sNorm + wNorm;
        if (iter >= maxIter && maxIter >= 0) work.clear();
        // TreeAdaptor::splitNodeVector (src/treebuilders/TreeAdaptor.h:41-52)
        std::vector<int> next;
        for (int u : work) {
            if (s.hasChildren(u)) continue;
            if (s.scale(u) + 2 > maxScale) continue;
            if (h.split(s, u)) {
                int c0 = s.createChildren(u, false);
                for (int c = 0; c < NodeStore<D>::TD; c++) next.push_back(c0 + c);
            }
        }
        work.swap(next);
    }
    return levels;
}

// tree_utils::split_check (src/utils/tree_utils.cpp:47-65), split factor 1
// Wavelet-norm threshold of the split check at a scale.
inline double wavelet_scale_factor(int scale) { return std::pow(2.0, -(0.5 * 1.0 * (scale + 1))); }

inline double wavelet_threshold(double prec, double tree_sq_norm, bool abs_prec, int scale) {
    double t_norm = 1.0;
    if (tree_sq_norm > 0.0 && !abs_prec) t_norm = std::sqrt(tree_sq_norm);
    const double scale_fac = wavelet_scale_factor(scale);
    return std::max(2.0 * MachinePrec, prec * t_norm * scale_fac);
}

inline bool wavelet_split(double prec, double tree_sq_norm, bool abs_prec, int scale, double wavelet_sq_norm) {
    if (prec <= 0.0) return false;
    return std::sqrt(wavelet_sq_norm) > wavelet_threshold(prec, tree_sq_norm, abs_prec, scale);
}

} // namespace detail
} // namespace mrcpp
