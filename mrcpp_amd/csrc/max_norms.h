// Largest scaled norms of a tree at any node index, as MRCPP's getNode +
// getMaxSquareNorm / getMaxWSquareNorm see them (src/trees/MWNode.h:84-85,
// src/trees/MWNode.cpp:1257-1268, src/trees/MWTree.cpp:536-543): for a real
// node, the largest 2^(3n) |c|^2 (and 2^(3n) |w|^2) of the node and its real
// descendants; for an index below the tree's leaves (MWTree::getNode
// generates it from its nearest ancestor), the generated node's own scaled
// square norm (its scaling block only) and a zero wavelet norm.  Generated
// nodes are kept in a private cache, so the tree itself is not changed.
#pragma once

#include <array>
#include <cmath>
#include <map>
#include <vector>

#include "trees.h"

namespace mrcpp {
namespace detail {

class MaxNorms {
public:
    struct Info {
        double maxSq;  // getMaxSquareNorm
        double maxWSq; // getMaxWSquareNorm
        bool leaf;     // isLeafNode (no children)
    };

    explicit MaxNorms(const NodeStore<3> &s)
            : s_(&s) {
        maxSq_.assign(s.size(), -1.0);
        maxWSq_.assign(s.size(), -1.0);
        const std::vector<int> order = real_nodes(s); // parents before children
        for (auto it = order.rbegin(); it != order.rend(); ++it) {
            const int u = *it;
            double m = scaled(s.scale(u), s.squareNorm(u));
            double w = scaled(s.scale(u), s.waveletNorm(u));
            if (s.hasRealChildren(u))
                for (int c = 0; c < 8; c++) {
                    m = std::max(m, maxSq_[s.child0(u) + c]);
                    w = std::max(w, maxWSq_[s.child0(u) + c]);
                }
            maxSq_[u] = m;
            maxWSq_[u] = w;
        }
    }

    Info at(int n, const std::array<int, 3> &l) {
        const NodeStore<3> &s = *s_;
        const int u = s.findReal(n, l);
        if (u >= 0) {
            const double m = maxSq_[u] > 0.0 ? maxSq_[u] : scaled(n, s.squareNorm(u));
            const double w = maxWSq_[u] > 0.0 ? maxWSq_[u] : scaled(n, s.waveletNorm(u));
            return {m, w, !s.hasRealChildren(u)};
        }
        return {scaled(n, generated(n, l)), 0.0, true};
    }

private:
    using Key = std::pair<int, std::array<int, 3>>;
    struct Gen {
        double sq;
        std::vector<double> block; // scaling block
    };
    const NodeStore<3> *s_;
    std::vector<double> maxSq_, maxWSq_;
    std::map<Key, Gen> gen_;

    static double scaled(int n, double sq) { return std::pow(2.0, 3 * n) * sq; }

    // square norm of the generated node (n, l): its nearest real or generated
    // ancestor's blocks reconstructed down to scale n
    double generated(int n, const std::array<int, 3> &l) {
        const NodeStore<3> &s = *s_;
        auto hit = gen_.find({n, l});
        if (hit != gen_.end()) return hit->second.sq;
        int an = n, real = -1;
        std::array<int, 3> al = l;
        const Gen *g = nullptr;
        for (;;) {
            if (an <= s.rootScale()) throw Error("node outside the world box");
            an--;
            for (int d = 0; d < 3; d++) al[d] >>= 1;
            if ((real = s.findReal(an, al)) >= 0) break;
            auto gi = gen_.find({an, al});
            if (gi != gen_.end()) {
                g = &gi->second;
                break;
            }
        }
        const int kd = s.kp1_d();
        std::vector<double> kids((size_t)8 * kd), parent;
        if (g) parent = g->block;
        for (int sc = an + 1; sc <= n; sc++) {
            const bool scalingOnly = real < 0 || sc > an + 1;
            children_scaling_blocks(s, (sc == an + 1 && real >= 0) ? s.coefs(real) : parent.data(), scalingOnly, kids.data());
            std::array<int, 3> pl;
            for (int d = 0; d < 3; d++) pl[d] = l[d] >> (n - sc + 1);
            for (int c = 0; c < 8; c++) {
                std::array<int, 3> cl;
                for (int d = 0; d < 3; d++) cl[d] = 2 * pl[d] + ((c >> d) & 1);
                Gen ng;
                ng.block.assign(kids.begin() + (size_t)c * kd, kids.begin() + (size_t)(c + 1) * kd);
                double ss = 0.0;
                for (double x : ng.block) ss += x * x;
                const double cn = std::sqrt(ss); // MWNode::calcNorms of a generated node: the scaling component only
                ng.sq = cn * cn;
                gen_.emplace(Key{sc, cl}, std::move(ng));
            }
            std::array<int, 3> ol;
            for (int d = 0; d < 3; d++) ol[d] = l[d] >> (n - sc);
            parent = gen_.at({sc, ol}).block;
        }
        return gen_.at({n, l}).sq;
    }
};

} // namespace detail
} // namespace mrcpp
