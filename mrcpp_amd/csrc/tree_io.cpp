// Tree files: the MADNESS-convention text format of
// FunctionTree::saveTreeTXT / loadTreeTXT (src/trees/FunctionTree.cpp:127-373)
// and a binary node dump for FunctionTree::saveTree / loadTree
// (src/trees/FunctionTree.cpp:375-435).
//
// MRCPP's binary file is its node allocator's memory written out as is (node
// objects with their pointers, repaired by NodeAllocator::reassemble on load),
// so it is readable only by the same MRCPP build.  The binary file here holds
// the same information in a layout of its own: a header with the MRA, then
// every real node depth first (scale, translation, child flag, the 8
// component norms and the square norm, the 2^3 blocks of (k+1)^3 coefficients).
#include <charconv>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <unordered_map>

#include "function_tree.h"

namespace mrcpp {

namespace {

using detail::NodeStore;
constexpr int TD = 8;

// MADNESS orders a node's values z fastest and each index from k down to 0;
// MRCPP x fastest from 0 (FunctionTree.cpp:168-176, :333-340).  Entry i of a
// text record is MRCPP's value map[i]; the map is its own inverse.
std::vector<int> madness_order(int kp1) {
    std::vector<int> map;
    for (int x = kp1 - 1; x >= 0; x--)
        for (int y = kp1 - 1; y >= 0; y--)
            for (int z = kp1 - 1; z >= 0; z--) map.push_back(z * kp1 * kp1 + y * kp1 + x);
    return map;
}

// MWNode::cvTransform (src/trees/MWNode.cpp:448-490) on the 8 blocks of c for
// a node at scale n: Forward maps scaling coefficients of the children to
// values at their quadrature points, Backward the reverse.
void cv_transform(const NodeStore<3> &s, const core::SqMat &map, int n, bool forward, double *c) {
    const int kd = s.kp1_d(), m = kd / s.kp1();
    std::vector<double> tmp((size_t)TD * kd);
    double *src = c, *dst = tmp.data();
    for (int pass = 0; pass < 3; pass++) {
        for (int t = 0; t < TD; t++) core::filter_pass(dst + (size_t)t * kd, src + (size_t)t * kd, map, m, false);
        std::swap(src, dst);
    }
    const auto &sf = s.mra().getWorldBox().getScalingFactors();
    double sf_prod = sf[0] * sf[1] * sf[2];
    if (sf_prod <= MachineZero) sf_prod = 1.0;
    double two_fac = std::pow(2.0, 3 * (n + 1)) / sf_prod;
    two_fac = forward ? std::sqrt(two_fac) : std::sqrt(1.0 / two_fac);
    for (int i = 0; i < TD * kd; i++) c[i] = two_fac * src[i];
}

// MWTree::getNode(idx, create = true): the node, creating real children (zero
// coefficients) on the way down from its root box.
int get_or_create(NodeStore<3> &s, int n, const NodeStore<3>::Index &l) {
    int u = s.find(n, l);
    if (u >= 0 && !s.isGen(u)) return u;
    const int r = s.mra().getWorldBox().rootIndex(n, l);
    if (r < 0) throw Error("tree file: node outside the world box");
    u = s.roots()[r];
    while (s.scale(u) < n) {
        if (!s.hasChildren(u)) {
            const int c0 = s.createChildren(u, false);
            for (int c = 0; c < TD; c++) {
                std::memset(s.coefs(c0 + c), 0, sizeof(double) * s.ncoef());
                s.clearNorms(c0 + c);
            }
        }
        const int rel = n - s.scale(u) - 1;
        int c = 0;
        for (int d = 0; d < 3; d++) c |= ((l[d] >> rel) & 1) << d;
        u = s.child0(u) + c;
    }
    return u;
}

// 2^p for the box arithmetic of the text format, as a double like std::pow
double two_pow(int p) { return std::ldexp(1.0, p); }

// Text of the 8 records of one end node (the reference writer's
// "std::setprecision(14)" numbers: std::to_chars with 14 significant digits in
// the general format prints the same characters).
void append_records(const NodeStore<3> &s, const core::SqMat &cv, const std::vector<int> &map, int L, int u, std::string &text) {
    const int ncoefs = s.kp1_d(), rscale = s.rootScale();
    std::vector<double> values((size_t)TD * ncoefs);
    std::memcpy(values.data(), s.coefs(u), sizeof(double) * values.size());
    const int n = s.scale(u);
    detail::local_transform(s, values.data(), Reconstruction);
    cv_transform(s, cv, n, true, values.data());
    int l[3];
    for (int d = 0; d < 3; d++) l[d] = 2 * (int)(s.transl(u)[d] + two_pow(n) * L);
    char buf[64];
    auto put_int = [&](long long v) {
        auto r = std::to_chars(buf, buf + sizeof(buf), v);
        text.append(buf, r.ptr);
        text.push_back(' ');
    };
    for (int c = 0; c < TD; c++) {
        put_int(n - rscale + 2);
        for (int d = 0; d < 3; d++) put_int(l[d] + ((c >> d) & 1));
        text.push_back('\n');
        for (int i = 0; i < ncoefs; i++) {
            auto r = std::to_chars(buf, buf + sizeof(buf), values[(size_t)c * ncoefs + map[i]], std::chars_format::general, 14);
            text.append(buf, r.ptr);
            text.push_back(' ');
        }
        text.push_back('\n');
    }
}

// whitespace-separated number at p (leading '+' accepted, as operator>> does)
template <typename T> const char *parse_number(const char *p, const char *end, T &v) {
    while (p < end && std::isspace((unsigned char)*p)) p++;
    if (p < end && *p == '+') p++;
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) throw Error("loadTreeTXT: malformed number");
    return r.ptr;
}

} // namespace

void FunctionTree<3>::saveTreeTXT(const std::string &fname) {
    NodeStore<3> &s = nodes();
    std::ofstream out(fname, std::ios::binary);
    if (!out) throw Error("Unable to open file " + fname);
    std::ostringstream head;
    head << std::setprecision(14);
    head << 3 << std::endl;
    const int rscale = s.rootScale();
    const auto &sf = mra_.getWorldBox().getScalingFactors();
    double lm = 1.0;
    for (int i = 0; i > rscale; i--) lm *= 2;
    for (int d = 0; d < 3; d++) head << -sf[d] * lm << " " << sf[d] * lm << std::endl;
    head << s.kp1() << std::endl;
    const std::vector<int> ends = detail::end_nodes(s);
    head << TD * ends.size() << std::endl;
    out << head.str();
    const std::vector<int> map = madness_order(s.kp1());
    const core::SqMat &cv = mra_.basis().cvMap();
    const int L = (int)two_pow(-rscale);
    // end nodes formatted in parallel, a block at a time, written in order
    constexpr int kBlock = 512;
    std::vector<std::string> text(kBlock);
    for (size_t b0 = 0; b0 < ends.size(); b0 += kBlock) {
        const int nb = (int)std::min<size_t>(kBlock, ends.size() - b0);
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nb; j++) {
            text[j].clear();
            append_records(s, cv, map, L, ends[b0 + j], text[j]);
        }
        for (int j = 0; j < nb; j++) out.write(text[j].data(), (std::streamsize)text[j].size());
    }
    if (!out) throw Error("write failed: " + fname);
}

void FunctionTree<3>::loadTreeTXT(const std::string &fname) {
    std::ifstream in(fname, std::ios::binary | std::ios::ate);
    if (!in) throw Error("Unable to open file " + fname);
    std::string text((size_t)in.tellg(), '\0');
    in.seekg(0);
    in.read(&text[0], (std::streamsize)text.size());
    if (!in) throw Error("loadTreeTXT: read failed: " + fname);
    const char *p = text.data(), *end = text.data() + text.size();
    int ndim = 0, kp1 = 0;
    p = parse_number(p, end, ndim);
    if (ndim != 3) throw Error("loadTreeTXT: the file holds a " + std::to_string(ndim) + "D function");
    double coord[3][2];
    for (int d = 0; d < 3; d++) {
        p = parse_number(p, end, coord[d][0]);
        p = parse_number(p, end, coord[d][1]);
    }
    const int rscale = mra_.getRootScale();
    int L = 1;
    for (int i = rscale; i < 0; i++) L *= 2;
    // the world must be the cube [-L, L] of this MRA (FunctionTree.cpp:144-153)
    for (int d = 0; d < 3; d++)
        if (std::abs(coord[d][0] + L) > 1.0e-14 || std::abs(coord[d][1] - L) > 1.0e-14)
            throw Error("loadTreeTXT: the file's world box is not [-L, L] of this MRA");
    p = parse_number(p, end, kp1);
    if (kp1 != mra_.getKp1()) throw Error("loadTreeTXT: the file's order differs from this MRA's");
    const int ncoefs = mra_.getKp1_d();
    const std::vector<int> map = madness_order(kp1);
    long long nread = 0;
    p = parse_number(p, end, nread);
    if (nread < 0) throw Error("loadTreeTXT: bad record count");

    // records: 4 integers and ncoefs values each.  Written two lines a
    // record (both writers do), the value lines are parsed in parallel;
    // otherwise token by token.
    std::vector<std::array<int, 4>> ids((size_t)nread);
    std::vector<double> vals((size_t)nread * ncoefs);
    std::vector<const char *> lines;
    lines.reserve(2 * (size_t)nread + 1);
    for (const char *q = p; q < end;) {
        const char *nl = static_cast<const char *>(std::memchr(q, '\n', end - q));
        const char *e = nl ? nl : end;
        bool blank = true;
        for (const char *t = q; t < e; t++)
            if (!std::isspace((unsigned char)*t)) {
                blank = false;
                break;
            }
        if (!blank) lines.push_back(q);
        q = nl ? nl + 1 : end;
    }
    if (lines.size() == 2 * (size_t)nread) {
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
        for (long long r = 0; r < nread; r++) {
            try {
                const char *q = lines[2 * r];
                for (int t = 0; t < 4; t++) q = parse_number(q, end, ids[r][t]);
                q = lines[2 * r + 1];
                for (int i = 0; i < ncoefs; i++) q = parse_number(q, end, vals[(size_t)r * ncoefs + i]);
            } catch (const Error &) {
                bad = 1;
            }
        }
        if (bad) throw Error("loadTreeTXT: malformed record");
    } else {
        for (long long r = 0; r < nread; r++) {
            for (int t = 0; t < 4; t++) p = parse_number(p, end, ids[r][t]);
            for (int i = 0; i < ncoefs; i++) p = parse_number(p, end, vals[(size_t)r * ncoefs + i]);
        }
    }

    clear();
    NodeStore<3> &s = *store_;
    for (int u : s.roots()) s.zeroCoefs(u);
    std::unordered_map<int, int> mp; // values read (or children completed) per node
    std::vector<std::vector<int>> table(64);
    int nmax = 0;
    for (long long r = 0; r < nread; r++) {
        const int n_in = ids[r][0];
        const std::array<int, 3> l_in = {ids[r][1], ids[r][2], ids[r][3]};
        std::array<int, 3> l, lp;
        if (n_in < 2 || n_in >= (int)table.size()) throw Error("loadTreeTXT: scale out of range");
        const int n = n_in + rscale - 1; // scale of the record's box; values sit in its parent
        for (int d = 0; d < 3; d++) {
            l[d] = (int)(l_in[d] - two_pow(n) * L);
            lp[d] = (int)(l_in[d] / 2 - two_pow(n - 1) * L);
        }
        const int u = get_or_create(s, n - 1, lp);
        int c_ix = 0;
        for (int d = 0; d < 3; d++)
            if (std::abs(l[d]) % 2 == 1) c_ix += 1 << d;
        if (mp[u] == 0) {
            s.zeroCoefs(u);
            const int pu = s.parent(u);
            if (pu >= 0)
                for (int c = 0; c < TD; c++)
                    if (mp[s.child0(pu) + c] == 0) s.zeroCoefs(s.child0(pu) + c);
        }
        double *values = s.coefs(u) + (size_t)c_ix * ncoefs;
        const double *src = vals.data() + (size_t)r * ncoefs;
        for (int i = 0; i < ncoefs; i++) values[map[i]] = src[i];
        mp[u]++;
        nmax = std::max(nmax, n_in);
        if (mp[u] == 1) table[n_in].push_back(u);
    }
    // values -> scaling coefficients of the children
    const core::SqMat &vc = mra_.basis().vcMap();
    {
        std::vector<int> all;
        for (int n = nmax; n > -1; n--) all.insert(all.end(), table[n].begin(), table[n].end());
#pragma omp parallel for schedule(dynamic, 16)
        for (size_t i = 0; i < all.size(); i++) {
            cv_transform(s, vc, s.scale(all[i]), false, s.coefs(all[i]));
            detail::compute_norms(s, all[i]);
        }
    }
    auto complete = [&mp](int v) {
        auto it = mp.find(v);
        return it != mp.end() && it->second == TD;
    };
    // complete nodes: compress and pass the scaling block up; a missing child of
    // a complete parent takes its scaling block from the parent.  The nodes of
    // a level are transformed in parallel, the bookkeeping follows in order.
    for (int n = nmax; n > -1; n--) {
        const std::vector<int> &lv = table[n];
#pragma omp parallel for schedule(dynamic, 16)
        for (size_t i = 0; i < lv.size(); i++) {
            const int u = lv[i];
            if (!complete(u)) continue; // incomplete node: left as read
            if (s.hasChildren(u)) {
                for (int c = 0; c < TD; c++) {
                    const int v = s.child0(u) + c;
                    if (complete(v)) continue;
                    double *cc = s.coefs(v);
                    std::memcpy(cc, s.coefs(u) + (size_t)c * ncoefs, sizeof(double) * ncoefs);
                    std::memset(cc + ncoefs, 0, sizeof(double) * (TD - 1) * ncoefs);
                    s.setHasCoefs(v, true);
                    detail::compute_norms(s, v);
                }
            }
            detail::local_transform(s, u, Compression);
            s.setHasCoefs(u, true);
            detail::compute_norms(s, u);
        }
        for (size_t i = 0; i < lv.size(); i++) {
            const int u = lv[i];
            const int p = s.parent(u);
            if (!complete(u) || p < 0) continue;
            if (mp[p] == 0) table[n - 1].push_back(p);
            const int my = u - s.child0(p);
            std::memcpy(s.coefs(p) + (size_t)my * ncoefs, s.coefs(u), sizeof(double) * ncoefs);
            mp[p]++;
        }
    }
    detail::calc_square_norm(s);
}

namespace {
constexpr char kMagic[8] = {'M', 'W', 'T', 'R', 'E', 'E', '3', '\0'};
constexpr int32_t kVersion = 1;

template <typename T> void put(std::ostream &o, const T &v) { o.write(reinterpret_cast<const char *>(&v), sizeof(T)); }
template <typename T> T get(std::istream &i) {
    T v{};
    i.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!i) throw Error("loadTree: truncated file");
    return v;
}
} // namespace

void FunctionTree<3>::saveTree(const std::string &file) {
    NodeStore<3> &s = nodes();
    const std::string fname = file + ".tree";
    std::ofstream out(fname, std::ios::binary);
    if (!out) throw Error("Unable to open file " + fname);
    const auto &box = mra_.getWorldBox();
    out.write(kMagic, sizeof(kMagic));
    put<int32_t>(out, kVersion);
    put<int32_t>(out, mra_.getOrder());
    put<int32_t>(out, mra_.getMaxDepth());
    put<int32_t>(out, box.getScale());
    for (int d = 0; d < 3; d++) put<int32_t>(out, box.getCornerIndex()[d]);
    for (int d = 0; d < 3; d++) put<int32_t>(out, box.getNBoxes()[d]);
    for (int d = 0; d < 3; d++) put<double>(out, box.getScalingFactor(d));
    const std::vector<int> nodes = detail::real_nodes(s);
    put<int64_t>(out, (int64_t)nodes.size());
    const size_t nc = s.ncoef();
    for (int u : nodes) {
        put<int32_t>(out, s.scale(u));
        for (int d = 0; d < 3; d++) put<int32_t>(out, s.transl(u)[d]);
        put<int32_t>(out, s.hasRealChildren(u) ? 1 : 0);
        out.write(reinterpret_cast<const char *>(s.norms(u)), sizeof(double) * TD);
        put<double>(out, s.squareNorm(u));
        out.write(reinterpret_cast<const char *>(s.coefs(u)), sizeof(double) * nc);
    }
    if (!out) throw Error("write failed: " + fname);
}

void FunctionTree<3>::loadTree(const std::string &file) {
    const std::string fname = file + ".tree";
    std::ifstream in(fname, std::ios::binary);
    if (!in) throw Error("Unable to open file " + fname);
    char magic[8];
    in.read(magic, sizeof(magic));
    if (!in || std::memcmp(magic, kMagic, sizeof(kMagic)) != 0) throw Error("loadTree: not a tree file of this library");
    if (get<int32_t>(in) != kVersion) throw Error("loadTree: unknown file version");
    const auto &box = mra_.getWorldBox();
    bool same = get<int32_t>(in) == mra_.getOrder();
    same &= get<int32_t>(in) == mra_.getMaxDepth();
    same &= get<int32_t>(in) == box.getScale();
    for (int d = 0; d < 3; d++) same &= get<int32_t>(in) == box.getCornerIndex()[d];
    for (int d = 0; d < 3; d++) same &= get<int32_t>(in) == box.getNBoxes()[d];
    for (int d = 0; d < 3; d++) same &= get<double>(in) == box.getScalingFactor(d);
    // the tree must have the MRA it was saved with (FunctionTree.cpp:402-404)
    if (!same) throw Error("loadTree: the file's MRA differs from this tree's");
    const int64_t count = get<int64_t>(in);
    clear();
    NodeStore<3> &s = *store_;
    const size_t nc = s.ncoef();
    std::vector<char> seen;
    for (int64_t k = 0; k < count; k++) {
        const int n = get<int32_t>(in);
        NodeStore<3>::Index l;
        for (int d = 0; d < 3; d++) l[d] = get<int32_t>(in);
        get<int32_t>(in); // child flag: the topology follows from the nodes present
        if (n < mra_.getRootScale() || n > mra_.getMaxScale()) throw Error("loadTree: scale out of range");
        const int u = get_or_create(s, n, l);
        in.read(reinterpret_cast<char *>(s.norms(u)), sizeof(double) * TD);
        s.setSquareNorm(u, get<double>(in));
        in.read(reinterpret_cast<char *>(s.coefs(u)), sizeof(double) * nc);
        if (!in) throw Error("loadTree: truncated file");
        s.setHasCoefs(u, true);
    }
    detail::calc_square_norm(s);
}

} // namespace mrcpp
