// MRCPP apply-path ORACLE -- test infrastructure, never product code.
//
// A plain FP64 CPU restatement of MRCPP's convolution-operator application
// path, used only by tests/, the smoke check and bench.py's CPU baseline to
// check and time the MI355X implementation in mrcpp_amd/.  Each routine cites
// the MRCPP file:line whose behaviour it restates (paths relative to the
// MRCPP source tree).  Parallelism is OpenMP over the nodes of one tree level,
// as in src/treebuilders/TreeCalculator.h:39-50.
//
// Where MRCPP delegates arithmetic to Eigen 3.4 (small GEMMs, matrix-vector
// products, lpNorm), the summation order here is a plain sequential loop; the
// reference's exact rounding at those points is not reproducible without
// Eigen and is marked "parity unpinned" in the test documentation.  Node
// traversal uses child-index (Lebesgue) order where MRCPP uses a Hilbert
// path (src/trees/TreeIterator.cpp:55-60); this only changes the summation
// order of tree norms.

#include <omp.h>

#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <iomanip>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// api/constants.h:30-57
constexpr double MachinePrec = 1.0e-15;
constexpr double MachineZero = 1.0e-14;
constexpr int MaxDepth = 30;
constexpr int MaxSepRank = 1000;
constexpr double pi = 3.1415926535897932384626433832795;
constexpr double root_pi = 1.7724538509055160273;

static std::string g_filter_dir;
static std::string g_last_error;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ----------------------------------------------------------------------------
// Legendre polynomials in the power basis (src/functions/LegendrePoly.cpp:62-86)
// ----------------------------------------------------------------------------
static std::mutex g_cache_mtx;

static const std::vector<double> &legendre_coefs(int k) {
    static std::map<int, std::vector<double>> cache;
    {
        std::lock_guard<std::mutex> lk(g_cache_mtx);
        auto it = cache.find(k);
        if (it != cache.end()) return it->second;
    }
    std::vector<double> c(k + 1, 0.0);
    if (k == 0) {
        c[0] = 1.0;
    } else if (k == 1) {
        c[0] = 0.0;
        c[1] = 1.0;
    } else {
        const std::vector<double> &m1 = legendre_coefs(k - 1);
        const std::vector<double> &m2 = legendre_coefs(k - 2);
        double K = (double)k;
        c[0] = -(K - 1.0) * m2[0] / K;
        for (int j = 1; j < k + 1; j++) {
            double cm1 = m1[j - 1];
            if (j <= k - 2) {
                c[j] = (2.0 * K - 1.0) * cm1 / K - (K - 1.0) * m2[j] / K;
            } else {
                c[j] = (2.0 * K - 1.0) * cm1 / K;
            }
        }
    }
    std::lock_guard<std::mutex> lk(g_cache_mtx);
    return cache.emplace(k, std::move(c)).first->second;
}

// LegendrePoly::firstDerivative (src/functions/LegendrePoly.cpp:88-125), on [-1,1]
static void legendre_deriv(int order, double q, double &val, double &dval) {
    if (order == 0) {
        val = 1.0;
        dval = 0.0;
        return;
    }
    if (order == 1) {
        val = q;
        dval = 1.0;
        return;
    }
    double y = q, dy = 1.0, yp = 1.0, dyp = 0.0;
    for (int i = 2; i < order + 1; i++) {
        double c1 = (double)i;
        double c2 = c1 * 2.0 - 1.0;
        double c4 = c1 - 1.0;
        double ym = y;
        y = (c2 * q * y - c4 * yp) / c1;
        yp = ym;
        double dym = dy;
        dy = (c2 * q * dy - c4 * dyp + c2 * yp) / c1;
        dyp = dym;
    }
    val = y;
    dval = dy;
}

// Gauss-Legendre nodes on [0,1] (src/core/GaussQuadrature.cpp:157-194, scaled
// by calcScaledPtsWgts :116-133; QuadratureCache bounds [0,1] :41-45).
struct Quadrature {
    std::vector<double> roots, weights;
};

static const Quadrature &quadrature(int order) {
    static std::map<int, Quadrature> cache;
    {
        std::lock_guard<std::mutex> lk(g_cache_mtx);
        auto it = cache.find(order);
        if (it != cache.end()) return it->second;
    }
    const double EPS = 3.0e-12;
    const int NewtonMaxIter = 10;
    std::vector<double> ur(order), uw(order);
    int K = (order % 2 == 0) ? order / 2 : (order + 1) / 2;
    double a = -1.0, b = 1.0;
    double xm = (b + a) * 0.5, xl = (b - a) * 0.5;
    for (int i = 0; i < K; i++) {
        double z = std::cos(pi * (i + 0.75) / (order + 0.5));
        double p0 = 0, p1 = 0;
        int iter;
        for (iter = 0; iter < NewtonMaxIter; iter++) {
            legendre_deriv(order, z, p0, p1);
            double z1 = z;
            z = z1 - p0 / p1;
            if (std::abs(z - z1) <= EPS) break;
        }
        if (iter == NewtonMaxIter) throw Error("Gauss-Legendre Newton iteration failed");
        ur[i] = xm - xl * z;
        ur[order - 1 - i] = xm + xl * z;
        uw[i] = 2.e0 * xl / ((1.e0 - z * z) * p1 * p1);
        uw[order - 1 - i] = uw[i];
    }
    Quadrature q;
    q.roots.resize(order);
    q.weights.resize(order);
    double A = 0.0, B = 1.0;
    double transl = (B - A) / 1.0;
    double pos = A, hl = transl * 0.5;
    for (int j = 0; j < order; j++) {
        q.roots[j] = ur[j] * hl + pos + hl;
        q.weights[j] = uw[j] * hl;
    }
    std::lock_guard<std::mutex> lk(g_cache_mtx);
    return cache.emplace(order, std::move(q)).first->second;
}

// Legendre scaling basis phi_j(x) = sqrt(2j+1) P_j(2x-1) on [0,1]
// (src/core/LegendreBasis.cpp:67-73) and its value-to-coefficient map
// vcMap(i,j) = phi_j(x_i) w_i (src/core/LegendreBasis.cpp:98-109).
struct Basis {
    int order;
    std::vector<std::vector<double>> funcs; // power-basis coefficients in q = 2x-1
    std::vector<double> vcMap;              // column-major (kp1 x kp1): [i + kp1*j]

    explicit Basis(int k)
            : order(k) {
        for (int j = 0; j <= k; j++) {
            std::vector<double> c = legendre_coefs(j);
            double s = std::sqrt(2.0 * j + 1.0);
            for (auto &x : c) x = s * x;
            funcs.push_back(c);
        }
        int kp1 = k + 1;
        const Quadrature &q = quadrature(kp1);
        vcMap.assign(kp1 * kp1, 0.0);
        for (int j = 0; j < kp1; j++)
            for (int i = 0; i < kp1; i++) vcMap[i + kp1 * j] = evalf(j, q.roots[i]) * q.weights[i];
    }
    // Polynomial::evalf (src/functions/Polynomial.cpp:89-101), bounded to [0,1]
    double evalf(int j, double x) const {
        if (x < 0.0 or x > 1.0) return 0.0;
        const std::vector<double> &c = funcs[j];
        double xp = 1.0, y = 0.0;
        for (size_t k = 0; k < c.size(); k++) {
            y += (xp * c[k]);
            xp *= 2.0 * x - 1.0;
        }
        return y;
    }
};

static const Basis &get_basis(int k) {
    static std::map<int, std::unique_ptr<Basis>> cache;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto it = cache.find(k);
    if (it != cache.end()) return *it->second;
    return *cache.emplace(k, std::make_unique<Basis>(k)).first->second;
}

static std::vector<double> read_doubles(const std::string &path, size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open filter table " + path);
    std::vector<double> v(n);
    f.read(reinterpret_cast<char *>(v.data()), n * sizeof(double));
    if ((size_t)f.gcount() != n * sizeof(double)) throw Error("short filter table " + path);
    return v;
}

// MWFilter (src/core/MWFilter.cpp:200-251): H0/G0 read row-major, H1/G1 by
// Legendre symmetry, sub-filters indexed 2*gtbit+ftbit (:100-133).
struct Filter {
    int order;
    // column-major K x K
    std::vector<double> H0, H1, G0, G1, H0t, H1t, G0t, G1t;
    explicit Filter(int k)
            : order(k) {
        int K = k + 1;
        auto h = read_doubles(g_filter_dir + "/L_H0_" + std::to_string(k), K * K);
        auto g = read_doubles(g_filter_dir + "/L_G0_" + std::to_string(k), K * K);
        H0.assign(K * K, 0);
        G0 = H1 = G1 = H0t = H1t = G0t = G1t = H0;
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                G0[i + K * j] = g[i * K + j];
                H0[i + K * j] = h[i * K + j];
            }
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                G1[i + K * j] = std::pow(-1.0, i + j + K) * G0[i + K * j];
                H1[i + K * j] = std::pow(-1.0, i + j) * H0[i + K * j];
            }
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) {
                G0t[i + K * j] = G0[j + K * i];
                G1t[i + K * j] = G1[j + K * i];
                H0t[i + K * j] = H0[j + K * i];
                H1t[i + K * j] = H1[j + K * i];
            }
    }
    enum { Compression = 0, Reconstruction = 1 };
    const double *sub(int i, int op) const {
        if (op == Compression) {
            const std::vector<double> *t[4] = {&H0t, &H1t, &G0t, &G1t};
            return t[i]->data();
        }
        const std::vector<double> *t[4] = {&H0, &G0, &H1, &G1};
        return t[i]->data();
    }
};

static const Filter &get_filter(int k) {
    static std::map<int, std::unique_ptr<Filter>> cache;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto it = cache.find(k);
    if (it != cache.end()) return *it->second;
    return *cache.emplace(k, std::make_unique<Filter>(k)).first->second;
}

// CrossCorrelation::readCCCBin (src/core/CrossCorrelation.cpp:99-124):
// (K^2) x (2K) row-major, entries below MachinePrec set to zero.
struct CrossCorr {
    int order;
    std::vector<double> L, R; // row-major [row*2K + col]
    explicit CrossCorr(int k)
            : order(k) {
        int K = k + 1;
        L = read_doubles(g_filter_dir + "/L_c_left_" + std::to_string(k), K * K * 2 * K);
        R = read_doubles(g_filter_dir + "/L_c_right_" + std::to_string(k), K * K * 2 * K);
        for (auto &x : L)
            if (std::abs(x) < MachinePrec) x = 0.0;
        for (auto &x : R)
            if (std::abs(x) < MachinePrec) x = 0.0;
    }
};

static const CrossCorr &get_ccc(int k) {
    static std::map<int, std::unique_ptr<CrossCorr>> cache;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto it = cache.find(k);
    if (it != cache.end()) return *it->second;
    return *cache.emplace(k, std::make_unique<CrossCorr>(k)).first->second;
}

// math_utils::apply_filter (src/utils/math_utils.cpp:175-194):
// out(r, j) (=|+=) sum_i in(i, r) * F(i, j), with in viewed (kp1 x kp1_dm1)
// and out (kp1_dm1 x kp1), both column-major.
static inline void apply_filter(double *out, const double *in, const double *F, int kp1, int kp1_dm1, bool add) {
    for (int j = 0; j < kp1; j++) {
        const double *Fj = F + kp1 * j;
        double *oj = out + kp1_dm1 * j;
        for (int r = 0; r < kp1_dm1; r++) {
            const double *ir = in + kp1 * r;
            double s = 0.0;
            for (int i = 0; i < kp1; i++) s += ir[i] * Fj[i];
            if (add)
                oj[r] += s;
            else
                oj[r] = s;
        }
    }
}

static inline int ipow(int m, int e) {
    int r = 1;
    for (int i = 0; i < e; i++) r *= m;
    return r;
}

// ----------------------------------------------------------------------------
// Gaussians (src/functions/Gaussian.cpp:116-135, GaussFunc.cpp:47-66)
// ----------------------------------------------------------------------------
template <int D> struct GaussFunc {
    std::array<double, D> alpha;
    double coef;
    std::array<double, D> pos;
    double evalf(const std::array<double, D> &r) const {
        double q2 = 0.0, p2 = 1.0;
        for (int d = 0; d < D; d++) {
            double q = r[d] - pos[d];
            q2 += alpha[d] * q * q;
        }
        return coef * p2 * std::exp(-q2);
    }
    bool isVisibleAtScale(int scale, int nQuadPts) const {
        for (auto &alp : alpha) {
            double stdDeviation = std::pow(2.0 * alp, -0.5);
            auto visibleScale = static_cast<int>(-std::floor(std::log2(nQuadPts * 0.5 * stdDeviation)));
            if (scale < visibleScale) return false;
        }
        return true;
    }
    bool isZeroOnInterval(const double *a, const double *b) const {
        for (int i = 0; i < D; i++) {
            double stdDeviation = std::pow(2.0 * alpha[i], -0.5);
            double gaussBoxMin = pos[i] - 5.0 * stdDeviation;
            double gaussBoxMax = pos[i] + 5.0 * stdDeviation;
            if (a[i] > gaussBoxMax or b[i] < gaussBoxMin) return true;
        }
        return false;
    }
};

// GaussExp::evalf / isVisibleAtScale / isZeroOnInterval (src/functions/GaussExp.cpp:80-98)
template <int D> struct GaussExp {
    std::vector<GaussFunc<D>> funcs;
    double evalf(const std::array<double, D> &r) const {
        double v = 0.0;
        for (auto &g : funcs) v += g.evalf(r);
        return v;
    }
};

// ----------------------------------------------------------------------------
// MRA and trees (src/trees/BoundingBox.cpp:47-80, MultiResolutionAnalysis.h:53-65)
// ----------------------------------------------------------------------------
template <int D> struct MRA {
    int order = 0, kp1 = 0, kp1_d = 0;
    int n0 = 0;
    std::array<int, D> corner{}, nbox{};
    std::array<double, D> sf{};
    int maxDepth = MaxDepth;
    const Basis *basis = nullptr;
    const Filter *filter = nullptr;

    int maxScale() const { return n0 + maxDepth; }
    int nRoots() const {
        int t = 1;
        for (int d = 0; d < D; d++) t *= nbox[d];
        return t;
    }
    void finish() {
        kp1 = order + 1;
        kp1_d = ipow(kp1, D);
        basis = &get_basis(order);
        filter = &get_filter(order);
    }
    // BoundingBox(std::array<int,2>) (src/trees/BoundingBox.cpp:47-80), uniform in all dims
    static MRA from_box(int lo, int hi, int order, int depth = MaxDepth) {
        MRA m;
        int n = 0;
        double size = 1.0 * hi;
        while (size >= 2.0) {
            size /= 2.0;
            n--;
        }
        m.n0 = n;
        for (int d = 0; d < D; d++) {
            m.corner[d] = (lo == 0) ? 0 : -1;
            m.nbox[d] = (lo == 0) ? 1 : 2;
            m.sf[d] = size;
        }
        m.order = order;
        m.maxDepth = depth;
        m.finish();
        return m;
    }
    double unitLength(int d) const { return sf[d] * std::pow(2.0, -n0); }
    double lowerBound(int d) const { return corner[d] * unitLength(d); }
    double upperBound(int d) const { return lowerBound(d) + unitLength(d) * nbox[d]; }
    // BoundingBox::getNodeIndex (src/trees/BoundingBox.cpp:263-281), x fastest
    std::array<int, D> rootL(int r) const {
        std::array<int, D> l;
        for (int d = 0; d < D; d++) {
            l[d] = r % nbox[d] + corner[d];
            r /= nbox[d];
        }
        return l;
    }
    // BoundingBox::getBoxIndex(NodeIndex) (src/trees/BoundingBox.cpp:309-330)
    int boxIndex(int n, const std::array<int, D> &l) const {
        int rel = n - n0;
        if (rel < 0) return -1;
        int b = 0, nc = 1;
        for (int d = 0; d < D; d++) {
            int t = (l[d] >> rel) - corner[d];
            if (t < 0 or t >= nbox[d]) return -1;
            b += nc * t;
            nc *= nbox[d];
        }
        return b;
    }
};

template <int D> struct Tree;

template <int D> struct Node {
    static constexpr int TD = 1 << D;
    using Index = std::array<int, D>;
    Tree<D> *tree = nullptr;
    Node *parent = nullptr;
    std::atomic<Node *> kids{nullptr}; // array of TD children (real or generated)
    int n = 0;
    std::array<int, D> l{};
    bool gen = false;
    bool has_coefs = false;
    std::vector<double> coefs;
    double cn[TD];
    double sq = -1.0;
    double maxSq = -1.0;  // MWNode::maxSquareNorm (src/trees/MWNode.h:180)
    double maxWSq = -1.0; // MWNode::maxWSquareNorm
    omp_lock_t lock;

    Node() {
        omp_init_lock(&lock);
        clearNorms();
    }
    ~Node() {
        Node *k = kids.load();
        delete[] k;
        omp_destroy_lock(&lock);
    }
    int depth() const;
    bool isBranch() const { return kids.load(std::memory_order_acquire) != nullptr; }
    bool hasRealKids() const {
        Node *k = kids.load(std::memory_order_acquire);
        return k != nullptr and not k[0].gen;
    }
    bool isEnd() const { return not gen and not hasRealKids(); }
    Node &child(int i) { return kids.load(std::memory_order_acquire)[i]; }
    void clearNorms() {
        sq = -1.0;
        for (int i = 0; i < TD; i++) cn[i] = -1.0;
    }
    void zeroNorms() {
        sq = 0.0;
        for (int i = 0; i < TD; i++) cn[i] = 0.0;
    }
    void zeroCoefs() {
        std::fill(coefs.begin(), coefs.end(), 0.0);
        zeroNorms();
        has_coefs = true;
    }
    std::array<int, D> childL(int c) const {
        std::array<int, D> r;
        for (int d = 0; d < D; d++) r[d] = 2 * l[d] + ((c >> d) & 1);
        return r;
    }
    double scalingNorm() const { return (cn[0] >= 0.0) ? cn[0] * cn[0] : -1.0; }
    // MWNode::getWaveletNorm (src/trees/MWNode.cpp:629-640)
    double waveletNorm() const {
        double w = 0.0;
        for (int i = 1; i < TD; i++) {
            double ni = cn[i];
            if (ni >= 0.0)
                w += ni * ni;
            else
                w = -1.0;
        }
        return w;
    }
};

template <int D> struct Tree {
    const MRA<D> *mra;
    std::vector<std::unique_ptr<Node<D>>> roots;
    double squareNorm = -1.0;
    bool isOperator = false;
    double normPrec = 0.0; // OperatorTree only
    std::vector<int> nodesAtDepth{0};

    explicit Tree(const MRA<D> &m, bool oper = false, double np = 0.0)
            : mra(&m)
            , isOperator(oper)
            , normPrec(np) {
        int nr = m.nRoots();
        for (int r = 0; r < nr; r++) {
            auto nd = std::make_unique<Node<D>>();
            nd->tree = this;
            nd->n = m.n0;
            nd->l = m.rootL(r);
            nd->coefs.assign((size_t)Node<D>::TD * m.kp1_d, 0.0);
            incrementNodeCount(nd->n);
            roots.push_back(std::move(nd));
        }
    }
    int rootScale() const { return mra->n0; }
    int getDepth() const { return (int)nodesAtDepth.size(); }
    // MWTree::incrementNodeCount (src/trees/MWTree.cpp:238-252), non-negative depths only
    void incrementNodeCount(int scale) {
        int depth = scale - rootScale();
        int n = (int)nodesAtDepth.size() - 1;
        if (depth > n)
            for (int i = 0; i < depth - n; i++) nodesAtDepth.push_back(0);
        nodesAtDepth[depth]++;
    }
};

template <int D> int Node<D>::depth() const { return n - tree->rootScale(); }

// MWNode::calcComponentNorm (src/trees/MWNode.cpp:644-655)
template <int D> static double function_component_norm(const Node<D> &nd, int i) {
    if (nd.gen and i != 0) return 0.0;
    int size = nd.tree->mra->kp1_d;
    const double *c = nd.coefs.data() + i * size;
    double s = 0.0;
    for (int j = 0; j < size; j++) s += c[j] * c[j];
    return std::sqrt(s);
}

// OperatorNode::calcComponentNorm (src/trees/OperatorNode.cpp:56-80) with
// math_utils::matrix_norm_{inf,1,2} (src/utils/math_utils.cpp:60-78);
// lpNorm<2> of a matrix is the Frobenius norm.
static double operator_component_norm(const Node<2> &nd, int i) {
    int depth = nd.depth();
    double prec = nd.tree->normPrec;
    double thrs = std::max(MachinePrec, prec / (8.0 * (1 << depth)));
    int kp1 = nd.tree->mra->kp1;
    int kp1_d = kp1 * kp1;
    const double *v = nd.coefs.data() + i * kp1_d;
    double vecNorm = 0.0;
    for (int j = 0; j < kp1_d; j++) vecNorm += v[j] * v[j];
    vecNorm = std::sqrt(vecNorm);
    double norm = 0.0;
    if (vecNorm > thrs) {
        double infNorm = 0.0, oneNorm = 0.0;
        for (int r = 0; r < kp1; r++) {
            double s = 0.0;
            for (int c = 0; c < kp1; c++) s += std::abs(v[r + kp1 * c]);
            infNorm = std::max(infNorm, s);
        }
        for (int c = 0; c < kp1; c++) {
            double s = 0.0;
            for (int r = 0; r < kp1; r++) s += std::abs(v[r + kp1 * c]);
            oneNorm = std::max(oneNorm, s);
        }
        if (std::sqrt(infNorm * oneNorm) > thrs) {
            double twoNorm = vecNorm; // Frobenius
            if (twoNorm > thrs) norm = twoNorm;
        }
    }
    return norm;
}

// MWNode::calcNorms (src/trees/MWNode.cpp:609-616)
template <int D> static void calc_norms(Node<D> &nd) {
    nd.sq = 0.0;
    for (int i = 0; i < Node<D>::TD; i++) {
        double ni;
        if constexpr (D == 2) {
            ni = nd.tree->isOperator ? operator_component_norm(nd, i) : function_component_norm(nd, i);
        } else {
            ni = function_component_norm(nd, i);
        }
        nd.cn[i] = ni;
        nd.sq += ni * ni;
    }
}

// MWNode::mwTransform (src/trees/MWNode.cpp:557-594): node-local transform
// between the 2^D child scaling blocks and the s/d blocks.
template <int D> static void node_mw_transform(Node<D> &nd, int op) {
    const MRA<D> &m = *nd.tree->mra;
    int kp1 = m.kp1, kp1_dm1 = ipow(kp1, D - 1), kp1_d = m.kp1_d;
    constexpr int TD = Node<D>::TD;
    std::vector<double> o_vec(TD * kp1_d);
    double *out_vec = o_vec.data();
    double *in_vec = nd.coefs.data();
    for (int i = 0; i < D; i++) {
        int mask = 1 << i;
        for (int gt = 0; gt < TD; gt++) {
            double *out = out_vec + gt * kp1_d;
            bool add = false;
            for (int ft = 0; ft < TD; ft++) {
                if ((gt | mask) == (ft | mask)) {
                    const double *in = in_vec + ft * kp1_d;
                    int fIdx = 2 * ((gt >> i) & 1) + ((ft >> i) & 1);
                    apply_filter(out, in, m.filter->sub(fIdx, op), kp1, kp1_dm1, add);
                    add = true;
                }
            }
        }
        std::swap(in_vec, out_vec);
    }
    if (D % 2 == 1)
        for (int i = 0; i < TD * kp1_d; i++) nd.coefs[i] = in_vec[i];
}

// tree_utils::mw_transform (src/utils/tree_utils.cpp:113-216): parent s/d
// (or scaling only, for generated parents) -> child scaling blocks.
template <int D> static void tree_mw_transform(const MRA<D> &m, const double *coeff_in, Node<D> *kids, bool readOnlyScaling, bool overwrite) {
    int kp1 = m.kp1, kp1_d = m.kp1_d, kp1_dm1 = ipow(kp1, D - 1);
    constexpr int TD = 1 << D;
    std::vector<double> t1(kp1_d * TD, 0.0), t2(kp1_d * TD, 0.0);
    int ftlim = TD, ftlim2 = TD, ftlim3 = TD;
    if (readOnlyScaling) {
        ftlim = 1;
        ftlim2 = 2;
        ftlim3 = 4;
    }
    const double *src = coeff_in;
    std::vector<double> *dst = &t1;
    int lims[3] = {ftlim, ftlim2, ftlim3};
    for (int i = 0; i < D; i++) {
        int mask = 1 << i;
        bool last3 = (D == 3 and i == 2);
        for (int gt = 0; gt < TD; gt++) {
            double *out = last3 ? kids[gt].coefs.data() : dst->data() + gt * kp1_d;
            bool add = last3 ? (not overwrite) : false;
            for (int ft = 0; ft < lims[i]; ft++) {
                if ((gt | mask) == (ft | mask)) {
                    const double *in = src + ft * kp1_d;
                    int fIdx = 2 * ((gt >> i) & 1) + ((ft >> i) & 1);
                    apply_filter(out, in, m.filter->sub(fIdx, Filter::Reconstruction), kp1, kp1_dm1, add);
                    add = true;
                }
            }
        }
        if (not last3) {
            src = dst->data();
            dst = (dst == &t1) ? &t2 : &t1;
        }
    }
    if (D < 3) {
        for (int j = 0; j < TD; j++)
            for (int i = 0; i < kp1_d; i++) {
                if (overwrite)
                    kids[j].coefs[i] = src[i + j * kp1_d];
                else
                    kids[j].coefs[i] += src[i + j * kp1_d];
            }
    }
}

// MWNode::giveChildrenCoefs (src/trees/MWNode.cpp:312-335)
template <int D> static void give_children_coefs(Node<D> &nd, bool overwrite) {
    constexpr int TD = Node<D>::TD;
    Node<D> *k = nd.kids.load(std::memory_order_acquire);
    if (overwrite)
        for (int i = 0; i < TD; i++) k[i].zeroCoefs();
    tree_mw_transform<D>(*nd.tree->mra, nd.coefs.data(), k, nd.gen, overwrite);
    for (int i = 0; i < TD; i++) {
        k[i].has_coefs = true;
        calc_norms(k[i]);
    }
}

// FunctionNode::createChildren / OperatorNode::createChildren
// (src/trees/FunctionNode.cpp:255-291, src/trees/OperatorNode.cpp:109-142)
template <int D> static Node<D> *make_kids(Node<D> &nd, bool gen) {
    constexpr int TD = Node<D>::TD;
    const MRA<D> &m = *nd.tree->mra;
    Node<D> *k = new Node<D>[TD];
    for (int c = 0; c < TD; c++) {
        k[c].tree = nd.tree;
        k[c].parent = &nd;
        k[c].n = nd.n + 1;
        k[c].l = nd.childL(c);
        k[c].gen = gen;
        k[c].coefs.assign(gen ? m.kp1_d : TD * m.kp1_d, 0.0);
        if (not gen) nd.tree->incrementNodeCount(k[c].n);
    }
    return k;
}

template <int D> static void create_children(Node<D> &nd) {
    if (nd.isBranch()) throw Error("Node already has children");
    nd.kids.store(make_kids(nd, false), std::memory_order_release);
}

// MWNode::threadSafeGenChildren (src/trees/MWNode.cpp:410-418): generated
// children for FunctionTrees (FunctionNode::genChildren :293-328), real
// children for OperatorTrees (OperatorNode::genChildren :144-147); the
// children are published only after their coefficients are computed.
template <int D> static void thread_safe_gen_children(Node<D> &nd) {
    if (nd.isBranch()) return;
    omp_set_lock(&nd.lock);
    if (not nd.isBranch()) {
        bool gen = not nd.tree->isOperator;
        Node<D> *k;
        if (gen) {
            k = make_kids(nd, true);
        } else {
#pragma omp critical(oracle_count)
            k = make_kids(nd, false);
        }
        constexpr int TD = Node<D>::TD;
        for (int i = 0; i < TD; i++) k[i].zeroCoefs();
        tree_mw_transform<D>(*nd.tree->mra, nd.coefs.data(), k, nd.gen, true);
        for (int i = 0; i < TD; i++) {
            k[i].has_coefs = true;
            calc_norms(k[i]);
        }
        nd.kids.store(k, std::memory_order_release);
    }
    omp_unset_lock(&nd.lock);
}

// MWNode::getChildIndex for a descendant index
template <int D> static int child_index(const Node<D> &nd, int n, const typename Node<D>::Index &l) {
    int rel = n - nd.n - 1;
    int c = 0;
    for (int d = 0; d < D; d++) c |= ((l[d] >> rel) & 1) << d;
    return c;
}

// MWTree::getNode (src/trees/MWTree.cpp:340-352) -> MWNode::retrieveNode
// (src/trees/MWNode.cpp:1097-1120): always returns the node, generating it
// from its ancestors when needed.
template <int D> static Node<D> &get_node(Tree<D> &t, int n, const typename Node<D>::Index &l) {
    int r = t.mra->boxIndex(n, l);
    if (r < 0) throw Error("node outside world box");
    Node<D> *nd = t.roots[r].get();
    while (nd->n < n) {
        thread_safe_gen_children(*nd);
        nd = &nd->child(child_index(*nd, n, l));
    }
    return *nd;
}

// MWTree::findNode (src/trees/MWTree.cpp:294-303): real nodes only
template <int D> static Node<D> *find_node(Tree<D> &t, int n, const typename Node<D>::Index &l) {
    int r = t.mra->boxIndex(n, l);
    if (r < 0) return nullptr;
    Node<D> *nd = t.roots[r].get();
    while (nd->n < n) {
        if (not nd->hasRealKids()) return nullptr;
        nd = &nd->child(child_index(*nd, n, l));
    }
    return nd;
}

// Depth-first node table, roots in box order, children in index order
// (tree_utils::make_node_table, src/utils/tree_utils.cpp:69-105); generated
// nodes are excluded.
template <int D> static void collect(Node<D> &nd, std::vector<Node<D> *> &out) {
    out.push_back(&nd);
    if (nd.hasRealKids())
        for (int i = 0; i < Node<D>::TD; i++) collect(nd.child(i), out);
}

template <int D> static std::vector<Node<D> *> node_table(Tree<D> &t) {
    std::vector<Node<D> *> out;
    for (auto &r : t.roots) collect(*r, out);
    return out;
}

template <int D> static std::vector<std::vector<Node<D> *>> node_table_by_depth(Tree<D> &t) {
    std::vector<std::vector<Node<D> *>> out;
    for (Node<D> *nd : node_table(t)) {
        int depth = nd->depth();
        if (depth + 1 > (int)out.size()) out.resize(depth + 1);
        out[depth].push_back(nd);
    }
    return out;
}

template <int D> static std::vector<Node<D> *> end_nodes(Tree<D> &t) {
    std::vector<Node<D> *> out;
    for (Node<D> *nd : node_table(t))
        if (nd->isEnd()) out.push_back(nd);
    return out;
}

// MWNode::reCompress / FunctionNode<3>::reCompress (src/trees/MWNode.cpp:658-667,
// src/trees/FunctionNode.cpp:391-409 via tree_utils::mw_transform_back
// :229-301, which performs the same three compression passes).
template <int D> static void re_compress(Node<D> &nd) {
    if (not nd.hasRealKids()) return;
    int kp1_d = nd.tree->mra->kp1_d;
    for (int c = 0; c < Node<D>::TD; c++) std::copy(nd.child(c).coefs.begin(), nd.child(c).coefs.begin() + kp1_d, nd.coefs.begin() + c * kp1_d);
    node_mw_transform(nd, Filter::Compression);
    nd.has_coefs = true;
    calc_norms(nd);
}

// MWTree::mwTransformUp (src/trees/MWTree.cpp:166-181; serial variant for
// operators, src/trees/OperatorTree.cpp:254-265)
template <int D> static void mw_transform_up(Tree<D> &t) {
    auto table = node_table_by_depth(t);
    for (int n = (int)table.size() - 2; n >= 0; n--) {
        auto &v = table[n];
        int nn = (int)v.size();
#pragma omp parallel for schedule(guided) if (not t.isOperator)
        for (int i = 0; i < nn; i++)
            if (v[i]->hasRealKids()) re_compress(*v[i]);
    }
}

// MWTree::mwTransformDown (src/trees/MWTree.cpp:194-216), non-negative scales
template <int D> static void mw_transform_down(Tree<D> &t, bool overwrite) {
    auto table = node_table_by_depth(t);
    for (auto &v : table) {
        int nn = (int)v.size();
#pragma omp parallel for schedule(guided)
        for (int i = 0; i < nn; i++)
            if (v[i]->hasRealKids()) give_children_coefs(*v[i], overwrite);
    }
}

// MWTree::calcSquareNorm (src/trees/MWTree.cpp:109-118)
template <int D> static void calc_square_norm(Tree<D> &t) {
    double s = 0.0;
    for (Node<D> *nd : end_nodes(t)) s += nd->sq;
    t.squareNorm = s;
}

// FunctionTree::deleteGenerated: drop generated children everywhere
template <int D> static void delete_generated(Node<D> &nd) {
    Node<D> *k = nd.kids.load();
    if (k == nullptr) return;
    if (k[0].gen) {
        nd.kids.store(nullptr);
        delete[] k;
        return;
    }
    for (int i = 0; i < Node<D>::TD; i++) delete_generated(k[i]);
}

template <int D> static void delete_generated(Tree<D> &t) {
    for (auto &r : t.roots) delete_generated(*r);
}

// ----------------------------------------------------------------------------
// TreeBuilder (src/treebuilders/TreeBuilder.cpp:38-87, 150-166)
// ----------------------------------------------------------------------------
template <int D> struct Calculator {
    virtual ~Calculator() = default;
    // TreeCalculator::getInitialWorkVector default: end nodes (TreeCalculator.h:37)
    virtual std::vector<Node<D> *> initial(Tree<D> &t) { return end_nodes(t); }
    virtual void calcNode(Node<D> &nd) = 0;
    virtual bool parallel() const { return true; }
    virtual void calcVector(std::vector<Node<D> *> &v) {
        int nn = (int)v.size();
        if (parallel()) {
#pragma omp parallel for schedule(guided)
            for (int i = 0; i < nn; i++) calcNode(*v[i]);
        } else {
            for (int i = 0; i < nn; i++) calcNode(*v[i]);
        }
    }
};

template <int D> struct Adaptor {
    int maxScale;
    explicit Adaptor(int ms)
            : maxScale(ms) {}
    virtual ~Adaptor() = default;
    virtual bool split(const Node<D> &nd) const = 0;
    // TreeAdaptor::splitNodeVector (src/treebuilders/TreeAdaptor.h:41-52)
    void splitVector(std::vector<Node<D> *> &out, std::vector<Node<D> *> &inp) const {
        for (Node<D> *nd : inp) {
            if (nd->isBranch()) continue;
            if (nd->n + 2 > maxScale) continue;
            if (split(*nd)) {
                create_children(*nd);
                for (int i = 0; i < Node<D>::TD; i++) out.push_back(&nd->child(i));
            }
        }
    }
};

struct BuildStats {
    std::vector<int> level_nodes;
};

template <int D> static void build(Tree<D> &t, Calculator<D> &calc, Adaptor<D> &adaptor, int maxIter, BuildStats *st = nullptr) {
    std::vector<Node<D> *> work = calc.initial(t);
    double sNorm = 0.0, wNorm = 0.0;
    int iter = 0;
    while (not work.empty()) {
        if (st) st->level_nodes.push_back((int)work.size());
        calc.calcVector(work);
        if (iter == 0) {
            sNorm = 0.0;
            for (Node<D> *nd : work)
                if (nd->depth() >= 0) sNorm += nd->scalingNorm();
        }
        // the step's wavelet norms summed first, then added once
        // (wNorm += calcWaveletNorm(*workVec), src/treebuilders/TreeBuilder.cpp:58,156-163)
        double stepW = 0.0;
        for (Node<D> *nd : work)
            if (nd->depth() >= 0) stepW += nd->waveletNorm();
        wNorm += stepW;
        if (sNorm < 0.0 or wNorm < 0.0)
            t.squareNorm = -1.0;
        else
            t.squareNorm = sNorm + wNorm;
        std::vector<Node<D> *> next;
        if (iter >= maxIter and maxIter >= 0) work.clear();
        adaptor.splitVector(next, work);
        work.swap(next);
        iter++;
    }
}

// DefaultCalculator (src/treebuilders/DefaultCalculator.h:34-46)
template <int D> struct DefaultCalculator : Calculator<D> {
    bool parallel() const override { return false; }
    void calcNode(Node<D> &nd) override {
        nd.has_coefs = false;
        nd.clearNorms();
    }
};

// AnalyticAdaptor (src/treebuilders/AnalyticAdaptor.h:42-50) for one GaussFunc
template <int D> struct AnalyticAdaptor : Adaptor<D> {
    const GaussFunc<D> *func;
    AnalyticAdaptor(const GaussFunc<D> &f, int ms)
            : Adaptor<D>(ms)
            , func(&f) {}
    bool split(const Node<D> &nd) const override {
        const MRA<D> &m = *nd.tree->mra;
        if (func->isVisibleAtScale(nd.n, m.kp1)) return false;
        double lb[D], ub[D];
        double two_n = std::pow(2.0, -nd.n);
        for (int d = 0; d < D; d++) {
            lb[d] = m.sf[d] * two_n * nd.l[d];
            ub[d] = m.sf[d] * two_n * (nd.l[d] + 1);
        }
        if (func->isZeroOnInterval(lb, ub)) return false;
        return true;
    }
};

// tree_utils::split_check (src/utils/tree_utils.cpp:47-65)
template <int D> static bool split_check(const Node<D> &nd, double prec, double split_fac, bool abs_prec) {
    bool split = false;
    if (prec > 0.0) {
        double t_norm = 1.0;
        double sq_norm = nd.tree->squareNorm;
        if (sq_norm > 0.0 and not abs_prec) t_norm = std::sqrt(sq_norm);
        double scale_fac = 1.0;
        if (split_fac > MachineZero) {
            double expo = 0.5 * split_fac * (nd.n + 1);
            scale_fac = std::pow(2.0, -expo);
        }
        double w_thrs = std::max(2.0 * MachinePrec, prec * t_norm * scale_fac);
        double w_norm = std::sqrt(nd.waveletNorm());
        if (w_norm > w_thrs) split = true;
    }
    return split;
}

// Local precision factor of a node (the precFunc of apply with precTrees,
// src/treebuilders/apply.cpp:222-229): 1 / the largest sqrt(getMaxSquareNorm)
// of the node's index over the precision trees, generating nodes below their
// leaves (MWTree::getNode).  getMaxSquareNorm (src/trees/MWNode.h:84) falls
// back to the node's own scaled square norm 2^(D n) |c|^2 (:212) when no
// positive maximum was set.
template <int D> static double scaled_square_norm(const Node<D> &nd) { return std::pow(2.0, D * nd.n) * nd.sq; }
template <int D> static double max_square_norm(const Node<D> &nd) { return nd.maxSq > 0.0 ? nd.maxSq : scaled_square_norm(nd); }
// MWNode::calcScaledWSquareNorm / getMaxWSquareNorm (src/trees/MWNode.h:85,213)
template <int D> static double scaled_wsquare_norm(const Node<D> &nd) { return std::pow(2.0, D * nd.n) * nd.waveletNorm(); }
template <int D> static double max_wsquare_norm(const Node<D> &nd) { return nd.maxWSq > 0.0 ? nd.maxWSq : scaled_wsquare_norm(nd); }

// MWNode::setMaxSquareNorm (src/trees/MWNode.cpp:1257-1268) from every root
// (MWTree::makeMaxSquareNorms, src/trees/MWTree.cpp:536-543).
template <int D> static void set_max_square_norm(Node<D> &nd) {
    nd.maxSq = scaled_square_norm(nd);
    nd.maxWSq = scaled_wsquare_norm(nd);
    if (nd.isEnd()) return;
    for (int i = 0; i < Node<D>::TD; i++) {
        Node<D> &c = nd.child(i);
        set_max_square_norm(c);
        nd.maxSq = std::max(nd.maxSq, c.maxSq);
        nd.maxWSq = std::max(nd.maxWSq, c.maxWSq);
    }
}
template <int D> static void make_max_square_norms(Tree<D> &t) {
    for (auto &r : t.roots) set_max_square_norm(*r);
}

template <int D> struct PrecFunc {
    std::vector<Tree<D> *> trees;
    double operator()(int n, const typename Node<D>::Index &l) const {
        double maxNorm = trees.empty() ? 1.0 : 0.0;
        for (Tree<D> *t : trees) maxNorm = std::max(maxNorm, std::sqrt(max_square_norm(get_node(*t, n, l))));
        return 1.0 / maxNorm;
    }
};

// WaveletAdaptor (src/treebuilders/WaveletAdaptor.h:36-54)
template <int D> struct WaveletAdaptor : Adaptor<D> {
    double prec;
    bool absPrec;
    const PrecFunc<D> *precFunc = nullptr; // setPrecFunction (:43); 1.0 when unset (:49)
    WaveletAdaptor(double p, int ms, bool ap)
            : Adaptor<D>(ms)
            , prec(p)
            , absPrec(ap) {}
    bool split(const Node<D> &nd) const override {
        const double precFac = precFunc ? (*precFunc)(nd.n, nd.l) : 1.0;
        return split_check(nd, prec * precFac, 1.0, absPrec);
    }
};

// ProjectionCalculator::calcNode (src/treebuilders/ProjectionCalculator.cpp:34-51)
// with MWNode::getPrimitiveChildPts/getExpandedChildPts (src/trees/MWNode.cpp:854-925)
// and MWNode::cvTransform (:448-490).
template <int D, typename F> struct ProjectionCalculator : Calculator<D> {
    const F *func;
    explicit ProjectionCalculator(const F &f)
            : func(&f) {}
    void calcNode(Node<D> &nd) override {
        const MRA<D> &m = *nd.tree->mra;
        int kp1 = m.kp1, kp1_d = m.kp1_d;
        constexpr int TD = Node<D>::TD;
        const Quadrature &q = quadrature(kp1);
        double sFac = std::pow(2.0, -(nd.n + 1));
        std::vector<double> prim[D][2];
        for (int d = 0; d < D; d++) {
            prim[d][0].resize(kp1);
            prim[d][1].resize(kp1);
            for (int i = 0; i < kp1; i++) {
                prim[d][0][i] = sFac * (q.roots[i] + 2.0 * static_cast<double>(nd.l[d]));
                prim[d][1][i] = sFac * (q.roots[i] + 2.0 * static_cast<double>(nd.l[d]) + 1.0);
            }
        }
        for (int t = 0; t < TD; t++) {
            for (int p = 0; p < kp1_d; p++) {
                std::array<double, D> r;
                int rem = p;
                for (int d = 0; d < D; d++) {
                    int id = rem % kp1;
                    rem /= kp1;
                    r[d] = m.sf[d] * prim[d][(t >> d) & 1][id];
                }
                nd.coefs[t * kp1_d + p] = func->evalf(r);
            }
        }
        // cvTransform(Backward)
        int kp1_dm1 = ipow(kp1, D - 1);
        std::vector<double> o_vec(TD * kp1_d);
        double *out_vec = o_vec.data();
        double *in_vec = nd.coefs.data();
        for (int i = 0; i < D; i++) {
            for (int t = 0; t < TD; t++) apply_filter(out_vec + t * kp1_d, in_vec + t * kp1_d, m.basis->vcMap.data(), kp1, kp1_dm1, false);
            std::swap(in_vec, out_vec);
        }
        double sf_prod = 1.0;
        for (int d = 0; d < D; d++) sf_prod *= m.sf[d];
        if (sf_prod <= MachineZero) sf_prod = 1.0;
        int np1 = nd.n + 1;
        double two_fac = std::pow(2.0, D * np1) / sf_prod;
        two_fac = std::sqrt(1.0 / two_fac);
        if (D % 2 == 1) {
            for (int i = 0; i < TD * kp1_d; i++) nd.coefs[i] = two_fac * in_vec[i];
        } else {
            for (int i = 0; i < TD * kp1_d; i++) nd.coefs[i] *= two_fac;
        }
        node_mw_transform(nd, Filter::Compression);
        nd.has_coefs = true;
        calc_norms(nd);
    }
};

// build_grid(FunctionTree, GaussExp) (src/treebuilders/grid.cpp:106-123)
template <int D> static void build_grid(Tree<D> &t, const GaussExp<D> &g) {
    for (auto &f : g.funcs) {
        DefaultCalculator<D> calc;
        AnalyticAdaptor<D> ad(f, t.mra->maxScale());
        build(t, calc, ad, -1);
    }
}

// project (src/treebuilders/project.cpp:85-102)
template <int D, typename F> static void project(double prec, Tree<D> &t, const F &func, int maxIter = -1, bool absPrec = false) {
    ProjectionCalculator<D, F> calc(func);
    WaveletAdaptor<D> ad(prec, t.mra->maxScale(), absPrec);
    build(t, calc, ad, maxIter);
    mw_transform_up(t);
    calc_square_norm(t);
}

// ----------------------------------------------------------------------------
// Kernels (src/operators/PoissonKernel.cpp:48-86, HelmholtzKernel.cpp:47-80)
// ----------------------------------------------------------------------------
struct Term {
    double alpha, coef;
};

static std::vector<Term> poisson_kernel(double epsilon, double r_min, double r_max) {
    std::vector<Term> out;
    double r0 = r_min / r_max;
    double r1 = r_max;
    double t1 = 1.0L;
    while ((2.0 * t1 * std::exp(-t1)) > epsilon) t1 *= 1.1L;
    double t2 = 1.0L;
    while ((std::sqrt(t2) * std::exp(-t2) / r0) > epsilon) t2 *= 1.1L;
    double s1 = -std::log(2.0 * t1);
    double s2 = std::log(t2 / (r0 * r0)) / 2.0;
    double h = 1.0 / (0.2L - 0.47L * std::log10(epsilon));
    int n_exp = static_cast<int>(std::ceil((s2 - s1) / h) + 1);
    if (n_exp > MaxSepRank) throw Error("Maximum separation rank exceeded.");
    for (int i = 0; i < n_exp; i++) {
        double arg = s1 + h * i;
        double sinharg = std::sinh(arg);
        double cosharg = std::cosh(arg);
        double onepexp = 1.0 + std::exp(-sinharg);
        double expo = 4.0L * (sinharg + std::log(onepexp)) * (sinharg + std::log(onepexp));
        double coef = h * (4.0L / root_pi) * cosharg / onepexp;
        expo *= 1.0 / (r1 * r1);
        coef *= 1.0 / r1;
        if (i == 0 or i == (n_exp - 1)) coef *= 1.0 / 2.0;
        out.push_back({expo, coef});
    }
    return out;
}

static std::vector<Term> helmholtz_kernel(double mu, double epsilon, double r_min, double r_max) {
    std::vector<Term> out;
    double r0 = r_min / r_max;
    double r1 = r_max;
    double mu_tilde = mu * r1;
    double t = std::max((-2.5L * std::log(epsilon)), 5.0L);
    double s1 = -std::log(4 * t / (mu_tilde * mu_tilde)) / 2;
    double s2 = std::log(t / (r0 * r0)) / 2;
    double h = 1.0 / (0.20L - 0.47L * std::log10(epsilon));
    int n_exp = static_cast<int>(std::ceil((s2 - s1) / h) + 1);
    if (n_exp > MaxSepRank) throw Error("Maximum separation rank exceeded.");
    for (int i = 0; i < n_exp; i++) {
        double arg = s1 + h * i;
        double temp = -arg * 2.0;
        double temp2 = -mu_tilde * mu_tilde * std::exp(temp) / 4.0 + arg;
        double beta = (h * (2.0 / root_pi) * std::exp(temp2));
        double temp3 = 2.0L * arg;
        double alpha = std::exp(temp3);
        alpha *= 1.0 / (r1 * r1);
        beta *= 1.0 / r1;
        if (i == 0 or i == (n_exp - 1)) beta *= 1.0 / 2.0;
        out.push_back({alpha, beta});
    }
    return out;
}

// ----------------------------------------------------------------------------
// Operator trees (src/trees/OperatorTree.cpp, src/treebuilders/CrossCorrelationCalculator.cpp)
// ----------------------------------------------------------------------------
struct OperTree {
    std::unique_ptr<MRA<2>> mra;
    std::unique_ptr<Tree<2>> tree;
    // per depth: translation cache for l in [-max_transl, max_transl]
    std::vector<int> max_transl;
    std::vector<std::vector<Node<2> *>> cache; // cache[n][l + max_transl[n]]
    // BandWidth (src/trees/BandWidth.h:37-61): rows 0..depth, columns T,C,B,A,max
    std::vector<std::array<int, 5>> bw;

    Node<2> &get(int depth, int l) { return *cache[depth][l + max_transl[depth]]; }
    int bwDepth() const { return (int)bw.size() - 1; }
    int width(int depth, int idx) const { return (depth > bwDepth()) ? -1 : bw[depth][idx]; }
    int maxWidth(int depth) const { return (depth > bwDepth()) ? -1 : bw[depth][4]; }
};

// CrossCorrelationCalculator::calcNode/applyCcc (CrossCorrelationCalculator.cpp:36-96)
struct CrossCorrelationCalculator : Calculator<2> {
    Tree<1> *kernel;
    explicit CrossCorrelationCalculator(Tree<1> &k)
            : kernel(&k) {}
    void calcNode(Node<2> &nd) override {
        nd.zeroCoefs();
        const MRA<2> &m = *nd.tree->mra;
        const CrossCorr &ccc = get_ccc(m.order);
        int K = m.kp1, kp1_d = m.kp1_d, K2 = 2 * K;
        int scale = nd.n + 1;
        std::vector<double> vec_o(4 * kp1_d, 0.0);
        for (int i = 0; i < 4; i++) {
            auto l = nd.childL(i);
            int l_a = l[1] - l[0] - 1;
            int l_b = l[1] - l[0];
            const Node<1> &na = get_node(*kernel, scale, {l_a});
            const Node<1> &nb = get_node(*kernel, scale, {l_b});
            for (int r = 0; r < kp1_d; r++) {
                double sa = 0.0, sb = 0.0;
                for (int p = 0; p < K2; p++) {
                    sa += ccc.L[r * K2 + p] * na.coefs[p];
                    sb += ccc.R[r * K2 + p] * nb.coefs[p];
                }
                vec_o[i * kp1_d + r] = sa + sb;
            }
        }
        double two_n = std::pow(2.0, -scale / 2.0);
        for (int i = 0; i < 4 * kp1_d; i++) nd.coefs[i] = std::sqrt(m.sf[0]) * two_n * vec_o[i];
        node_mw_transform(nd, Filter::Compression);
        nd.has_coefs = true;
        calc_norms(nd);
    }
};

// OperatorAdaptor (src/treebuilders/OperatorAdaptor.h:38-48)
struct OperatorAdaptor : Adaptor<2> {
    explicit OperatorAdaptor(int ms)
            : Adaptor<2>(ms) {}
    bool split(const Node<2> &nd) const override {
        bool chkTransl = (nd.l[0] == 0 or nd.l[1] == 0);
        bool chkCompNorm = false;
        for (int i = 1; i < 4; i++)
            if (nd.cn[i] > 0.0) chkCompNorm = true;
        return chkTransl and chkCompNorm;
    }
};

// OperatorTree::getMaxTranslations (src/trees/OperatorTree.cpp:181-192)
static std::vector<int> max_translations(Tree<2> &t) {
    std::vector<int> mt(t.getDepth(), 0);
    for (Node<2> *nd : node_table(t)) {
        int n = nd->depth();
        mt[n] = std::max(mt[n], std::abs(nd->l[0]));
        mt[n] = std::max(mt[n], std::abs(nd->l[1]));
    }
    return mt;
}

// OperatorTree::setupOperNodeCache (src/trees/OperatorTree.cpp:200-238):
// l > 0 -> node (l, 0), l <= 0 -> node (0, -l); missing nodes are generated.
static void setup_oper_cache(OperTree &ot) {
    Tree<2> &t = *ot.tree;
    int nScales = t.getDepth();
    ot.max_transl = max_translations(t);
    ot.cache.assign(nScales, {});
    for (int n = 0; n < nScales; n++) {
        int scale = t.rootScale() + n;
        int nt = ot.max_transl[n];
        auto &v = ot.cache[n];
        for (int i = nt; i >= 0; i--) v.push_back(&get_node(t, scale, {0, i}));
        for (int i = 1; i <= nt; i++) v.push_back(&get_node(t, scale, {i, 0}));
    }
}

// OperatorTree::calcBandWidth (src/trees/OperatorTree.cpp:109-134)
static void calc_band_width(OperTree &ot, double prec) {
    Tree<2> &t = *ot.tree;
    int depthN = t.getDepth();
    ot.bw.assign(depthN + 1, {-1, -1, -1, -1, -1});
    std::vector<int> mt = max_translations(t);
    if (prec < 0.0) prec = t.normPrec;
    for (int depth = 0; depth < depthN; depth++) {
        int l = 0;
        bool done = false;
        while (not done) {
            done = true;
            if (l > ot.max_transl[depth]) throw Error("band width beyond the operator node cache");
            Node<2> &nd = ot.get(depth, l);
            double thrs = std::max(MachinePrec, prec / (8.0 * (1 << depth)));
            for (int k = 0; k < 4; k++) {
                if (nd.cn[k] > thrs) {
                    ot.bw[depth][k] = l;
                    if (l > ot.bw[depth][4]) ot.bw[depth][4] = l;
                    done = false;
                }
            }
            if (++l > mt[depth]) break;
        }
    }
}

// ConvolutionOperator<3> (src/operators/ConvolutionOperator.cpp:48-142,
// MWOperator.cpp:35-129): one OperatorTree per Gaussian term, used in all
// three directions.
struct Operator {
    const MRA<3> *mra;
    std::vector<Term> kernel;
    std::vector<std::unique_ptr<OperTree>> terms;
    std::vector<int> band_max;
    double build_prec;
    double ms_build = 0.0;

    int maxBandWidth(int depth) const {
        if (depth < 0) return *std::max_element(band_max.begin(), band_max.end());
        if (depth < (int)band_max.size()) return band_max[depth];
        return -1;
    }
};

static void build_operator(Operator &op, const MRA<3> &mra, const std::vector<Term> &kernel, double prec) {
    auto t0 = std::chrono::steady_clock::now();
    op.mra = &mra;
    op.kernel = kernel;
    op.build_prec = prec;
    double o_prec = prec, k_prec = prec / 10.0;
    // getKernelMRA (:110-142): order 2k+1, box [-reach, reach) at the root scale
    int reach = -10 + 1;
    if (reach < 0)
        for (int d = 0; d < 3; d++)
            if (mra.nbox[d] > reach) reach = mra.nbox[d];
    auto k_mra = std::make_unique<MRA<1>>();
    k_mra->order = 2 * mra.order + 1;
    k_mra->n0 = mra.n0;
    k_mra->corner = {-reach};
    k_mra->nbox = {2 * reach};
    k_mra->sf = {mra.sf[0]};
    k_mra->finish();
    for (auto &term : kernel) {
        GaussFunc<1> kf;
        kf.alpha = {term.alpha};
        kf.coef = std::copysign(std::pow(std::abs(term.coef), 1.0 / 3), term.coef);
        kf.pos = {0.0};
        Tree<1> k_tree(*k_mra);
        GaussExp<1> ge;
        ge.funcs.push_back(kf);
        build_grid(k_tree, ge);
        project(k_prec, k_tree, kf);

        auto ot = std::make_unique<OperTree>();
        // getOperatorMRA (MWOperator.cpp:110-129)
        ot->mra = std::make_unique<MRA<2>>();
        ot->mra->order = mra.order;
        ot->mra->n0 = mra.n0;
        ot->mra->corner = {0, 0};
        ot->mra->nbox = {reach, reach};
        ot->mra->sf = {mra.sf[0], mra.sf[0]};
        ot->mra->finish();
        ot->tree = std::make_unique<Tree<2>>(*ot->mra, true, o_prec);
        CrossCorrelationCalculator calc(k_tree);
        OperatorAdaptor ad(ot->mra->maxScale());
        build(*ot->tree, calc, ad, -1);
        mw_transform_up(*ot->tree);
        calc_square_norm(*ot->tree);
        setup_oper_cache(*ot);
        op.terms.push_back(std::move(ot));
    }
    op.ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// MWOperator::calcBandWidths (src/operators/MWOperator.cpp:78-108)
static void calc_band_widths(Operator &op, double prec) {
    int maxDepth = 0;
    for (auto &t : op.terms) {
        calc_band_width(*t, prec);
        maxDepth = std::max(maxDepth, t->bwDepth());
    }
    op.band_max.assign(maxDepth + 1, -1);
    for (auto &t : op.terms)
        for (int n = 0; n <= t->bwDepth(); n++)
            for (int j = 0; j < 4; j++) op.band_max[n] = std::max(op.band_max[n], t->width(n, j));
}

// ----------------------------------------------------------------------------
// Convolution apply (src/treebuilders/ConvolutionCalculator.cpp)
// ----------------------------------------------------------------------------
struct ApplyStats {
    std::atomic<long long> fCount{0}, gCount{0}, genCount{0};
    // work-shape counters: band pairs, (g,f) pairs with work, (g,f,ft,term) groups,
    // distinct x-sweeps and y-sweeps when sweeps are shared across gt
    std::atomic<long long> nBand{0}, nPairs{0}, nGroups{0}, nXs{0}, nYs{0};
    std::vector<int> level_nodes;
    double seconds = 0.0;
};

struct ConvolutionCalculator : Calculator<3> {
    double prec;
    Operator *oper;
    Tree<3> *fTree;
    ApplyStats *stats;
    const PrecFunc<3> *precFunc = nullptr; // setPrecFunction (ConvolutionCalculator.h:43)
    // bandSizes per term: MaxDepth x 65 (:105-139)
    std::vector<std::vector<int>> bandSizes;
    int kp1, kp1_2, kp1_d, kp1_dm1;

    ConvolutionCalculator(double p, Operator &o, Tree<3> &f, ApplyStats *s)
            : prec(p)
            , oper(&o)
            , fTree(&f)
            , stats(s) {
        kp1 = f.mra->kp1;
        kp1_2 = kp1 * kp1;
        kp1_d = f.mra->kp1_d;
        kp1_dm1 = kp1_2;
        for (auto &t : oper->terms) {
            std::vector<int> bs(MaxDepth * 65, 0);
            for (int depth = 0; depth < MaxDepth; depth++) {
                int mx = 0;
                for (int gt = 0; gt < 8; gt++)
                    for (int ft = 0; ft < 8; ft++) {
                        int k = gt * 8 + ft;
                        int totNodes = 1;
                        for (int i = 0; i < 3; i++) {
                            int oIdx = 2 * ((gt >> i) & 1) + ((ft >> i) & 1);
                            int width = t->width(depth, oIdx);
                            if (width < 0) continue; // the quirk: only the factor is skipped
                            totNodes *= 2 * width + 1;
                        }
                        bs[depth * 65 + k] = totNodes * 64;
                        mx = std::max(mx, bs[depth * 65 + k]);
                    }
                bs[depth * 65 + 64] = mx;
            }
            bandSizes.push_back(std::move(bs));
        }
    }

    std::vector<Node<3> *> initial(Tree<3> &t) override { return node_table(t); }

    // tensorApplyOperComp (:333-382): three sweeps with cyclic index
    // rotation; accumulation into g happens only in the last sweep.
    void tensor_apply(const double *f, double *g, const double *const *o, double *s1, double *s2) const {
        apply_filter(s1, f, o[0], kp1, kp1_dm1, false);
        apply_filter(s2, s1, o[1], kp1, kp1_dm1, false);
        apply_filter(g, s2, o[2], kp1, kp1_dm1, true);
    }

    void calcNode(Node<3> &g) override {
        g.zeroCoefs();
        stats->gCount++;
        int o_depth = g.n - oper->mra->n0;
        // makeOperBand (:142-179): box g_l +- width clamped to the world
        std::vector<Node<3> *> band;
        std::vector<std::array<int, 3>> bidx;
        int width = oper->maxBandWidth(o_depth);
        if (width >= 0) {
            const MRA<3> &m = *fTree->mra;
            std::array<int, 3> s, e;
            for (int i = 0; i < 3; i++) {
                s[i] = g.l[i] - width;
                e[i] = g.l[i] + width;
                int nboxes = m.nbox[i] * (1 << o_depth);
                int c_i = m.corner[i] * (1 << o_depth);
                if (s[i] < c_i) s[i] = c_i;
                if (e[i] > c_i + nboxes - 1) e[i] = c_i + nboxes - 1;
            }
            // fillOperBand (:182-222): z outermost, x innermost
            for (int z = s[2]; z <= e[2]; z++)
                for (int y = s[1]; y <= e[1]; y++)
                    for (int x = s[0]; x <= e[0]; x++) {
                        std::array<int, 3> l{x, y, z};
                        band.push_back(&get_node(*fTree, g.n, l));
                        bidx.push_back(l);
                    }
        }
        double gThrs = g.tree->squareNorm;
        if (gThrs > 0.0) {
            double nTerms = (double)oper->terms.size();
            const double precFac = precFunc ? (*precFunc)(g.n, g.l) : 1.0;
            gThrs = prec * precFac * std::sqrt(gThrs / nTerms);
        }
        std::vector<double> scr(2 * kp1_d);
        long long applied = 0, genApplied = 0, nPairs = 0, nGroups = 0, nXs = 0, nYs = 0;
        const int M = (int)oper->terms.size();
        std::vector<unsigned char> gmask(8 * M);
        for (size_t b = 0; b < band.size(); b++) {
            std::fill(gmask.begin(), gmask.end(), 0);
            long long before = applied;
            Node<3> &f = *band[b];
            const auto &fl = bidx[b];
            int maxDeltaL = 0;
            for (int d = 0; d < 3; d++) maxDeltaL = std::max(maxDeltaL, std::abs(fl[d] - g.l[d]));
            for (int ft = 0; ft < 8; ft++) {
                double fNorm = f.cn[ft];
                if (fNorm < MachineZero) continue;
                for (int gt = 0; gt < 8; gt++) {
                    if (not(o_depth == 0 or gt != 0 or ft != 0)) continue;
                    // applyOperComp (:277-288) and applyOperator (:297-329)
                    for (int i = 0; i < M; i++) {
                        OperTree &ot = *oper->terms[i];
                        if (maxDeltaL > ot.maxWidth(o_depth)) continue;
                        double fThreshold = bandSizes[i][o_depth * 65 + gt * 8 + ft] * fNorm;
                        double oNorm = 1.0;
                        const double *od[3];
                        bool outside = false;
                        for (int d = 0; d < 3; d++) {
                            int oTransl = fl[d] - g.l[d];
                            int a = (gt >> d) & 1, bb = (ft >> d) & 1;
                            int idx = (a << 1) + bb;
                            if (std::abs(oTransl) > ot.width(o_depth, idx)) {
                                outside = true;
                                break;
                            }
                            Node<2> &on = ot.get(o_depth, oTransl);
                            oNorm *= on.cn[idx];
                            od[d] = on.coefs.data() + idx * kp1_2;
                        }
                        if (outside) continue;
                        double upperBound = oNorm * fThreshold;
                        if (upperBound > gThrs) {
                            gmask[ft * M + i] |= (unsigned char)(1 << gt);
                            applied++;
                            if (f.gen) genApplied++;
                            tensor_apply(f.coefs.data() + ft * kp1_d, g.coefs.data() + gt * kp1_d, od, scr.data(), scr.data() + kp1_d);
                        }
                    }
                }
            }
            if (applied > before) nPairs++;
            for (int q = 0; q < 8 * M; q++) {
                unsigned m = gmask[q];
                if (!m) continue;
                nGroups++;
                unsigned xs = 0, ys = 0;
                for (int gt = 0; gt < 8; gt++)
                    if (m >> gt & 1) {
                        xs |= 1u << (gt & 1);
                        ys |= 1u << (gt & 3);
                    }
                nXs += __builtin_popcount(xs);
                nYs += __builtin_popcount(ys);
            }
        }
        stats->nBand += (long long)band.size();
        stats->nPairs += nPairs;
        stats->nGroups += nGroups;
        stats->nXs += nXs;
        stats->nYs += nYs;
        stats->fCount += applied;
        stats->genCount += genApplied;
        calc_norms(g);
    }
};

// apply (src/treebuilders/apply.cpp:68-93)
static void apply(double prec, Tree<3> &out, Operator &oper, Tree<3> &inp, int maxIter, bool absPrec, ApplyStats &st) {
    if (out.mra != inp.mra) throw Error("Incompatible MRA");
    auto t0 = std::chrono::steady_clock::now();
    calc_band_widths(oper, prec);
    WaveletAdaptor<3> adaptor(prec, out.mra->maxScale(), absPrec);
    ConvolutionCalculator calc(prec, oper, inp, &st);
    BuildStats bst;
    build(out, calc, adaptor, maxIter, &bst);
    st.level_nodes = bst.level_nodes;
    mw_transform_down(out, false);
    mw_transform_up(out);
    calc_square_norm(out);
    delete_generated(inp);
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// apply with precision trees (src/treebuilders/apply.cpp:214-251): the
// screening threshold and the split test of node idx use prec * precFunc(idx).
static void apply_prec(double prec, Tree<3> &out, Operator &oper, Tree<3> &inp, std::vector<Tree<3> *> &precTrees, int maxIter, bool absPrec,
                       ApplyStats &st) {
    auto t0 = std::chrono::steady_clock::now();
    calc_band_widths(oper, prec);
    for (Tree<3> *t : precTrees) make_max_square_norms(*t);
    PrecFunc<3> pf{precTrees};
    WaveletAdaptor<3> adaptor(prec, out.mra->maxScale(), absPrec);
    adaptor.precFunc = &pf;
    ConvolutionCalculator calc(prec, oper, inp, &st);
    calc.precFunc = &pf;
    BuildStats bst;
    build(out, calc, adaptor, maxIter, &bst);
    st.level_nodes = bst.level_nodes;
    mw_transform_down(out, false);
    mw_transform_up(out);
    calc_square_norm(out);
    delete_generated(inp);
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// dot (src/treebuilders/multiply.cpp:286-318), serial
static double dot(Tree<3> &bra, Tree<3> &ket) {
    double result = 0.0;
    int kp1_d = bra.mra->kp1_d;
    for (Node<3> *b : node_table(bra)) {
        Node<3> *k = find_node(ket, b->n, b->l);
        if (k == nullptr) continue;
        if (b->depth() == 0)
            for (int i = 0; i < kp1_d; i++) result += b->coefs[i] * k->coefs[i];
        double w = 0.0;
        for (int i = kp1_d; i < 8 * kp1_d; i++) w += b->coefs[i] * k->coefs[i];
        result += w;
    }
    return result;
}

// FunctionTree::evalf (src/trees/FunctionTree.cpp:524-546) with
// FunctionNode::evalScaling (src/trees/FunctionNode.cpp:60-88)
static double evalf(Tree<3> &t, const std::array<double, 3> &r) {
    const MRA<3> &m = *t.mra;
    std::array<double, 3> arg;
    for (int d = 0; d < 3; d++) arg[d] = r[d] / m.sf[d];
    for (int d = 0; d < 3; d++) {
        double lo = m.corner[d] * std::pow(2.0, -m.n0), hi = lo + m.nbox[d] * std::pow(2.0, -m.n0);
        if (arg[d] < lo or arg[d] > hi) return 0.0;
    }
    // root containing the point, then descend to the end node
    std::array<int, 3> l;
    for (int d = 0; d < 3; d++) {
        double x = arg[d] * std::pow(2.0, m.n0);
        int li = (int)std::floor(x);
        if (li >= m.corner[d] + m.nbox[d]) li = m.corner[d] + m.nbox[d] - 1;
        l[d] = li;
    }
    Node<3> *nd = t.roots[m.boxIndex(m.n0, l)].get();
    while (nd->hasRealKids()) {
        int c = 0;
        for (int d = 0; d < 3; d++) {
            double x = arg[d] * std::pow(2.0, nd->n + 1);
            int li = (int)std::floor(x);
            li = std::min(std::max(li, 2 * nd->l[d]), 2 * nd->l[d] + 1);
            c |= (li - 2 * nd->l[d]) << d;
        }
        nd = &nd->child(c);
    }
    double nf = std::pow(2.0, nd->n);
    double val[3][64];
    for (int d = 0; d < 3; d++) {
        double a = arg[d] * nf - (double)nd->l[d];
        for (int k = 0; k < m.kp1; k++) val[d][k] = m.basis->evalf(k, a);
    }
    double res = 0.0;
    for (int i = 0; i < m.kp1_d; i++) {
        int ix = i % m.kp1, iy = (i / m.kp1) % m.kp1, iz = i / (m.kp1 * m.kp1);
        res += nd->coefs[i] * val[0][ix] * val[1][iy] * val[2][iz];
    }
    double two_n = std::pow(2.0, (3 * nd->n) / 2.0);
    double coef = 1.0;
    for (int d = 0; d < 3; d++) coef /= std::sqrt(m.sf[d]);
    return coef * two_n * res;
}

// ----------------------------------------------------------------------------
// Tree arithmetic: add and multiply (src/treebuilders/add.cpp:43-80,
// src/treebuilders/multiply.cpp:52-100) with WaveletAdaptor refinement
// ----------------------------------------------------------------------------
using TreeVec = std::vector<std::pair<double, Tree<3> *>>;

// AdditionCalculator::calcNode (src/treebuilders/AdditionCalculator.h:45-70):
// the linear combination of the inputs' nodes, generated from their ancestors
// when missing (a generated node holds its scaling block only).
struct AdditionCalculator : Calculator<3> {
    const TreeVec &inp;
    explicit AdditionCalculator(const TreeVec &v)
            : inp(v) {}
    void calcNode(Node<3> &out) override {
        out.zeroCoefs();
        for (const auto &ci : inp) {
            Node<3> &nd = get_node(*ci.second, out.n, out.l);
            for (size_t j = 0; j < nd.coefs.size(); j++) out.coefs[j] += ci.first * nd.coefs[j];
        }
        out.has_coefs = true;
        calc_norms(out);
    }
};

// MultiplicationCalculator::calcNode (src/treebuilders/MultiplicationCalculator.h:45-83):
// each input node (a copy) reconstructed to its children and mapped to values
// at their quadrature points, the values multiplied (times each input's
// coefficient), mapped back and compressed.
struct MultiplicationCalculator : Calculator<3> {
    const TreeVec &inp;
    std::vector<double> cv;
    explicit MultiplicationCalculator(const TreeVec &v, const Basis &b)
            : inp(v)
            , cv(cv_map_of(b)) {}
    static std::vector<double> cv_map_of(const Basis &b);
    void calcNode(Node<3> &out) override {
        const MRA<3> &m = *out.tree->mra;
        const int nc = 8 * m.kp1_d;
        for (int j = 0; j < nc; j++) out.coefs[j] = 1.0;
        Node<3> tmp;
        tmp.tree = out.tree;
        tmp.n = out.n;
        for (const auto &ci : inp) {
            Node<3> &nd = get_node(*ci.second, out.n, out.l);
            tmp.coefs.assign(nc, 0.0);
            std::copy(nd.coefs.begin(), nd.coefs.end(), tmp.coefs.begin());
            node_mw_transform(tmp, Filter::Reconstruction);
            cv_transform_node(m, out.n, cv, true, tmp.coefs.data());
            for (int j = 0; j < nc; j++) out.coefs[j] *= ci.first * tmp.coefs[j];
        }
        cv_transform_node(m, out.n, m.basis->vcMap, false, out.coefs.data());
        node_mw_transform(out, Filter::Compression);
        out.has_coefs = true;
        calc_norms(out);
    }
    static void cv_transform_node(const MRA<3> &m, int n, const std::vector<double> &S, bool forward, double *coefs);
};

template <int D> struct WaveletSplit : Adaptor<D> {
    double prec;
    bool absPrec;
    WaveletSplit(double p, int ms, bool ap)
            : Adaptor<D>(ms)
            , prec(p)
            , absPrec(ap) {}
    bool split(const Node<D> &nd) const override { return split_check(nd, prec, 1.0, absPrec); }
};

static void add(double prec, Tree<3> &out, const TreeVec &inp, int maxIter, bool absPrec) {
    for (const auto &ci : inp)
        if (ci.second->mra != out.mra) throw Error("Incompatible MRA");
    AdditionCalculator calc(inp);
    WaveletSplit<3> ad(prec, out.mra->maxScale(), absPrec);
    build(out, calc, ad, maxIter);
    mw_transform_up(out);
    calc_square_norm(out);
    for (const auto &ci : inp) delete_generated(*ci.second);
}

// MultiplicationAdaptor::splitNode (src/treebuilders/MultiplicationAdaptor.h:46-66)
struct MultiplicationAdaptor : Adaptor<3> {
    double prec;
    Tree<3> *t0, *t1;
    MultiplicationAdaptor(double p, int ms, Tree<3> *a, Tree<3> *b)
            : Adaptor<3>(ms)
            , prec(p)
            , t0(a)
            , t1(b) {}
    bool split(const Node<3> &nd) const override {
        Node<3> &p0 = get_node(*t0, nd.n, nd.l);
        Node<3> &p1 = get_node(*t1, nd.n, nd.l);
        const double maxW0 = std::sqrt(max_wsquare_norm(p0)), maxW1 = std::sqrt(max_wsquare_norm(p1));
        const double maxS0 = std::sqrt(max_square_norm(p0)), maxS1 = std::sqrt(max_square_norm(p1));
        const double multNorm = maxW0 * maxS1 + maxW1 * maxS0 + maxW0 * maxW1;
        return multNorm > prec && !(!p0.isBranch() && !p1.isBranch());
    }
};

static void multiply(double prec, Tree<3> &out, const TreeVec &inp, int maxIter, bool absPrec, bool useMaxNorms = false) {
    for (const auto &ci : inp)
        if (ci.second->mra != out.mra) throw Error("Incompatible MRA");
    MultiplicationCalculator calc(inp, *out.mra->basis);
    if (useMaxNorms) {
        // multiply.cpp:110-114
        if (inp.size() != 2) throw Error("Invalid tree vec size");
        for (const auto &ci : inp) make_max_square_norms(*ci.second);
        MultiplicationAdaptor ad(prec, out.mra->maxScale(), inp[0].second, inp[1].second);
        build(out, calc, ad, maxIter);
    } else {
        WaveletSplit<3> ad(prec, out.mra->maxScale(), absPrec);
        build(out, calc, ad, maxIter);
    }
    mw_transform_up(out);
    calc_square_norm(out);
    for (const auto &ci : inp) delete_generated(*ci.second);
}

// ----------------------------------------------------------------------------
// Text files in MADNESS conventions (FunctionTree::saveTreeTXT / loadTreeTXT,
// src/trees/FunctionTree.cpp:127-373)
// ----------------------------------------------------------------------------

// LegendreBasis::calcCVMaps (src/core/LegendreBasis.cpp:98-109): cvMap is the
// inverse of vcMap, here by Gauss-Jordan elimination with partial pivoting.
static std::vector<double> cv_map(const Basis &b) {
    const int n = b.order + 1;
    std::vector<double> a(b.vcMap), inv(n * n, 0.0);
    for (int i = 0; i < n; i++) inv[i + n * i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::abs(a[r + n * c]) > std::abs(a[piv + n * c])) piv = r;
        if (piv != c)
            for (int j = 0; j < n; j++) {
                std::swap(a[c + n * j], a[piv + n * j]);
                std::swap(inv[c + n * j], inv[piv + n * j]);
            }
        const double d = a[c + n * c];
        for (int j = 0; j < n; j++) {
            a[c + n * j] /= d;
            inv[c + n * j] /= d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a[r + n * c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) {
                a[r + n * j] -= f * a[c + n * j];
                inv[r + n * j] -= f * inv[c + n * j];
            }
        }
    }
    return inv;
}

// MWNode::cvTransform (src/trees/MWNode.cpp:448-490) on a node's 8 blocks
static void cv_transform(const MRA<3> &m, int n, const std::vector<double> &S, bool forward, double *coefs) {
    const int kp1 = m.kp1, kp1_d = m.kp1_d, kp1_dm1 = kp1 * kp1;
    std::vector<double> o_vec(8 * kp1_d);
    double *out_vec = o_vec.data(), *in_vec = coefs;
    for (int i = 0; i < 3; i++) {
        for (int t = 0; t < 8; t++) apply_filter(out_vec + t * kp1_d, in_vec + t * kp1_d, S.data(), kp1, kp1_dm1, false);
        std::swap(in_vec, out_vec);
    }
    double sf_prod = m.sf[0] * m.sf[1] * m.sf[2];
    if (sf_prod <= MachineZero) sf_prod = 1.0;
    double two_fac = std::pow(2.0, 3 * (n + 1)) / sf_prod;
    two_fac = forward ? std::sqrt(two_fac) : std::sqrt(1.0 / two_fac);
    for (int i = 0; i < 8 * kp1_d; i++) coefs[i] = two_fac * in_vec[i];
}

// MADNESS index order (FunctionTree.cpp:168-176)
static std::vector<int> madness_map(int kp1) {
    std::vector<int> map;
    for (int x = kp1 - 1; x >= 0; x--)
        for (int y = kp1 - 1; y >= 0; y--)
            for (int z = kp1 - 1; z >= 0; z--) map.push_back(z * kp1 * kp1 + y * kp1 + x);
    return map;
}

// FunctionTree::saveTreeTXT (src/trees/FunctionTree.cpp:310-373)
static void save_tree_txt(Tree<3> &t, const std::string &fname) {
    const MRA<3> &m = *t.mra;
    std::ofstream out(fname);
    if (!out) throw Error("Unable to open file");
    out << std::setprecision(14);
    out << 3 << std::endl;
    const int rscale = m.n0;
    double LMRChem = 1.0;
    for (int i = 0; i > rscale; i--) LMRChem *= 2;
    for (int d = 0; d < 3; d++) out << -m.sf[d] * LMRChem << " " << m.sf[d] * LMRChem << std::endl;
    const int kp1 = m.kp1, ncoefs = m.kp1_d;
    out << kp1 << std::endl;
    auto ends = end_nodes(t);
    out << 8 * ends.size() << std::endl;
    const std::vector<int> mapMRC = madness_map(kp1);
    const std::vector<double> cv = cv_map(*m.basis);
    const int L = (int)std::pow(2, -rscale);
    std::vector<double> values(8 * ncoefs);
    Node<3> tmp;
    tmp.tree = &t;
    for (Node<3> *nd : ends) {
        tmp.n = nd->n;
        tmp.coefs = nd->coefs;
        node_mw_transform(tmp, Filter::Reconstruction);
        cv_transform(m, nd->n, cv, true, tmp.coefs.data());
        std::array<int, 3> l;
        for (int i = 0; i < 3; i++) l[i] = 2 * (nd->l[i] + std::pow(2, nd->n) * L);
        for (int cix = 0; cix < 8; cix++) {
            out << nd->n - rscale + 2 << " ";
            for (int i = 0; i < 3; i++) out << l[i] + ((cix >> i) & 1) << " ";
            out << std::endl;
            for (int i = 0; i < ncoefs; i++) out << tmp.coefs[cix * ncoefs + mapMRC[i]] << " ";
            out << std::endl;
        }
    }
}

// MWTree::getNode(idx, true) (src/trees/MWTree.cpp:340-352): real children created on the way
static Node<3> &create_node(Tree<3> &t, int n, const std::array<int, 3> &l) {
    int r = t.mra->boxIndex(n, l);
    if (r < 0) throw Error("node outside world box");
    Node<3> *nd = t.roots[r].get();
    while (nd->n < n) {
        if (not nd->isBranch()) create_children(*nd);
        nd = &nd->child(child_index(*nd, n, l));
    }
    return *nd;
}

// FunctionTree::loadTreeTXT (src/trees/FunctionTree.cpp:127-308)
static void load_tree_txt(Tree<3> &t, const std::string &fname) {
    const MRA<3> &m = *t.mra;
    std::ifstream in(fname);
    if (!in) throw Error("Unable to open file");
    int NDIM, k;
    in >> NDIM;
    if (NDIM != 3) throw Error("dimension");
    double coord[3][2];
    for (int d = 0; d < 3; d++) in >> coord[d][0] >> coord[d][1];
    int p = 1;
    const int rscale = m.n0;
    for (int i = rscale; i < 0; i++) p *= 2;
    const int L = p;
    for (int d = 0; d < 3; d++)
        if (std::abs(coord[d][0] + L) > 1.0e-14 or std::abs(coord[d][1] - L) > 1.0e-14) throw Error("world box");
    in >> k;
    if (k != m.kp1) throw Error("order");
    const int kp1 = k, ncoefs = m.kp1_d;
    const std::vector<int> mapMRC = madness_map(kp1);
    // fresh roots
    for (auto &r : t.roots) {
        Node<3> *kids = r->kids.load();
        r->kids.store(nullptr);
        delete[] kids;
        r->zeroCoefs();
    }
    t.nodesAtDepth.assign(1, (int)t.roots.size());
    std::vector<std::vector<Node<3> *>> NodeTable(50);
    std::map<Node<3> *, int> mp;
    int nmax = 0;
    int nread;
    in >> nread;
    while (nread-- > 0) {
        int n_in;
        in >> n_in;
        const int n = n_in + rscale - 1;
        std::array<int, 3> l_in, l, lp;
        for (int i = 0; i < 3; i++) in >> l_in[i];
        for (int i = 0; i < 3; i++) {
            l[i] = l_in[i] - std::pow(2, n) * L;
            lp[i] = l_in[i] / 2 - std::pow(2, n - 1) * L;
        }
        Node<3> *node = &create_node(t, n - 1, lp);
        int c_ix = 0, q = 1;
        for (int i = 0; i < 3; i++) {
            if (std::abs(l[i]) % 2 == 1) c_ix += q;
            q *= 2;
        }
        if (mp[node] == 0) {
            node->zeroCoefs();
            if (node->parent != nullptr)
                for (int c = 0; c < 8; c++)
                    if (mp[&node->parent->child(c)] == 0) node->parent->child(c).zeroCoefs();
        }
        double *values = node->coefs.data() + c_ix * ncoefs;
        for (int i = 0; i < ncoefs; i++) in >> values[mapMRC[i]];
        mp[node]++;
        nmax = std::max(nmax, n_in);
        if (mp[node] == 1) NodeTable[n_in].push_back(node);
    }
    for (int n = nmax; n > -1; n--)
        for (Node<3> *node : NodeTable[n]) {
            cv_transform(m, node->n, m.basis->vcMap, false, node->coefs.data());
            calc_norms(*node);
        }
    for (int n = nmax; n > -1; n--) {
        for (size_t i = 0; i < NodeTable[n].size(); i++) {
            Node<3> *node = NodeTable[n][i];
            if (mp[node] != 8) continue;
            if (not node->isBranch()) {
                node_mw_transform(*node, Filter::Compression);
                node->has_coefs = true;
                calc_norms(*node);
            } else {
                for (int c = 0; c < 8; c++) {
                    Node<3> &cn = node->child(c);
                    if (mp[&cn] == 8) continue;
                    for (int j = 0; j < ncoefs; j++) cn.coefs[j] = node->coefs[j + c * ncoefs];
                    for (int j = ncoefs; j < 8 * ncoefs; j++) cn.coefs[j] = 0.0;
                    cn.has_coefs = true;
                    calc_norms(cn);
                }
                node_mw_transform(*node, Filter::Compression);
                node->has_coefs = true;
                calc_norms(*node);
            }
            if (node->parent != nullptr) {
                Node<3> *parent = node->parent;
                if (mp[parent] == 0) NodeTable[n - 1].push_back(parent);
                int my_ix = -1;
                for (int c = 0; c < 8; c++)
                    if (&parent->child(c) == node) my_ix = c;
                for (int j = 0; j < ncoefs; j++) parent->coefs[j + my_ix * ncoefs] = node->coefs[j];
                mp[parent]++;
            }
        }
    }
    calc_square_norm(t);
}

std::vector<double> MultiplicationCalculator::cv_map_of(const Basis &b) { return cv_map(b); }
void MultiplicationCalculator::cv_transform_node(const MRA<3> &m, int n, const std::vector<double> &S, bool forward, double *coefs) {
    cv_transform(m, n, S, forward, coefs);
}

} // namespace orc

// ============================================================================
// C API (ctypes) -- for tests/, the smoke check and the CPU-baseline timing
// ============================================================================
using namespace orc;

struct OrcTree {
    std::unique_ptr<Tree<3>> t;
};

extern "C" {

const char *orc_last_error() { return g_last_error.c_str(); }

void orc_set_filter_dir(const char *dir) { g_filter_dir = dir; }

void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

int orc_max_threads() { return omp_get_max_threads(); }

void *orc_mra_create(int box_lo, int box_hi, int order, int depth) {
    try {
        auto *m = new MRA<3>(MRA<3>::from_box(box_lo, box_hi, order, depth));
        return m;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void orc_mra_destroy(void *m) { delete static_cast<MRA<3> *>(m); }

void orc_mra_info(void *mp, int *n0, double *sf, int *max_scale) {
    auto *m = static_cast<MRA<3> *>(mp);
    *n0 = m->n0;
    *sf = m->sf[0];
    *max_scale = m->maxScale();
}

void *orc_tree_create(void *mp) { return new OrcTree{std::make_unique<Tree<3>>(*static_cast<MRA<3> *>(mp))}; }

void orc_tree_destroy(void *t) { delete static_cast<OrcTree *>(t); }

// f = sum_i c_i exp(-a_i |r - R_i|^2): build_grid + project
int orc_tree_project_gaussians(void *tp, double prec, int n, const double *coefs, const double *alphas, const double *centers) {
    try {
        Tree<3> &t = *static_cast<OrcTree *>(tp)->t;
        GaussExp<3> ge;
        for (int i = 0; i < n; i++) {
            GaussFunc<3> g;
            g.alpha = {alphas[i], alphas[i], alphas[i]};
            g.coef = coefs[i];
            g.pos = {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]};
            ge.funcs.push_back(g);
        }
        build_grid(t, ge);
        project(prec, t, ge);
        return 0;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

// build_grid only (an empty grid refined around each Gaussian)
int orc_tree_build_grid_gaussians(void *tp, int n, const double *coefs, const double *alphas, const double *centers) {
    try {
        Tree<3> &t = *static_cast<OrcTree *>(tp)->t;
        GaussExp<3> ge;
        for (int i = 0; i < n; i++) {
            GaussFunc<3> g;
            g.alpha = {alphas[i], alphas[i], alphas[i]};
            g.coef = coefs[i];
            g.pos = {centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]};
            ge.funcs.push_back(g);
        }
        build_grid(t, ge);
        return 0;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

long long orc_tree_num_nodes(void *tp) { return (long long)node_table(*static_cast<OrcTree *>(tp)->t).size(); }

double orc_tree_square_norm(void *tp) { return static_cast<OrcTree *>(tp)->t->squareNorm; }

// nodes in depth-first order: scale, lx, ly, lz, is_end; coefficients 8*(k+1)^3 each
int orc_tree_export(void *tp, int *idx, double *coefs, double *norms) {
    Tree<3> &t = *static_cast<OrcTree *>(tp)->t;
    auto tab = node_table(t);
    int kp1_d = t.mra->kp1_d;
    for (size_t i = 0; i < tab.size(); i++) {
        Node<3> *nd = tab[i];
        idx[5 * i + 0] = nd->n;
        idx[5 * i + 1] = nd->l[0];
        idx[5 * i + 2] = nd->l[1];
        idx[5 * i + 3] = nd->l[2];
        idx[5 * i + 4] = nd->isEnd() ? 1 : 0;
        if (coefs) std::memcpy(coefs + i * 8 * (size_t)kp1_d, nd->coefs.data(), sizeof(double) * 8 * kp1_d);
        if (norms)
            for (int c = 0; c < 8; c++) norms[9 * i + c] = nd->cn[c];
        if (norms) norms[9 * i + 8] = nd->sq;
    }
    return (int)tab.size();
}

// out = sum_i c_i trees_i (mode 0) or prod_i c_i trees_i (mode 1)
// (mode 2: the product with useMaxNorms)
int orc_arith(int mode, double prec, void *outp, int n, const double *c, void **trees, int maxIter, int absPrec) {
    try {
        TreeVec v;
        for (int i = 0; i < n; i++) v.push_back({c[i], static_cast<OrcTree *>(trees[i])->t.get()});
        Tree<3> &out = *static_cast<OrcTree *>(outp)->t;
        if (mode == 0)
            add(prec, out, v, maxIter, absPrec != 0);
        else
            multiply(prec, out, v, maxIter, absPrec != 0, mode == 2);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

int orc_tree_save_txt(void *tp, const char *path) {
    try {
        save_tree_txt(*static_cast<OrcTree *>(tp)->t, path);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

int orc_tree_load_txt(void *tp, const char *path) {
    try {
        load_tree_txt(*static_cast<OrcTree *>(tp)->t, path);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

double orc_tree_evalf(void *tp, double x, double y, double z) { return evalf(*static_cast<OrcTree *>(tp)->t, {x, y, z}); }

double orc_tree_dot(void *a, void *b) { return dot(*static_cast<OrcTree *>(a)->t, *static_cast<OrcTree *>(b)->t); }

double orc_tree_integrate(void *tp) {
    Tree<3> &t = *static_cast<OrcTree *>(tp)->t;
    double r = 0.0;
    for (auto &nd : t.roots) r += std::pow(2.0, -(3 * nd->n) / 2.0) * nd->coefs[0];
    double sf = 1.0;
    for (int d = 0; d < 3; d++) sf *= std::sqrt(t.mra->sf[d]);
    return sf * r;
}

// kind 0: Poisson, 1: Helmholtz.  Builds the kernel and the operator trees.
void *orc_operator_create(void *mp, int kind, double mu, double prec) {
    try {
        auto *m = static_cast<MRA<3> *>(mp);
        double k_prec = prec / 10.0;
        double r_min = std::sqrt(k_prec * std::pow(2.0, -m->maxScale()));
        double r_max = 0.0;
        for (int d = 0; d < 3; d++) {
            double dd = m->upperBound(d) - m->lowerBound(d);
            r_max += dd * dd;
        }
        r_max = std::sqrt(r_max);
        std::vector<Term> kern = (kind == 0) ? poisson_kernel(k_prec, r_min, r_max) : helmholtz_kernel(mu, k_prec, r_min, r_max);
        auto *op = new Operator();
        build_operator(*op, *m, kern, prec);
        return op;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void orc_operator_destroy(void *op) { delete static_cast<Operator *>(op); }

int orc_operator_size(void *op) { return (int)static_cast<Operator *>(op)->terms.size(); }

double orc_operator_build_ms(void *op) { return static_cast<Operator *>(op)->ms_build; }

void orc_operator_kernel(void *op, double *alpha, double *coef) {
    auto *o = static_cast<Operator *>(op);
    for (size_t i = 0; i < o->kernel.size(); i++) {
        alpha[i] = o->kernel[i].alpha;
        coef[i] = o->kernel[i].coef;
    }
}

// Kernel expansion alone (for the term-count and accuracy checks)
int orc_kernel_terms(int kind, double mu, double eps, double r_min, double r_max, double *alpha, double *coef, int cap) {
    try {
        auto k = (kind == 0) ? poisson_kernel(eps, r_min, r_max) : helmholtz_kernel(mu, eps, r_min, r_max);
        for (size_t i = 0; i < k.size() and (int)i < cap; i++) {
            alpha[i] = k[i].alpha;
            coef[i] = k[i].coef;
        }
        return (int)k.size();
    } catch (std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

// band widths after calc_band_widths(prec): out[term][depth][5]
int orc_operator_bandwidths(void *op, double prec, int *out, int max_depth) {
    auto *o = static_cast<Operator *>(op);
    calc_band_widths(*o, prec);
    for (size_t i = 0; i < o->terms.size(); i++)
        for (int d = 0; d < max_depth; d++)
            for (int c = 0; c < 5; c++) {
                int w = (c < 4) ? o->terms[i]->width(d, c) : o->terms[i]->maxWidth(d);
                out[(i * max_depth + d) * 5 + c] = w;
            }
    return (int)o->band_max.size();
}

// One operator node component: term i, depth, translation l, component c
// (0=T,1=C,2=B,3=A), column-major (k+1)^2 matrix and its thresholded norm.
int orc_operator_node(void *op, int term, int depth, int l, int comp, double *mat, double *norm) {
    auto *o = static_cast<Operator *>(op);
    OperTree &ot = *o->terms[term];
    if (depth >= (int)ot.cache.size() or std::abs(l) > ot.max_transl[depth]) return -1;
    Node<2> &nd = ot.get(depth, l);
    int kp1 = ot.mra->kp1;
    std::memcpy(mat, nd.coefs.data() + comp * kp1 * kp1, sizeof(double) * kp1 * kp1);
    *norm = nd.cn[comp];
    return 0;
}

// apply; stats = [N_applied, N_gen_applied, N_g, seconds, n_levels, band pairs,
// pairs with work, (g,f,ft,term) groups, shared x-sweeps, shared y-sweeps]
// precTrees: n_prec tree handles, or n_prec < 0 for the plain apply
int orc_apply(double prec, void *outp, void *op, void *inp, int maxIter, int absPrec, double *stats, int *level_nodes, int max_levels,
              void **prec_trees, int n_prec) {
    try {
        ApplyStats st;
        if (n_prec < 0) {
            apply(prec, *static_cast<OrcTree *>(outp)->t, *static_cast<Operator *>(op), *static_cast<OrcTree *>(inp)->t, maxIter, absPrec != 0, st);
        } else {
            std::vector<Tree<3> *> pt;
            for (int i = 0; i < n_prec; i++) pt.push_back(static_cast<OrcTree *>(prec_trees[i])->t.get());
            apply_prec(prec, *static_cast<OrcTree *>(outp)->t, *static_cast<Operator *>(op), *static_cast<OrcTree *>(inp)->t, pt, maxIter,
                       absPrec != 0, st);
        }
        stats[0] = (double)st.fCount.load();
        stats[1] = (double)st.genCount.load();
        stats[2] = (double)st.gCount.load();
        stats[3] = st.seconds;
        stats[4] = (double)st.level_nodes.size();
        stats[5] = (double)st.nBand.load();
        stats[6] = (double)st.nPairs.load();
        stats[7] = (double)st.nGroups.load();
        stats[8] = (double)st.nXs.load();
        stats[9] = (double)st.nYs.load();
        for (size_t i = 0; i < st.level_nodes.size() and (int)i < max_levels; i++) level_nodes[i] = st.level_nodes[i];
        return 0;
    } catch (std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

// precFunc(n, l) over the given trees after makeMaxSquareNorms (the local
// precision factor of apply with precTrees); generates nodes in the trees
double orc_prec_factor(void **trees, int n_trees, int n, const int *l) {
    try {
        PrecFunc<3> pf;
        for (int i = 0; i < n_trees; i++) {
            pf.trees.push_back(static_cast<OrcTree *>(trees[i])->t.get());
            make_max_square_norms(*pf.trees.back());
        }
        return pf(n, {l[0], l[1], l[2]});
    } catch (std::exception &e) {
        g_last_error = e.what();
        return -1.0;
    }
}

} // extern "C"
