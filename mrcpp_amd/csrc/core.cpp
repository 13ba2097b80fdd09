// Numerical core: basis, quadrature, filters, cross-correlation tables.
#include "core.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>

#include <dlfcn.h>
#include <sys/stat.h>

namespace mrcpp {

namespace {
std::string &filter_dir_storage() {
    static std::string d;
    return d;
}
std::mutex &core_mutex() {
    static std::mutex m;
    return m;
}

std::vector<double> load_table(const std::string &name, size_t count) {
    std::string path = filter_dir() + "/" + name;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error("cannot open filter table " + path);
    std::vector<double> v(count);
    in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(count * sizeof(double)));
    if ((size_t)in.gcount() != count * sizeof(double)) throw Error("truncated filter table " + path);
    return v;
}
} // namespace

void set_filter_dir(const std::string &dir) {
    std::lock_guard<std::mutex> lk(core_mutex());
    filter_dir_storage() = dir;
}

// Filter folder when none was set: $MWFILTERS_DIR if it names a folder (as
// details::find_filters, src/utils/details.cpp:53-68, looks first), else the
// share/mwfilters folder beside this library's lib/ folder.
const std::string &filter_dir() {
    std::lock_guard<std::mutex> lk(core_mutex());
    std::string &d = filter_dir_storage();
    if (d.empty()) {
        struct stat st {};
        const char *env = std::getenv("MWFILTERS_DIR");
        if (env && stat(env, &st) == 0 && S_ISDIR(st.st_mode)) {
            d = env;
        } else {
            Dl_info info{};
            if (dladdr(reinterpret_cast<const void *>(&set_filter_dir), &info) && info.dli_fname) {
                std::string so = info.dli_fname;
                const size_t cut = so.find_last_of('/');
                d = (cut == std::string::npos ? std::string(".") : so.substr(0, cut)) + "/../share/mwfilters";
            }
        }
    }
    return d;
}

namespace core {

// Three-term recurrence written on the power coefficients, exactly as the
// reference builds its cached Legendre polynomials (LegendrePoly.cpp:62-86).
const std::vector<double> &legendre_power_coefs(int k) {
    static std::map<int, std::vector<double>> memo;
    {
        std::lock_guard<std::mutex> lk(core_mutex());
        auto it = memo.find(k);
        if (it != memo.end()) return it->second;
    }
    std::vector<double> p(k + 1, 0.0);
    if (k == 0) {
        p[0] = 1.0;
    } else if (k == 1) {
        p[1] = 1.0;
    } else {
        const auto &a = legendre_power_coefs(k - 1);
        const auto &b = legendre_power_coefs(k - 2);
        const double K = (double)k;
        p[0] = -(K - 1.0) * b[0] / K;
        for (int m = 1; m <= k; m++) {
            double t = (2.0 * K - 1.0) * a[m - 1] / K;
            p[m] = (m <= k - 2) ? t - (K - 1.0) * b[m] / K : t;
        }
    }
    std::lock_guard<std::mutex> lk(core_mutex());
    return memo.emplace(k, std::move(p)).first->second;
}

void gauss_legendre_01(int npts, std::vector<double> &x, std::vector<double> &w) {
    const double tol = 3.0e-12;
    const int max_newton = 10;
    std::vector<double> ux(npts), uw(npts);
    const int half = (npts + 1) / 2;
    for (int i = 0; i < half; i++) {
        double z = std::cos(pi * (i + 0.75) / (npts + 0.5));
        double dp = 0.0;
        int it = 0;
        for (; it < max_newton; it++) {
            // P_n(z) and P_n'(z) by the recurrence of LegendrePoly::firstDerivative
            double p = z, d = 1.0, pm = 1.0, dm = 0.0;
            if (npts == 0) {
                p = 1.0;
                d = 0.0;
            } else if (npts >= 2) {
                for (int j = 2; j <= npts; j++) {
                    double c1 = (double)j, c2 = c1 * 2.0 - 1.0, c4 = c1 - 1.0;
                    double pold = p;
                    p = (c2 * z * p - c4 * pm) / c1;
                    pm = pold;
                    double dold = d;
                    d = (c2 * z * d - c4 * dm + c2 * pm) / c1;
                    dm = dold;
                }
            }
            dp = d;
            double zprev = z;
            z = zprev - p / d;
            if (std::abs(z - zprev) <= tol) break;
        }
        if (it == max_newton) throw Error("Gauss-Legendre Newton iteration did not converge");
        ux[i] = 0.0 - 1.0 * z;
        ux[npts - 1 - i] = 0.0 + 1.0 * z;
        uw[i] = 2.e0 * 1.0 / ((1.e0 - z * z) * dp * dp);
        uw[npts - 1 - i] = uw[i];
    }
    // map [-1,1] -> [0,1] as GaussQuadrature::calcScaledPtsWgts
    x.resize(npts);
    w.resize(npts);
    for (int j = 0; j < npts; j++) {
        x[j] = ux[j] * 0.5 + 0.0 + 0.5;
        w[j] = uw[j] * 0.5;
    }
}

LegendreTables::LegendreTables(int order)
        : order_(order) {
    for (int j = 0; j <= order; j++) {
        std::vector<double> c = legendre_power_coefs(j);
        const double norm = std::sqrt(2.0 * j + 1.0);
        for (double &v : c) v = norm * v;
        power_.push_back(std::move(c));
    }
    gauss_legendre_01(order + 1, nodes_, weights_);
    vc_ = SqMat(order + 1);
    for (int j = 0; j <= order; j++)
        for (int i = 0; i <= order; i++) vc_(i, j) = phi(j, nodes_[i]) * weights_[i];
    // cvMap = vcMap^-1 (LegendreBasis::calcCVMaps, src/core/LegendreBasis.cpp:98-109)
    // by Gauss-Jordan elimination with partial pivoting
    const int n = order + 1;
    SqMat a = vc_;
    cv_ = SqMat(n);
    for (int i = 0; i < n; i++) cv_(i, i) = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::abs(a(r, c)) > std::abs(a(piv, c))) piv = r;
        if (piv != c)
            for (int j = 0; j < n; j++) {
                std::swap(a(c, j), a(piv, j));
                std::swap(cv_(c, j), cv_(piv, j));
            }
        const double d = a(c, c);
        for (int j = 0; j < n; j++) {
            a(c, j) /= d;
            cv_(c, j) /= d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a(r, c);
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) {
                a(r, j) -= f * a(c, j);
                cv_(r, j) -= f * cv_(c, j);
            }
        }
    }
}

double LegendreTables::phi(int j, double x) const {
    if (x < 0.0 || x > 1.0) return 0.0;
    const std::vector<double> &c = power_[j];
    const double q = 2.0 * x - 1.0;
    double qp = 1.0, s = 0.0;
    for (double cm : c) {
        s += qp * cm;
        qp *= q;
    }
    return s;
}

FilterBank::FilterBank(int order)
        : order_(order) {
    const int K = order + 1;
    std::vector<double> h = load_table("L_H0_" + std::to_string(order), (size_t)K * K);
    std::vector<double> g = load_table("L_G0_" + std::to_string(order), (size_t)K * K);
    SqMat H0(K), H1(K), G0(K), G1(K);
    for (int i = 0; i < K; i++)
        for (int j = 0; j < K; j++) {
            H0(i, j) = h[(size_t)i * K + j];
            G0(i, j) = g[(size_t)i * K + j];
            // Legendre symmetry: H1 = (-1)^{i+j} H0, G1 = (-1)^{i+j+K} G0
            H1(i, j) = ((i + j) % 2 ? -1.0 : 1.0) * H0(i, j);
            G1(i, j) = ((i + j + K) % 2 ? -1.0 : 1.0) * G0(i, j);
        }
    auto transpose = [K](const SqMat &m) {
        SqMat t(K);
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) t(i, j) = m(j, i);
        return t;
    };
    comp_[0] = transpose(H0);
    comp_[1] = transpose(H1);
    comp_[2] = transpose(G0);
    comp_[3] = transpose(G1);
    reco_[0] = H0;
    reco_[1] = G0;
    reco_[2] = H1;
    reco_[3] = G1;
}

CrossCorrTables::CrossCorrTables(int order)
        : order_(order) {
    const int K = order + 1;
    const size_t cnt = (size_t)K * K * 2 * K;
    left_ = load_table("L_c_left_" + std::to_string(order), cnt);
    right_ = load_table("L_c_right_" + std::to_string(order), cnt);
    for (double &v : left_)
        if (std::abs(v) < MachinePrec) v = 0.0;
    for (double &v : right_)
        if (std::abs(v) < MachinePrec) v = 0.0;
}

template <typename T> static const T &cached(int order) {
    static std::map<int, std::unique_ptr<T>> memo;
    static std::mutex m;
    std::lock_guard<std::mutex> lk(m);
    auto &slot = memo[order];
    if (!slot) slot = std::make_unique<T>(order);
    return *slot;
}

const LegendreTables &legendre(int order) { return cached<LegendreTables>(order); }
const FilterBank &filters(int order) { return cached<FilterBank>(order); }
const CrossCorrTables &cross_correlation(int order) { return cached<CrossCorrTables>(order); }

void filter_pass(double *out, const double *in, const SqMat &F, int m, bool accumulate) {
    const int n = F.n;
    for (int j = 0; j < n; j++) {
        const double *f = F.data() + (size_t)n * j;
        double *o = out + (size_t)m * j;
        for (int r = 0; r < m; r++) {
            const double *x = in + (size_t)n * r;
            double s = 0.0;
            for (int i = 0; i < n; i++) s += x[i] * f[i];
            o[r] = accumulate ? o[r] + s : s;
        }
    }
}

} // namespace core
} // namespace mrcpp
