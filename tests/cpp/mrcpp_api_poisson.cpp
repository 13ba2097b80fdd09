// A host program written against MRCPP's public C++ API (the headers
// "MRCPP/MWFunctions", "MRCPP/MWOperators", "MRCPP/Gaussians",
// "MRCPP/Printer", "MRCPP/Timer"), built with g++ against this library.
// It follows the flow of MRCPP's Poisson example and tests
// (examples/poisson.cpp, tests/operators/poisson_operator.cpp:129-154):
// a unit-charge Gaussian projected on a Legendre basis, the Poisson operator
// applied, and the Coulomb self-energy compared with the closed form; then
// the precTrees and CompFunction overloads, add / multiply over
// FunctionTreeVector, and projection of a std::function.
// Exit status 0 when every check holds; one line per check on stdout.
#include <cmath>
#include <cstdio>
#include <functional>

#include "MRCPP/Gaussians"
#include "MRCPP/MWFunctions"
#include "MRCPP/MWOperators"
#include "MRCPP/Printer"
#include "MRCPP/Timer"

namespace {
int failures = 0;
void check(bool ok, const char *what, double a, double b) {
    std::printf("%s %-44s %.12e %.12e\n", ok ? "ok  " : "FAIL", what, a, b);
    if (!ok) failures++;
}
} // namespace

int run() {
    constexpr int D = 3;
    const double prec = 1.0e-5;
    mrcpp::Timer total;
    mrcpp::Printer::init(0);
    mrcpp::print::environment(0);
    mrcpp::print::header(0, "Poisson operator through the MRCPP C++ API");

    // world [-16, 16]^3: root scale -4, 2 x 2 x 2 root boxes from corner (-1,-1,-1)
    mrcpp::BoundingBox<D> world(-4, std::array<int, D>{-1, -1, -1}, std::array<int, D>{2, 2, 2});
    mrcpp::LegendreBasis basis(7);
    mrcpp::MultiResolutionAnalysis<D> mra(world, basis, 25);

    const double beta = 100.0;
    const double alpha = std::pow(beta / mrcpp::pi, 1.5);
    const mrcpp::Coord<D> pos{mrcpp::pi / 3.0, mrcpp::pi / 3.0, mrcpp::pi / 3.0};
    mrcpp::GaussFunc<D> gauss(beta, alpha, pos, std::array<int, D>{0, 0, 0});
    const double e_ana = gauss.calcCoulombEnergy(gauss);
    check(std::abs(e_ana - std::sqrt(2.0 * beta / mrcpp::pi)) <= 1e-14 * e_ana, "Coulomb self-energy closed form", e_ana,
          std::sqrt(2.0 * beta / mrcpp::pi));

    mrcpp::Timer t_op;
    mrcpp::PoissonOperator P(mra, prec);
    t_op.stop();
    mrcpp::print::time(0, "PoissonOperator", t_op);

    mrcpp::Timer t_proj;
    mrcpp::FunctionTree<D> f(mra);
    mrcpp::build_grid(f, gauss);
    mrcpp::project(prec, f, gauss);
    t_proj.stop();

    mrcpp::Timer t_apply;
    mrcpp::FunctionTree<D> g(mra);
    mrcpp::apply(prec, g, P, f);
    t_apply.stop();
    mrcpp::print::tree(0, "f", f, t_proj);
    mrcpp::print::tree(0, "g = P f", g, t_apply);

    const double q = f.integrate();
    const double e_num = mrcpp::dot(g, f);
    check(std::abs(q - 1.0) <= 10.0 * prec, "charge of the projected Gaussian", q, 1.0);
    check(std::abs(e_num - e_ana) <= 10.0 * prec * e_ana, "dot(P f, f) against the closed form", e_num, e_ana);
    check(g.getSquareNorm() > 0.0 && std::isfinite(g.getSquareNorm()), "square norm of P f is finite", g.getSquareNorm(), 0.0);

    // potential at a point 1 bohr from the centre: erf(sqrt(beta) r) / r
    const mrcpp::Coord<D> r1{pos[0] + 1.0, pos[1], pos[2]};
    const double v = g.evalf(r1);
    check(std::abs(v - std::erf(std::sqrt(beta))) <= 10.0 * prec, "P f at r = 1 against erf(sqrt(beta) r) / r", v, std::erf(std::sqrt(beta)));

    // precTrees overload with a FunctionTreeVector: precision relative to f's
    // local norms (src/treebuilders/apply.cpp:214-251); the result keeps the energy
    mrcpp::FunctionTreeVector<D> precTrees;
    precTrees.push_back(std::make_tuple(1.0, &f));
    mrcpp::FunctionTree<D> gp(mra);
    mrcpp::apply(prec, gp, P, f, precTrees);
    const double e_prec = mrcpp::dot(gp, f);
    check(std::abs(e_prec - e_ana) <= 10.0 * prec * e_ana, "precTrees apply: dot(P f, f)", e_prec, e_ana);
    check(mrcpp::get_n_nodes(precTrees) == f.getNNodes(), "FunctionTreeVector node count", mrcpp::get_n_nodes(precTrees), f.getNNodes());

    // CompFunction overload, one real component, metric 2 on the diagonal
    mrcpp::CompFunction<D> cf(mra, 1), cg(mra);
    mrcpp::project(prec, *cf.CompD[0], gauss);
    mrcpp::ComplexDouble metric[4][4] = {{2.0, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    mrcpp::apply(prec, cg, P, cf, metric);
    const double e_comp = mrcpp::dot(*cg.CompD[0], f);
    check(std::abs(e_comp - 2.0 * e_num) <= 1e-10 * e_num, "CompFunction apply with metric 2", e_comp, 2.0 * e_num);

    // add and multiply over FunctionTreeVector: (P f) - (P f) is zero,
    // f * 1 integrates to the charge
    mrcpp::FunctionTreeVector<D> terms;
    terms.push_back(std::make_tuple(1.0, &g));
    terms.push_back(std::make_tuple(-1.0, &g));
    mrcpp::FunctionTree<D> zero(mra);
    mrcpp::add(prec, zero, terms);
    check(zero.getSquareNorm() <= 1e-20, "add: P f - P f", zero.getSquareNorm(), 0.0);
    mrcpp::FunctionTree<D> one(mra);
    mrcpp::project(prec, one, [](const mrcpp::Coord<D> &) { return 1.0; });
    mrcpp::FunctionTree<D> prod(mra);
    mrcpp::multiply(prec, prod, 1.0, f, one);
    check(std::abs(prod.integrate() - q) <= 10.0 * prec, "multiply: f * 1 integrates to the charge", prod.integrate(), q);

    // std::function projection of the same Gaussian equals the GaussFunc one
    mrcpp::FunctionTree<D> h(mra);
    std::function<double(const mrcpp::Coord<D> &)> fn = [&](const mrcpp::Coord<D> &r) { return gauss.evalf(r); };
    mrcpp::build_grid(h, gauss);
    mrcpp::project(prec, h, fn);
    check(std::abs(h.integrate() - q) <= 1e-10, "std::function projection integrates like GaussFunc", h.integrate(), q);

    mrcpp::print::value(0, "Analytic energy", e_ana, "(au)");
    mrcpp::print::value(0, "Numerical energy", e_num, "(au)");
    mrcpp::print::value(0, "Error", (e_num - e_ana) / e_num, "(rel)", 1);
    mrcpp::print::footer(0, total, 2);
    std::printf("%s: %d failed checks\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}

int main() {
    try {
        return run();
    } catch (const std::exception &e) {
        std::printf("ERROR: %s\n", e.what());
        return 2;
    }
}
