// CompFunction<3> with real components and its apply overloads
// (src/utils/CompFunction.h, src/treebuilders/apply.cpp:119-139).
#include <set>

#include "function_tree.h"

namespace mrcpp {

CompFunction<3>::CompFunction(const MultiResolutionAnalysis<3> &mra, int ncomp)
        : mra_(mra) {
    reset(ncomp);
}

void CompFunction<3>::alloc_comp(int i) {
    if (i < 0 || i > 3) throw Error("CompFunction: component index outside 0..3");
    own_[i] = std::make_unique<FunctionTree<3>>(mra_);
    CompD[i] = own_[i].get();
    ncomp_ = std::max(ncomp_, i + 1);
}

void CompFunction<3>::reset(int ncomp) {
    if (ncomp < 0 || ncomp > 4) throw Error("CompFunction: 0..4 components");
    for (int i = 0; i < 4; i++) {
        own_[i].reset();
        CompD[i] = nullptr;
    }
    ncomp_ = 0;
    for (int i = 0; i < ncomp; i++) alloc_comp(i);
    ncomp_ = ncomp;
}

void apply_components(double prec, FunctionTree<3> *const *out, ConvolutionOperator<3> &oper, FunctionTree<3> *const *inp, int ncomp,
                      const ComplexDouble (*metric)[4], int maxIter, bool absPrec) {
    if (ncomp < 0 || ncomp > 4) throw Error("apply: 0..4 components");
    if (!metric) metric = defaultMetric;
    // (input, output) component pairs in the reference's order
    std::vector<std::pair<int, int>> sel;
    for (int ic = 0; ic < ncomp; ic++)
        for (int oc = 0; oc < 4; oc++) {
            if (!(std::norm(metric[ic][oc]) > MachinePrec)) continue;
            if (std::abs(metric[ic][oc].imag()) > MachinePrec) throw Error("apply: complex metric entries need complex components");
            if (!inp[ic]) throw Error("apply: missing input component");
            if (!out[oc]) throw Error("apply: missing output component");
            sel.push_back({ic, oc});
        }
    auto rescale = [&](const std::pair<int, int> &p) {
        const double m = metric[p.first][p.second].real();
        if (std::abs(m - 1.0) > MachinePrec) out[p.second]->rescale(m);
    };
    std::set<int> outs;
    for (const auto &p : sel) outs.insert(p.second);
    if (outs.size() == sel.size()) {
        // distinct outputs: one batch
        std::vector<ApplyTask> tasks;
        for (const auto &p : sel) tasks.push_back({prec, out[p.second], &oper, inp[p.first], maxIter, absPrec});
        apply_batch(tasks);
        for (const auto &p : sel) rescale(p);
    } else {
        // an output written by several inputs: apply and rescale pair by pair,
        // the last pair's result kept (as the reference does)
        for (const auto &p : sel) {
            apply(prec, *out[p.second], oper, *inp[p.first], maxIter, absPrec);
            rescale(p);
        }
    }
}

void apply(double prec, CompFunction<3> &out, ConvolutionOperator<3> &oper, const CompFunction<3> &inp, const ComplexDouble (*metric)[4], int maxIter,
           bool absPrec) {
    if (&out == &inp) throw Error("apply: output and input must differ");
    if (out.getMRA() != inp.getMRA()) throw Error("Incompatible MRA");
    if (!metric) metric = defaultMetric;
    out.reset(inp.Ncomp());
    // output components the metric selects beyond the input's count
    for (int ic = 0; ic < inp.Ncomp(); ic++)
        for (int oc = 0; oc < 4; oc++)
            if (std::norm(metric[ic][oc]) > MachinePrec && !out.CompD[oc]) out.alloc_comp(oc);
    apply_components(prec, out.CompD, oper, inp.CompD, inp.Ncomp(), metric, maxIter, absPrec);
}

void apply(double prec, CompFunction<3> &out, ConvolutionOperator<3> &oper, CompFunction<3> &inp, FunctionTreeVector<3> *precTrees,
           const ComplexDouble (*metric)[4], int maxIter, bool absPrec) {
    if (&out == &inp) throw Error("apply: output and input must differ");
    if (out.getMRA() != inp.getMRA()) throw Error("Incompatible MRA");
    if (!precTrees) throw Error("apply: null precision tree vectors");
    if (!metric) metric = defaultMetric;
    out.reset(inp.Ncomp());
    // (input, output) pairs in the reference's order, one apply each, the
    // output components being the ones paramCopy allocates
    for (int ic = 0; ic < inp.Ncomp(); ic++)
        for (int oc = 0; oc < 4; oc++) {
            if (!(std::norm(metric[ic][oc]) > MachinePrec)) continue;
            if (std::abs(metric[ic][oc].imag()) > MachinePrec) throw Error("apply: complex metric entries need complex components");
            if (!out.CompD[oc]) throw Error("apply: the metric selects an output component beyond the input's count");
            apply(prec, *out.CompD[oc], oper, *inp.CompD[ic], precTrees[ic], maxIter, absPrec);
            const double m = metric[ic][oc].real();
            if (std::abs(m - 1.0) > MachinePrec) out.CompD[oc]->rescale(m);
        }
}

} // namespace mrcpp
