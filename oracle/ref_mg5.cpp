// ref_mg5.cpp -- test infrastructure: a C driver for the reference's standalone MadGraph5 QCD
// 2 -> 2 matrix elements (pythia8317/plugins/mg5mes/procs_qcd_sm, MG5_aMC 2.6.0 output, built
// without PYTHIA8 so that the model parameters come from their SLHA card).  oracle/Makefile's
// target `ref` compiles this file together with those sources, where they lie, into
// oracle/_ref/libmg5ref.so.  Only the oracle's tests load it: it is an independent
// (helicity-amplitude) check of the oracle's Pythia-form sigmaHat, massless and massive.
#include <cmath>
#include <set>
#include <vector>

#include "PY8MEs_R110_P10_sm_gg_qq.h"
#include "PY8MEs_R110_P11_sm_gq_gq.h"
#include "PY8MEs_R110_P12_sm_qq_gg.h"
#include "PY8MEs_R110_P14_sm_qq_qq.h"
#include "PY8MEs_R110_P9_sm_gg_gg.h"
#include "PY8MEs_R120_P13_sm_gg_bbx.h"
#include "PY8MEs_R120_P15_sm_qq_bbx.h"
#include "Parameters_sm.h"
#include "read_slha.h"

using namespace PY8MEs_namespace;

namespace {
Parameters_sm *pars = nullptr;
PY8ME *procs[7] = {};
}

extern "C" {

// Reads the model card; 0 on success.
int mg5ref_init(const char *param_card)
{
    try {
        SLHAReader slha(param_card);
        pars = new Parameters_sm();
        pars->setIndependentParameters(slha);
        pars->setIndependentCouplings();
        pars->setDependentParameters();
        pars->setDependentCouplings();
        procs[0] = new PY8MEs_R110_P9_sm_gg_gg(pars);
        procs[1] = new PY8MEs_R110_P10_sm_gg_qq(pars);
        procs[2] = new PY8MEs_R110_P11_sm_gq_gq(pars);
        procs[3] = new PY8MEs_R110_P12_sm_qq_gg(pars);
        procs[4] = new PY8MEs_R110_P14_sm_qq_qq(pars);
        procs[5] = new PY8MEs_R120_P13_sm_gg_bbx(pars);
        procs[6] = new PY8MEs_R120_P15_sm_qq_bbx(pars);
    } catch (...) {
        return -1;
    }
    return 0;
}

// |M|^2 of id1 id2 -> id3 id4 (averaged over the initial helicities and colours, summed over the
// final ones, with the identical-particle factor) from process class `which` (0 gg_gg, 1 gg_qq,
// 2 gq_gq, 3 qq_gg, 4 qq_qq, 5 gg_bbx, 6 qq_bbx), in the partonic rest frame at sqrt(sHat) with
// particle 3 at polar angle cos_theta to particle 1's direction, at alphaS.  *m3 receives the
// final-state mass the process uses.  Returns -1 if the flavours are not a subprocess of `which`.
double mg5ref_me2(int which, int id1, int id2, int id3, int id4, double sHat, double cos_theta, double alpS,
                  double *m3)
{
    if (which < 0 || which > 6 || !procs[which]) return -1.0;
    PY8ME *P = procs[which];
    try {
    const std::pair<std::vector<int>, int> sel = P->getPY8ME(std::vector<int>{id1, id2},
                                                             std::vector<int>{id3, id4}, std::set<int>());
    if (sel.second < 0) return -1.0;
    P->setProcID(sel.second);
    P->setPermutation(sel.first);
    P->setColors(std::vector<int>());          // sum over colours and helicities
    P->setHelicities(std::vector<int>());
    const std::vector<double> ms = P->getMasses();
    const double m = ms.size() >= 4 ? std::fmax(ms[2], ms[3]) : 0.0;
    if (m3) *m3 = m;
    const double E = 0.5 * std::sqrt(sHat);
    const double p = std::sqrt(std::fmax(E * E - m * m, 0.0));
    const double sth = std::sqrt(std::fmax(1.0 - cos_theta * cos_theta, 0.0));
    std::vector<std::vector<double>> mom = {
        {E, 0.0, 0.0, E}, {E, 0.0, 0.0, -E}, {E, p * sth, 0.0, p * cos_theta}, {E, -p * sth, 0.0, -p * cos_theta}};
    P->setMomenta(mom);
    pars->aS = alpS;
    return P->sigmaKin();
    } catch (...) {
        return -2.0;
    }
}

}  // extern "C"
