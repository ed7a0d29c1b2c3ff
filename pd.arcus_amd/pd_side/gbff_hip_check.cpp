// gbff_hip_check.cpp -- drives the PD-side components of gbff_hip.h inside
// PD itself: builds a peptide with PD's NewProtein from a parameter file,
// evaluates PD's own Forcefield (FF_Bonded + FF_GeneralizedBorn_Still, then
// FF_NonBonded alone) and the same Forcefield with the _HIP components, and
// compares energies and forces (FP64: 1e-10 relative energy, 1e-9 force
// RMS). Needs a gfx950 device and PD's parameter file at run time.
//
// Usage: gbff_hip_check <param.ff> [sequence]
#include "system/system.h"
#include "system/genpolymer.h"
#include "forcefields/ffparam.h"
#include "forcefields/forcefield.h"
#include "gbff_hip.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using namespace Physics;

namespace {

struct Result {
    double epot;
    std::vector<double> f;
};

Result snapshot(WorkSpace &ws) {
    Result r;
    r.epot = ws.ene.epot;
    for (size_t i = 0; i < ws.atom.size(); ++i) {
        r.f.push_back(ws.cur.atom[i].f.x);
        r.f.push_back(ws.cur.atom[i].f.y);
        r.f.push_back(ws.cur.atom[i].f.z);
    }
    return r;
}

bool compare(const char *what, const Result &a, const Result &b) {
    double num = 0, den = 0;
    for (size_t k = 0; k < a.f.size(); ++k) {
        num += (a.f[k] - b.f[k]) * (a.f[k] - b.f[k]);
        den += b.f[k] * b.f[k];
    }
    const double de = std::fabs(a.epot - b.epot) / std::fabs(b.epot);
    const double df = std::sqrt(num / den);
    const bool ok = de <= 1e-10 && df <= 1e-9;
    std::printf("%s: epot %.12e (PD %.12e) rel %.2e, force rms rel %.2e %s\n", what, a.epot, b.epot, de, df,
                ok ? "ok" : "FAIL");
    return ok;
}

// PD's grammar for a linear peptide with charged termini
std::string grammar(const std::string &seq) {
    if (seq.size() == 1) return "*" + seq + "*";
    std::string mid;
    for (size_t k = 1; k + 1 < seq.size(); ++k) mid += seq[k];
    if (mid.empty()) return "*" + seq.substr(0, 1) + "-" + seq.substr(1, 1) + "*";
    return "*" + seq.substr(0, 1) + "-(" + mid + ")-" + seq.substr(seq.size() - 1) + "*";
}

} // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: gbff_hip_check <param.ff> [sequence]\n");
        return 2;
    }
    const std::string seq = argc > 2 ? argv[2] : "ACDEFGHIKLMNPQRSTVWY";
    FFParamSet ffps;
    ffps.readLib(argv[1]);
    bool ok = true;
    Result ref, gpu;
    for (int pass = 0; pass < 2; ++pass) { // FF_Bonded + GB: PD, then the _HIP components
        System sys(ffps);
        sys.add(NewProtein(ffps, grammar(seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(2.0);
        Forcefield ff(ws);
        FF_Bonded *bonded = pass ? new FF_Bonded_HIP(ws) : new FF_Bonded(ws);
        FF_GeneralizedBorn_Still *gb = pass ? new FF_GeneralizedBorn_Still_HIP(ws) : new FF_GeneralizedBorn_Still(ws);
        gb->Cutoff = 12; gb->InnerCutoff = 9; gb->VdwCutoff = 12; gb->VdwInnerCutoff = 9;
        gb->DielectricSolvent = 78.5;
        if (pass) {
            static_cast<FF_Bonded_HIP *>(bonded)->NlistPadding = 2.0;
            static_cast<FF_GeneralizedBorn_Still_HIP *>(gb)->NlistPadding = 2.0;
        }
        ff.add(*bonded);
        ff.add(*gb);
        ff.printEnergySummary(); // setup, list, calcForces
        if (pass) static_cast<FF_GeneralizedBorn_Still_HIP *>(gb)->calcNewList();
        ff.calcForces();
        (pass ? gpu : ref) = snapshot(ws);
    }
    ok &= compare("FF_Bonded + FF_GeneralizedBorn_Still", gpu, ref);
    for (int pass = 0; pass < 2; ++pass) { // FF_NonBonded alone, its own defaults
        System sys(ffps);
        sys.add(NewProtein(ffps, grammar(seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(2.0);
        Forcefield ff(ws);
        FF_NonBonded *nb = pass ? new FF_NonBonded_HIP(ws) : new FF_NonBonded(ws);
        if (pass) static_cast<FF_NonBonded_HIP *>(nb)->NlistPadding = 2.0;
        ff.add(*nb);
        ff.printEnergySummary();
        if (pass) static_cast<FF_NonBonded_HIP *>(nb)->calcNewList();
        ff.calcForces();
        (pass ? gpu : ref) = snapshot(ws);
    }
    ok &= compare("FF_NonBonded", gpu, ref);
    return ok ? 0 : 1;
}
