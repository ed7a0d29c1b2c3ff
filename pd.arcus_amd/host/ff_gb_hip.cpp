// ff_gb_hip.cpp -- see ff_gb_hip.h.
#include "ff_gb_hip.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace pdarcus {

namespace {
const double kJ2kcalNa = 6.0221418E23 / 4184.0; // PhysicsConst::J2kcal * Na

void check(int rc, pdgpu_ctx *ctx, const char *what) {
    if (rc != PDGPU_OK && rc != PDGPU_ERR_NONFINITE) {
        // PD reports procedure failures as exceptions (exception.h)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? pdgpu_last_error(ctx) : "no context"));
    }
}
} // namespace

FF_GeneralizedBorn_Still_HIP::FF_GeneralizedBorn_Still_HIP(WorkSpaceView &ws, int precision)
    : ForcefieldBase(ws), Precision(precision) {
    pdgpu_options o;
    pdgpu_options_default(&o);
    name = "Generalised Born Forcefield (HIP)";
    ShortName = "GB";
    Cutoff = o.cutoff; InnerCutoff = o.inner_cutoff;
    VdwCutoff = o.vdw_cutoff; VdwInnerCutoff = o.vdw_inner_cutoff;
    Vdw14Scaling = o.vdw14_scaling; Elec14Scaling = o.elec14_scaling;
    AutoScaling14 = true;
    EnergySwitch = true; ForceSwitch = false; DoVdw = true; DoElec = true;
    FastMode = true;
    DielectricSolvent = o.dielectric_solvent; DielectricSolute = o.dielectric_solute;
    GbsaStillCutoff = o.gbsa_still_cutoff; ExpApproxThreshold = o.exp_approx_threshold;
    DielectricOffset = o.dielectric_offset; BornRadiusOffset = o.born_radius_offset;
    P1 = o.p1; P2 = o.p2; P3 = o.p3; P4 = o.p4; P5 = o.p5;
    Padding = o.nlist_padding;
}

FF_GeneralizedBorn_Still_HIP::~FF_GeneralizedBorn_Still_HIP() {
    if (ctx) pdgpu_ctx_destroy(ctx);
}

void FF_GeneralizedBorn_Still_HIP::fillOptions(pdgpu_options &o) const {
    pdgpu_options_default(&o);
    o.precision = Precision;
    o.fast_mode = FastMode ? 1 : 0;
    o.cutoff = Cutoff; o.inner_cutoff = InnerCutoff;
    o.vdw_cutoff = VdwCutoff; o.vdw_inner_cutoff = VdwInnerCutoff;
    o.vdw14_scaling = Vdw14Scaling; o.elec14_scaling = Elec14Scaling;
    o.dielectric_solute = DielectricSolute; o.dielectric_solvent = DielectricSolvent;
    o.gbsa_still_cutoff = GbsaStillCutoff; o.exp_approx_threshold = ExpApproxThreshold;
    o.dielectric_offset = DielectricOffset; o.born_radius_offset = BornRadiusOffset;
    o.p1 = P1; o.p2 = P2; o.p3 = P3; o.p4 = P4; o.p5 = P5;
    o.nlist_padding = Padding;
    // FF_NonBonded::calcEnergies mode selection (forcefields/nonbonded.cpp:1173-1178)
    o.elec_mode = !DoElec ? 0 : ForceSwitch ? 4 : EnergySwitch ? 3 : 1;
    o.vdw_mode = DoVdw ? 2 : 0;
    o.components = PDGPU_FF_GB;
}

// FF_GeneralizedBorn_Still::setup (forcefields/gbff.cpp:40-87): parameters,
// bond orders, Still fixed terms; the list is built by calcNewList.
void FF_GeneralizedBorn_Still_HIP::setup() {
    const int n = (int)wspace.nAtoms();
    if (AutoScaling14) { Vdw14Scaling = ffVdw14; Elec14Scaling = ffElec14; }
    pdgpu_options o;
    fillOptions(o);
    if (ctx) { pdgpu_ctx_destroy(ctx); ctx = nullptr; }
    check(pdgpu_ctx_create(&o, &ctx), ctx, "pdgpu_ctx_create");
    std::vector<int32_t> resid(n), a12, b12, a13, b13, a14, b14;
    std::vector<double> q(n), r(n), e(n), g(n), m(n);
    for (int i = 0; i < n; ++i) {
        const AtomParams &p = wspace.atom[i];
        resid[i] = p.ir; q[i] = p.charge; r[i] = p.radius; e[i] = p.epsilon; g[i] = p.gb_radius; m[i] = p.mass;
    }
    for (auto &pr : wspace.cov12) { a12.push_back(pr.first); b12.push_back(pr.second); }
    for (auto &pr : wspace.cov13) { a13.push_back(pr.first); b13.push_back(pr.second); }
    for (auto &pr : wspace.cov14) { a14.push_back(pr.first); b14.push_back(pr.second); }
    int32_t mid = 0, nsp = 0;
    check(pdgpu_bondorder_build(n, resid.data(), (int)a12.size(), a12.data(), b12.data(), (int)a13.size(),
                                a13.data(), b13.data(), (int)a14.size(), a14.data(), b14.data(), &mid, nullptr,
                                nullptr, nullptr, nullptr, nullptr), ctx, "bondorder");
    std::vector<int8_t> band((size_t)mid * (n > 1 ? n - 1 : 0));
    const size_t cap = a12.size() + a13.size() + a14.size() + 1;
    std::vector<int32_t> si(cap), sj(cap), so(cap);
    check(pdgpu_bondorder_build(n, resid.data(), (int)a12.size(), a12.data(), b12.data(), (int)a13.size(),
                                a13.data(), b13.data(), (int)a14.size(), a14.data(), b14.data(), &mid, band.data(),
                                &nsp, si.data(), sj.data(), so.data()), ctx, "bondorder");
    std::vector<double> vj(n), t123(n);
    check(pdgpu_fixed_terms(n, g.data(), (int)a12.size(), a12.data(), b12.data(), wspace.d12.data(),
                            (int)a13.size(), a13.data(), b13.data(), wspace.d13.data(), &o, vj.data(), t123.data()),
          ctx, "fixed terms");
    pdgpu_topology t = {};
    t.natoms = n; t.max_index_delta = mid;
    t.charge = q.data(); t.radius = r.data(); t.epsilon = e.data(); t.gb_radius = g.data();
    t.vj = vj.data(); t.terms123 = t123.data(); t.mass = m.data(); t.band = band.data();
    t.nsparse = nsp; t.sparse_i = si.data(); t.sparse_j = sj.data(); t.sparse_order = so.data();
    check(pdgpu_instance_add(ctx, &t, &id), ctx, "instance_add");
    check(pdgpu_set_positions(ctx, id, wspace.pos.data()), ctx, "set_positions");
    needsetup = false;
    Active = true;
}

void FF_GeneralizedBorn_Still_HIP::calcNewList() {
    ensuresetup();
    check(pdgpu_set_positions(ctx, id, wspace.pos.data()), ctx, "set_positions");
    check(pdgpu_nlist_build(ctx, id), ctx, "nlist_build");
}

// Adds this component's forces and energies into the workspace the way the
// reference does (forcefields/gbff.cpp:1508-1736, nonbonded.cpp:660-668).
void FF_GeneralizedBorn_Still_HIP::accumulate(const pdgpu_energies &e, const std::vector<double> &f,
                                              bool force_path) {
    for (size_t k = 0; k < f.size(); ++k) wspace.force[k] += f[k];
    wspace.ene.epot_vdw = e.epot_vdw;
    wspace.ene.epot_elec = e.epot_elec;
    wspace.ene.epot += e.epot;
    if (force_path) {
        wspace.ene.epot_pol += e.epot_pol;
        wspace.ene.epot_pol_self += e.epot_pol_self;
        wspace.ene.epot_pol_cross += e.epot_pol_cross;
    } else {
        wspace.ene.InternalVirial += e.virial;
    }
    epot_pol = e.epot_pol;
    epot_pol_self = e.epot_pol_self;
    epot_pol_cross = e.epot_pol_cross;
    if (e.status != PDGPU_OK) wspace.ene.epot = NAN; // in-band instability signal
}

void FF_GeneralizedBorn_Still_HIP::calcForces() {
    ensuresetup();
    check(pdgpu_set_positions(ctx, id, wspace.pos.data()), ctx, "set_positions");
    pdgpu_energies e;
    std::vector<double> f(3 * wspace.nAtoms());
    check(pdgpu_calc_forces(ctx, id, &e, f.data()), ctx, "calc_forces");
    accumulate(e, f, true);
}

void FF_GeneralizedBorn_Still_HIP::calcEnergies() {
    ensuresetup();
    check(pdgpu_set_positions(ctx, id, wspace.pos.data()), ctx, "set_positions");
    pdgpu_energies e;
    std::vector<double> f(3 * wspace.nAtoms());
    check(pdgpu_calc_energies(ctx, id, &e, f.data()), ctx, "calc_energies");
    accumulate(e, f, false);
}

const std::vector<double> &FF_GeneralizedBorn_Still_HIP::bornRadii() {
    born.assign(wspace.nAtoms(), 0.0);
    if (ctx) check(pdgpu_get_born_radii(ctx, id, born.data()), ctx, "born radii");
    return born;
}

const char *FF_GeneralizedBorn_Still_HIP::lastError() const { return ctx ? pdgpu_last_error(ctx) : ""; }

void FF_GeneralizedBorn_Still_HIP::info() const { // gbff.cpp:89-102
    printf(" --Generalised Born Forcefield (HIP, %s)--\n", Precision == PDGPU_MIXED ? "mixed" : "fp64");
    printf(" Fast Mode:          %s\n", FastMode ? "yes" : "no");
    printf(" Solvent Dielectric: %lf \n", DielectricSolvent);
    printf(" Solute Dielectric:  %lf \n", DielectricSolute);
    printf(" Still Cutoff:       %lf \n", GbsaStillCutoff);
    printf(" Dielec. Offset:     %lf \n", DielectricOffset);
    printf(" Exp Threshold:      %lf \n", ExpApproxThreshold);
}

void FF_GeneralizedBorn_Still_HIP::infoLine() const { // gbff.cpp:104-117
    const Hamiltonian &h = wspace.ene;
    if (FastMode) printf("\t%6.1lf\t%6.1lf", h.epot_vdw * kJ2kcalNa, h.epot_elec * kJ2kcalNa);
    printf("\t%6.1lf\t%6.1lf\t%6.1lf\t%7.2lf", epot_pol_self * kJ2kcalNa, epot_pol_cross * kJ2kcalNa,
           epot_pol * kJ2kcalNa, (epot_pol + h.epot_elec) * kJ2kcalNa);
}

void FF_GeneralizedBorn_Still_HIP::infoLineHeader() const {
    if (FastMode) printf("\t%6s\t%6s", "EVdw", "Ecoul");
    printf("\t%6s\t%6s\t%6s\t%7s", "Eself", "Ecross", "Epol", "Epol+ES");
}

// FF_NonBonded::calcCheckSum (forcefields/nonbonded.cpp:1061-1067): a change
// of the cutoffs forces a new setup.
unsigned long FF_GeneralizedBorn_Still_HIP::calcCheckSum() {
    return (unsigned long)(Cutoff * VdwCutoff * 100000.0) + 1354231 * (int)AutoScaling14;
}

// ---------------------------------------------------------------- FF_Bonded

FF_Bonded_HIP::FF_Bonded_HIP(WorkSpaceView &ws, BondedTerms t, int precision)
    : ForcefieldBase(ws), Precision(precision), terms(std::move(t)) {
    name = "Bonded Forcefield (HIP)";
    ShortName = "Bonds";
}

FF_Bonded_HIP::~FF_Bonded_HIP() {
    if (ctx) pdgpu_ctx_destroy(ctx);
}

// FF_Bonded::setup: the term lists are given; the instance carries only
// what the bonded kernels read (positions). The GB parameters of the
// topology are placeholders: this context evaluates no GB.
void FF_Bonded_HIP::setup() {
    const int n = (int)wspace.nAtoms();
    pdgpu_options o;
    pdgpu_options_default(&o);
    o.precision = Precision;
    o.components = (DoBonds ? PDGPU_FF_BONDS : 0) | (DoAngles ? PDGPU_FF_ANGLES : 0) |
                   (DoTorsions ? PDGPU_FF_TORSIONS : 0) | (DoImpropers ? PDGPU_FF_IMPROPERS : 0);
    if (ctx) { pdgpu_ctx_destroy(ctx); ctx = nullptr; }
    check(pdgpu_ctx_create(&o, &ctx), ctx, "pdgpu_ctx_create");
    std::vector<double> zero(n, 0.0), one(n, 1.0);
    std::vector<int8_t> band(n > 1 ? n - 1 : 0, 0);
    pdgpu_topology t = {};
    t.natoms = n; t.max_index_delta = 1;
    t.charge = zero.data(); t.radius = one.data(); t.epsilon = zero.data(); t.gb_radius = one.data();
    t.vj = zero.data(); t.terms123 = one.data(); t.band = band.data();
    check(pdgpu_instance_add(ctx, &t, &id), ctx, "instance_add");
    pdgpu_bonded b = {};
    b.nbonds = (int32_t)terms.bond_lk.size() / 2;
    b.nangles = (int32_t)terms.angle_tk.size() / 2;
    b.ntorsions = (int32_t)terms.torsion_terms.size();
    b.nimpropers = (int32_t)terms.improper_terms.size();
    b.bond_ij = terms.bond_ij.data(); b.bond_lk = terms.bond_lk.data();
    b.angle_iaj = terms.angle_iaj.data(); b.angle_tk = terms.angle_tk.data();
    b.torsion_iabj = terms.torsion_iabj.data(); b.torsion_terms = terms.torsion_terms.data();
    b.torsion_vn = terms.torsion_vn.data(); b.torsion_n = terms.torsion_n.data();
    b.torsion_gamma = terms.torsion_gamma.data();
    b.improper_iabj = terms.improper_iabj.data(); b.improper_terms = terms.improper_terms.data();
    b.improper_vn = terms.improper_vn.data(); b.improper_n = terms.improper_n.data();
    b.improper_gamma = terms.improper_gamma.data();
    check(pdgpu_bonded_set(ctx, id, &b), ctx, "bonded_set");
    needsetup = false;
}

// FF_Bonded::calcForces (ffbonded.cpp:809-824): forces added to the
// workspace; bonds, angles and torsions (impropers included) added to the
// Hamiltonian, bond stretching to its InternalVirial.
void FF_Bonded_HIP::calcForces() {
    ensuresetup();
    check(pdgpu_set_positions(ctx, id, wspace.pos.data()), ctx, "set_positions");
    pdgpu_energies e;
    std::vector<double> f(3 * wspace.nAtoms());
    check(pdgpu_calc_forces(ctx, id, &e, f.data()), ctx, "calc_forces");
    for (size_t k = 0; k < f.size(); ++k) wspace.force[k] += f[k];
    epot_bond = e.epot_bond;
    epot_angle = e.epot_angle;
    epot_improper = e.epot_improper;
    epot_torsion = e.epot_torsion - e.epot_improper;
    epot = e.epot;
    wspace.ene.epot_bond += e.epot_bond;
    wspace.ene.epot_angle += e.epot_angle;
    wspace.ene.epot_torsion += e.epot_torsion;
    wspace.ene.InternalVirial += e.virial;
    wspace.ene.epot += e.epot;
    if (e.status != PDGPU_OK) wspace.ene.epot = NAN;
}

void FF_Bonded_HIP::info() const { // ffbonded.cpp:826-840
    printf(" Bonds:     %8zu\n", terms.bond_lk.size() / 2);
    printf(" Angles:    %8zu\n", terms.angle_tk.size() / 2);
    printf(" Torsions:  %8zu\n", terms.torsion_terms.size());
    printf(" Impropers: %8zu\n", terms.improper_terms.size());
}

void FF_Bonded_HIP::infoLine() const { // ffbonded.cpp:842-853
    const Hamiltonian &h = wspace.ene;
    if (DoBonds) printf("\t%6.1lf", h.epot_bond * kJ2kcalNa);
    if (DoAngles) printf("\t%6.1lf", h.epot_angle * kJ2kcalNa);
    if (DoTorsions) printf("\t%6.1lf", h.epot_torsion * kJ2kcalNa);
}

void FF_Bonded_HIP::infoLineHeader() const {
    if (DoBonds) printf("\t%6s", "Ebond");
    if (DoAngles) printf("\t%6s", "Eangle");
    if (DoTorsions) printf("\t%6s", "Edihed");
}

} // namespace pdarcus
