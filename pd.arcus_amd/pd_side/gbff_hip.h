// gbff_hip.h -- the PD-side binding of the MI355X path: forcefield
// components a PD maintainer adds next to forcefields/gbff.h. They derive
// from PD's own classes (FF_GeneralizedBorn_Still, forcefields/gbff.h:120-200;
// FF_Bonded, forcefields/ffbonded.h:103-210; FF_NonBonded,
// forcefields/nonbonded.h:45-230), reuse the parent's knobs and setup (which
// computes V_j, terms123, the 1-4 scalings and the bond-order band), and
// forward setup / calcForces / calcEnergies to the C ABI of include/pdgpu.h.
// PD's protocols then call them exactly like the CPU components they
// replace. Written in C++98, the dialect PD builds in; link with -lpdgpu.
#ifndef PDGPU_GBFF_HIP_H
#define PDGPU_GBFF_HIP_H

#include <cstring>
#include <string>
#include <vector>

#include "forcefields/ffbonded.h"
#include "forcefields/gbff.h"
#include "forcefields/nonbonded.h"
#include "workspace/bondorder.h"
#include "workspace/neighbourlist.h"
#include "workspace/workspace.h"
#include "pdgpu.h"

namespace Physics {

// Device context of one component over one workspace: uploads the
// positions before an evaluation and adds the results into the workspace,
// in PD's units (A, N, J/molecule).
class PdgpuComponentBase {
public:
    PdgpuComponentBase() : NlistPadding(0.5), ctx(0), id(-1) {}
    // a copy (PD's clone()) gets its own context at its own setup
    PdgpuComponentBase(const PdgpuComponentBase &o) : NlistPadding(o.NlistPadding), ctx(0), id(-1) {}
    PdgpuComponentBase &operator=(const PdgpuComponentBase &o) {
        if (this != &o) {
            if (ctx) pdgpu_ctx_destroy(ctx);
            NlistPadding = o.NlistPadding;
            ctx = 0;
            id = -1;
        }
        return *this;
    }
    ~PdgpuComponentBase() { if (ctx) pdgpu_ctx_destroy(ctx); }
    // NeighbourList::Padding of the workspace (protected in PD, so the
    // caller passes the value it gave setPadding; PD's default is 0.5 A,
    // workspace/neighbourlist.cpp:18). The device list holds the same
    // candidates as PD's only if the two agree.
    double NlistPadding;

protected:
    pdgpu_ctx *ctx;
    int32_t id;

    void create(const pdgpu_options &o) {
        if (ctx) pdgpu_ctx_destroy(ctx);
        ctx = 0;
        if (pdgpu_ctx_create(&o, &ctx) != PDGPU_OK) throw ProcedureException("pdgpu_ctx_create failed");
    }
    void check(int rc) const {
        if (rc != PDGPU_OK && rc != PDGPU_ERR_NONFINITE) throw ProcedureException(pdgpu_last_error(ctx));
    }
    static void positions(WorkSpace &ws, std::vector<double> &x) {
        const int n = (int)ws.atom.size();
        x.resize(3 * n);
        for (int i = 0; i < n; ++i) {
            x[3 * i] = ws.cur.atom[i].p.x;
            x[3 * i + 1] = ws.cur.atom[i].p.y;
            x[3 * i + 2] = ws.cur.atom[i].p.z;
        }
    }
    // the topology every component uploads: per-atom parameters of the
    // workspace and its BondOrder (band + off-diagonal pairs)
    void add_instance(WorkSpace &ws, const std::vector<double> &gb_radius, const std::vector<double> &vj,
                      const std::vector<double> &terms123) {
        const int n = (int)ws.atom.size();
        q.resize(n); r.resize(n); e.resize(n); m.resize(n);
        for (int i = 0; i < n; ++i) {
            q[i] = ws.atom[i].charge;
            r[i] = ws.atom[i].radius;
            e[i] = ws.atom[i].epsilon;
            m[i] = ws.atom[i].mass;
        }
        const BondOrder &bo = ws.bondorder();
        si.clear(); sj.clear(); so.clear();
        const std::vector<LongBondOrder> &sp = bo.getOffDiagData();
        for (size_t p = 0; p < sp.size(); ++p)
            for (size_t k = 0; k < sp[p].partner.size(); ++k) {
                si.push_back(sp[p].i);
                sj.push_back(sp[p].partner[k].j);
                so.push_back(sp[p].partner[k].order);
            }
        pdgpu_topology tp;
        std::memset(&tp, 0, sizeof(tp));
        tp.natoms = n;
        tp.max_index_delta = bo.getMaxIndexDelta();
        tp.charge = &q[0]; tp.radius = &r[0]; tp.epsilon = &e[0]; tp.mass = &m[0];
        tp.gb_radius = &gb_radius[0]; tp.vj = &vj[0]; tp.terms123 = &terms123[0];
        tp.band = bo.getDiagData().empty() ? 0 : (const int8_t *)&bo.getDiagData()[0];
        tp.nsparse = (int32_t)si.size();
        if (!si.empty()) { tp.sparse_i = &si[0]; tp.sparse_j = &sj[0]; tp.sparse_order = &so[0]; }
        check(pdgpu_instance_add(ctx, &tp, &id));
        std::vector<double> x;
        positions(ws, x);
        check(pdgpu_set_positions(ctx, id, &x[0]));
    }
    // one evaluation at the workspace's positions; forces are added to
    // cur.atom[].f, the energies of the record are returned
    pdgpu_energies evaluate(WorkSpace &ws, bool force_path) {
        std::vector<double> x, f;
        positions(ws, x);
        f.assign(x.size(), 0.0);
        check(pdgpu_set_positions(ctx, id, &x[0]));
        pdgpu_energies en;
        check(force_path ? pdgpu_calc_forces(ctx, id, &en, &f[0]) : pdgpu_calc_energies(ctx, id, &en, &f[0]));
        for (size_t i = 0; i < ws.atom.size(); ++i) {
            ws.cur.atom[i].f.x += f[3 * i];
            ws.cur.atom[i].f.y += f[3 * i + 1];
            ws.cur.atom[i].f.z += f[3 * i + 2];
        }
        return en;
    }
    // rebuild of the device tiles at the moment PD rebuilds its own list
    void rebuild(WorkSpace &ws) {
        std::vector<double> x;
        positions(ws, x);
        check(pdgpu_set_positions(ctx, id, &x[0]));
        check(pdgpu_nlist_build(ctx, id));
    }

private:
    std::vector<double> q, r, e, m;
    std::vector<int32_t> si, sj, so;
};

// Knobs of FF_NonBonded shared by the GB and the vacuum components
// (forcefields/nonbonded.h:55-126).
inline void pdgpu_nonbonded_knobs(const FF_NonBonded &nb, double padding, pdgpu_options &o) {
    o.cutoff = nb.Cutoff;
    o.inner_cutoff = nb.InnerCutoff;
    o.vdw_cutoff = nb.VdwCutoff;
    o.vdw_inner_cutoff = nb.VdwInnerCutoff;
    o.vdw14_scaling = nb.Vdw14Scaling;
    o.elec14_scaling = nb.Elec14Scaling;
    o.nlist_padding = padding;
}

// FF_GeneralizedBorn_Still on the device (FastMode, the reference default).
class PD_API FF_GeneralizedBorn_Still_HIP : public FF_GeneralizedBorn_Still, public PdgpuComponentBase {
public:
    FF_GeneralizedBorn_Still_HIP(WorkSpace &ws, int precision = PDGPU_FP64)
        : FF_GeneralizedBorn_Still(ws), Precision(precision) {}
    virtual FF_GeneralizedBorn_Still_HIP *clone() const {
        FF_GeneralizedBorn_Still_HIP *c = new FF_GeneralizedBorn_Still_HIP(*this);
        c->forceSetup(); // the copy has no device context yet
        return c;
    }
    int Precision; // PDGPU_FP64 or PDGPU_MIXED

    // to be called where PD rebuilds its list (ProtocolBase::refreshNeighborList,
    // Forcefield::printEnergySummary)
    void calcNewList() { rebuild(getWSpace()); }

protected:
    virtual void setup() {
        FF_GeneralizedBorn_Still::setup(); // V_j, terms123, GB type radii, 1-4 scalings
        WorkSpace &ws = getWSpace();
        pdgpu_options o;
        pdgpu_options_default(&o);
        o.precision = Precision;
        o.components = PDGPU_FF_GB;
        pdgpu_nonbonded_knobs(*this, NlistPadding, o);
        o.dielectric_solute = DielectricSolute;
        o.dielectric_solvent = DielectricSolvent;
        o.gbsa_still_cutoff = GbsaStillCutoff;
        o.exp_approx_threshold = ExpApproxThreshold;
        o.dielectric_offset = DielectricOffset;
        o.born_radius_offset = BornRadiusOffset;
        o.p1 = P1; o.p2 = P2; o.p3 = P3; o.p4 = P4; o.p5 = P5;
        create(o);
        const int n = (int)ws.atom.size();
        std::vector<double> g(n), vj(n), t(n);
        for (int i = 0; i < n; ++i) {
            g[i] = GBtype[ws.atom[i].FFType].radius;
            vj[i] = GB_atom_param_still[i].Vj;
            t[i] = GB_atom_param_still[i].terms123;
        }
        add_instance(ws, g, vj, t);
        needsetup = false;
    }
    virtual void calcForces() { add(evaluate(getWSpace(), true), true); }
    virtual void calcEnergies() { add(evaluate(getWSpace(), false), false); }

private:
    // what FF_GeneralizedBorn_Still leaves in the Hamiltonian (FastMode
    // writes the totals, gbff.cpp:1732-1736; the energy path also the
    // virial of its FF_NonBonded part)
    void add(const pdgpu_energies &en, bool force_path) {
        WorkSpace &ws = getWSpace();
        ws.ene.epot_vdw = en.epot_vdw;
        ws.ene.epot_elec = en.epot_elec;
        ws.ene.epot_pol += en.epot_pol;
        ws.ene.epot_pol_self += en.epot_pol_self;
        ws.ene.epot_pol_cross += en.epot_pol_cross;
        ws.ene.epot += en.epot;
        if (!force_path) ws.ene.InternalVirial += en.virial;
        epot_pol_self = en.epot_pol_self;
        epot_pol_cross = en.epot_pol_cross;
        if (en.status != PDGPU_OK) ws.ene.epot = Maths::dNAN(); // PD's in-band instability signal
    }
};

// FF_NonBonded on the device: vacuum LJ + Coulomb with the component's
// switching (force switch by default, nonbonded.cpp:986-987).
class PD_API FF_NonBonded_HIP : public FF_NonBonded, public PdgpuComponentBase {
public:
    FF_NonBonded_HIP(WorkSpace &ws, int precision = PDGPU_FP64) : FF_NonBonded(ws), Precision(precision) {}
    virtual FF_NonBonded_HIP *clone() const {
        FF_NonBonded_HIP *c = new FF_NonBonded_HIP(*this);
        c->forceSetup(); // the copy has no device context yet
        return c;
    }
    int Precision;
    void calcNewList() { rebuild(getWSpace()); }

protected:
    virtual void setup() {
        FF_NonBonded::setup(); // list cutoff request, AutoScaling14
        WorkSpace &ws = getWSpace();
        pdgpu_options o;
        pdgpu_options_default_nonbonded(&o);
        o.precision = Precision;
        o.components = PDGPU_FF_NONBONDED;
        pdgpu_nonbonded_knobs(*this, NlistPadding, o);
        o.dielectric_solute = Dielectric;
        o.elec_mode = !DoElec ? 0 : ForceSwitch ? 4 : EnergySwitch ? 3 : 1;
        o.vdw_mode = DoVdw ? 2 : 0;
        create(o);
        // GB inputs are not read by this component; any finite values do
        const int n = (int)ws.atom.size();
        std::vector<double> g(n, 1.5), vj(n, 0.0), t(n, -1.0);
        add_instance(ws, g, vj, t);
        needsetup = false;
    }
    virtual void calcForces() { add(evaluate(getWSpace(), true)); }
    virtual void calcEnergies() { add(evaluate(getWSpace(), true)); } // UsePartialRecalc off

private:
    void add(const pdgpu_energies &en) {
        WorkSpace &ws = getWSpace();
        ws.ene.epot_vdw = en.epot_vdw;
        ws.ene.epot_elec = en.epot_elec;
        ws.ene.epot += en.epot;
        ws.ene.InternalVirial += en.virial;
        epot_vdw = en.epot_vdw;
        epot_elec = en.epot_elec;
        if (en.status != PDGPU_OK) ws.ene.epot = Maths::dNAN();
    }
};

// FF_Bonded on the device: the term lists FF_Bonded::setup assembles
// (forcefields/ffbonded.cpp:13-486) go to pdgpu_bonded_set.
class PD_API FF_Bonded_HIP : public FF_Bonded, public PdgpuComponentBase {
public:
    FF_Bonded_HIP(WorkSpace &ws, int precision = PDGPU_FP64) : FF_Bonded(ws), Precision(precision) {}
    virtual FF_Bonded_HIP *clone() const {
        FF_Bonded_HIP *c = new FF_Bonded_HIP(*this);
        c->forceSetup(); // the copy has no device context yet
        return c;
    }
    int Precision;

protected:
    virtual void setup() {
        FF_Bonded::setup(); // bond, angle, torsion, improper lists
        WorkSpace &ws = getWSpace();
        pdgpu_options o;
        pdgpu_options_default(&o);
        o.precision = Precision;
        o.components = (DoBonds ? PDGPU_FF_BONDS : 0) | (DoAngles ? PDGPU_FF_ANGLES : 0) |
                       (DoTorsions ? PDGPU_FF_TORSIONS : 0) | (DoImpropers ? PDGPU_FF_IMPROPERS : 0);
        create(o);
        const int n = (int)ws.atom.size();
        std::vector<double> g(n, 1.5), vj(n, 0.0), t(n, -1.0);
        add_instance(ws, g, vj, t);
        std::vector<int32_t> bi, ai, ti, tt, ii, it;
        std::vector<double> bp, ap, tv, tn, tg, iv, in, ig;
        for (size_t k = 0; k < bond.size(); ++k) {
            bi.push_back(bond[k].i); bi.push_back(bond[k].j);
            bp.push_back(bond[k].l); bp.push_back(bond[k].k);
        }
        for (size_t k = 0; k < angle.size(); ++k) {
            ai.push_back(angle[k].i); ai.push_back(angle[k].a); ai.push_back(angle[k].j);
            ap.push_back(angle[k].theta0); ap.push_back(angle[k].k);
        }
        dihedrals(torsion, ti, tt, tv, tn, tg);
        dihedrals(improper, ii, it, iv, in, ig);
        pdgpu_bonded b;
        std::memset(&b, 0, sizeof(b));
        b.nbonds = (int32_t)bond.size(); b.nangles = (int32_t)angle.size();
        b.ntorsions = (int32_t)torsion.size(); b.nimpropers = (int32_t)improper.size();
        b.bond_ij = data(bi); b.bond_lk = data(bp); b.angle_iaj = data(ai); b.angle_tk = data(ap);
        b.torsion_iabj = data(ti); b.torsion_terms = data(tt);
        b.torsion_vn = data(tv); b.torsion_n = data(tn); b.torsion_gamma = data(tg);
        b.improper_iabj = data(ii); b.improper_terms = data(it);
        b.improper_vn = data(iv); b.improper_n = data(in); b.improper_gamma = data(ig);
        check(pdgpu_bonded_set(ctx, id, &b));
        needsetup = false;
    }
    virtual void calcForces() { add(evaluate(getWSpace(), true)); }
    virtual void calcEnergies() { add(evaluate(getWSpace(), true)); } // FF_Bonded::calcEnergies is calcForces

private:
    template <typename T> static const T *data(const std::vector<T> &v) { return v.empty() ? 0 : &v[0]; }
    static void dihedrals(const std::vector<Torsion> &d, std::vector<int32_t> &idx, std::vector<int32_t> &terms,
                          std::vector<double> &vn, std::vector<double> &nn, std::vector<double> &gamma) {
        for (size_t k = 0; k < d.size(); ++k) {
            idx.push_back(d[k].i); idx.push_back(d[k].a); idx.push_back(d[k].b); idx.push_back(d[k].j);
            terms.push_back(d[k].terms);
            for (int m = 0; m < 4; ++m) {
                const bool used = m < d[k].terms;
                vn.push_back(used ? d[k].Vn[m] : 0.0);
                nn.push_back(used ? d[k].n[m] : 0.0);
                gamma.push_back(used ? d[k].gamma[m] : 0.0);
            }
        }
    }
    void add(const pdgpu_energies &en) {
        WorkSpace &ws = getWSpace();
        ws.ene.epot_bond += en.epot_bond;
        ws.ene.epot_angle += en.epot_angle;
        ws.ene.epot_torsion += en.epot_torsion;
        ws.ene.epot += en.epot;
        ws.ene.InternalVirial += en.virial;
        epot_bond = en.epot_bond;
        epot_angle = en.epot_angle;
        epot_torsion = en.epot_torsion - en.epot_improper;
        epot_improper = en.epot_improper;
        if (en.status != PDGPU_OK) ws.ene.epot = Maths::dNAN();
    }
};

} // namespace Physics

#endif
