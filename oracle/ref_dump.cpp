// ref_dump.cpp -- drives the compiled reference (oracle/_ref/libmmlib.a) to
// write golden vectors for the hot path. Test infrastructure: it runs only in
// the build container, where /root/reference and its parameter files exist.
//
// Usage: ref_dump <param.ff> <outdir>
// For every case it builds a peptide with PD's NewProtein (extended builder,
// system/genpolymer.cpp:256-270), sets up FF_Bonded (only to obtain the
// equilibrium 1-2/1-3 distances GB's fixed terms read) and evaluates
// FF_GeneralizedBorn_Still alone, writing inputs and outputs as a flat
// record file <outdir>/<case>.bin read by tests/golden/make_golden.py.
#include "global.h"
#include "system/system.h"
#include "system/genpolymer.h"
#include "workspace/workspace.h"
#include "workspace/neighbourlist.h"
#include "workspace/bondorder.h"
#include "forcefields/ffparam.h"
#include "forcefields/forcefield.h"
#include "forcefields/ffbonded.h"
#include "forcefields/nonbonded.h"
#include "forcefields/gbff.h"
#include "forcefields/ffsoftvdw.h"
#include "protocols/minimise.h"
#include "protocols/md.h"
#include "fileio/psfdcd.h"
#include "fileio/tra.h"

#include <stdint.h>
#include <cstdio>
#include <string>
#include <vector>

using namespace Physics;

namespace {

// --- flat record writer: [u32 namelen][name][u8 dtype][u64 count][data] ---
struct Rec {
    FILE *f;
    explicit Rec(const std::string &path) { f = fopen(path.c_str(), "wb"); }
    ~Rec() { if (f) fclose(f); }
    void put(const std::string &name, char dt, const void *data, uint64_t count, size_t esz) {
        uint32_t nl = (uint32_t)name.size();
        fwrite(&nl, 4, 1, f);
        fwrite(name.data(), 1, nl, f);
        fwrite(&dt, 1, 1, f);
        fwrite(&count, 8, 1, f);
        if (count) fwrite(data, esz, count, f);
    }
    void d(const std::string &n, const std::vector<double> &v) { put(n, 'd', v.empty() ? 0 : &v[0], v.size(), 8); }
    void i(const std::string &n, const std::vector<int32_t> &v) { put(n, 'i', v.empty() ? 0 : &v[0], v.size(), 4); }
    void u(const std::string &n, const std::vector<uint32_t> &v) { put(n, 'u', v.empty() ? 0 : &v[0], v.size(), 4); }
    void b(const std::string &n, const std::vector<int8_t> &v) { put(n, 'b', v.empty() ? 0 : &v[0], v.size(), 1); }
    void s(const std::string &n, double x) { std::vector<double> v(1, x); d(n, v); }
};

// Gives the driver read access to GB internals it needs to dump.
class GBProbe : public FF_GeneralizedBorn_Still {
public:
    explicit GBProbe(WorkSpace &ws) : FF_GeneralizedBorn_Still(ws) {}
    double typeRadius(int fftype) const { return GBtype[fftype].radius; }
    double vj(int i) const { return GB_atom_param_still[i].Vj; }
    double terms123(int i) const { return GB_atom_param_still[i].terms123; }
    double born(int i) const { return GB_atom_param[i].bornradius; }
    double deda(int i) const { return GB_atom_param[i].deda; }
    double polSelf() const { return epot_pol_self; }
    double polCross() const { return epot_pol_cross; }
    double pol() const { return epot_pol; }
    double p(int k) const { const double v[5] = {P1, P2, P3, P4, P5}; return v[k]; }
};

// Gives the driver read access to FF_Bonded's assembled term lists
// (forcefields/ffbonded.h:204-207).
class BondedProbe : public FF_Bonded {
public:
    explicit BondedProbe(WorkSpace &ws) : FF_Bonded(ws) {}
    const std::vector<Bond> &bonds() const { return bond; }
    const std::vector<Angle> &angles() const { return angle; }
    const std::vector<Torsion> &torsions() const { return torsion; }
    const std::vector<Torsion> &impropers() const { return improper; }
};

// Uniform deviate in [-1, 1) from a SplitMix64 state.
double splitmix_unit(uint64_t &x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

void dump_dihedrals(Rec &r, const std::vector<Torsion> &t, const std::string &tag) {
    std::vector<int32_t> idx, terms;
    std::vector<double> vn, nn, gamma;
    for (size_t k = 0; k < t.size(); ++k) {
        idx.push_back(t[k].i); idx.push_back(t[k].a); idx.push_back(t[k].b); idx.push_back(t[k].j);
        terms.push_back(t[k].terms);
        for (int m = 0; m < 4; ++m) {
            const bool used = m < t[k].terms;
            vn.push_back(used ? t[k].Vn[m] : 0.0);
            nn.push_back(used ? t[k].n[m] : 0.0);
            gamma.push_back(used ? t[k].gamma[m] : 0.0);
        }
    }
    r.i(tag + "_idx", idx); r.i(tag + "_terms", terms);
    r.d(tag + "_vn", vn); r.d(tag + "_n", nn); r.d(tag + "_gamma", gamma);
}

void dump_bonded_terms(Rec &r, const BondedProbe &b) {
    std::vector<int32_t> bi, ai;
    std::vector<double> bp, ap;
    for (size_t k = 0; k < b.bonds().size(); ++k) {
        const Bond &x = b.bonds()[k];
        bi.push_back(x.i); bi.push_back(x.j); bp.push_back(x.l); bp.push_back(x.k);
    }
    for (size_t k = 0; k < b.angles().size(); ++k) {
        const Angle &x = b.angles()[k];
        ai.push_back(x.i); ai.push_back(x.a); ai.push_back(x.j); ap.push_back(x.theta0); ap.push_back(x.k);
    }
    r.i("bd_idx", bi); r.d("bd_lk", bp);
    r.i("an_idx", ai); r.d("an_tk", ap);
    dump_dihedrals(r, b.torsions(), "to");
    dump_dihedrals(r, b.impropers(), "im");
}

void dump_forces(Rec &r, WorkSpace &ws, const std::string &tag);

// The bytes of a file the reference wrote.
std::vector<int8_t> slurp(const std::string &path) {
    std::vector<int8_t> v;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return v;
    int ch;
    while ((ch = fgetc(f)) != EOF) v.push_back((int8_t)ch);
    fclose(f);
    return v;
}

void dump_bonded_eval(Rec &r, WorkSpace &ws, const BondedProbe &b, const std::string &tag) {
    dump_forces(r, ws, tag + "_force");
    r.s(tag + "_epot", ws.ene.epot);
    r.s(tag + "_bond", b.getEBond()); r.s(tag + "_angle", b.getEAngle());
    r.s(tag + "_torsion", b.getETorsion()); r.s(tag + "_improper", b.getEImproper());
}

// SplitMix64 over the 20 standard residues.
std::string random_sequence(uint64_t seed, int len) {
    static const char *aa = "ACDEFGHIKLMNPQRSTVWY";
    uint64_t x = seed;
    std::string s;
    for (int k = 0; k < len; ++k) {
        x += 0x9E3779B97F4A7C15ULL;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        s += aa[z % 20];
    }
    return s;
}

std::string pd_grammar(const std::string &seq) {
    if (seq.size() == 1) return "*" + seq + "*";
    if (seq.size() == 2) return "*" + seq.substr(0, 1) + "-" + seq.substr(1, 1) + "*";
    std::string mid;
    for (size_t k = 1; k + 1 < seq.size(); ++k) mid += seq[k];
    return "*" + seq.substr(0, 1) + "-(" + mid + ")-" + seq.substr(seq.size() - 1) + "*";
}

struct Case {
    std::string name, seq;
    double cutoff, inner, vdw_cutoff, vdw_inner, padding, eps_out;
    bool set_vdw;   // false: leave VdwCutoff/VdwInnerCutoff at PD's defaults (quirk Q6)
    int sd_steps;   // 0: no minimisation record
    int md_steps;   // 0: no molecular-dynamics record
};

void dump_inputs(Rec &r, WorkSpace &ws, GBProbe &gb) {
    const int n = (int)ws.atom.size();
    std::vector<double> xyz, q, rad, eps, gbr, vj, t123, mass;
    std::vector<int32_t> resid;
    std::vector<int32_t> a12, b12, a13, b13, a14, b14;
    std::vector<double> d12, d13;
    for (int i = 0; i < n; ++i) {
        xyz.push_back(ws.cur.atom[i].p.x);
        xyz.push_back(ws.cur.atom[i].p.y);
        xyz.push_back(ws.cur.atom[i].p.z);
        q.push_back(ws.atom[i].charge);
        rad.push_back(ws.atom[i].radius);
        eps.push_back(ws.atom[i].epsilon);
        gbr.push_back(gb.typeRadius(ws.atom[i].FFType));
        vj.push_back(gb.vj(i));
        t123.push_back(gb.terms123(i));
        mass.push_back(ws.atom[i].mass);
        resid.push_back(ws.atom[i].ir);
        for (size_t k = 0; k < ws.atom[i].cov12atom.size(); ++k) {
            int j = ws.atom[i].cov12atom[k].i;
            if (j > i) { a12.push_back(i); b12.push_back(j); d12.push_back(ws.atom[i].cov12atom[k].d); }
        }
        for (size_t k = 0; k < ws.atom[i].cov13atom.size(); ++k) {
            int j = ws.atom[i].cov13atom[k].i;
            if (j > i) { a13.push_back(i); b13.push_back(j); d13.push_back(ws.atom[i].cov13atom[k].d); }
        }
        for (size_t k = 0; k < ws.atom[i].cov14atom.size(); ++k) {
            int j = ws.atom[i].cov14atom[k].i;
            if (j > i) { a14.push_back(i); b14.push_back(j); }
        }
    }
    r.d("xyz", xyz); r.d("charge", q); r.d("radius", rad); r.d("epsilon", eps);
    r.d("gb_radius", gbr); r.d("vj", vj); r.d("terms123", t123); r.d("mass", mass);
    r.i("resid", resid);
    r.i("a12", a12); r.i("b12", b12); r.d("d12", d12);
    r.i("a13", a13); r.i("b13", b13); r.d("d13", d13);
    r.i("a14", a14); r.i("b14", b14);
    const BondOrder &bo = ws.bondorder();
    std::vector<int32_t> mid(1, bo.getMaxIndexDelta());
    r.i("mid", mid);
    const std::vector<char> &band = bo.getDiagData();
    std::vector<int8_t> b8(band.begin(), band.end());
    r.b("band", b8);
    std::vector<int32_t> spi, spj, spo;
    const std::vector<LongBondOrder> &sp = bo.getOffDiagData();
    for (size_t p = 0; p < sp.size(); ++p)
        for (size_t k = 0; k < sp[p].partner.size(); ++k) {
            spi.push_back(sp[p].i); spj.push_back(sp[p].partner[k].j); spo.push_back(sp[p].partner[k].order);
        }
    r.i("sp_i", spi); r.i("sp_j", spj); r.i("sp_o", spo);
}

void dump_nlist(Rec &r, WorkSpace &ws, const std::string &tag) {
    const int n = (int)ws.atom.size();
    const NeighbourData *nb = ws.nlist().getData();
    std::vector<int32_t> rs(1, 0);
    std::vector<uint32_t> ent;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < nb[i].n; ++k) ent.push_back((uint32_t)nb[i].i[k]);
        rs.push_back((int32_t)ent.size());
    }
    r.i(tag + "_row_start", rs);
    r.u(tag + "_entries", ent);
}

void dump_forces(Rec &r, WorkSpace &ws, const std::string &tag) {
    const int n = (int)ws.atom.size();
    std::vector<double> f;
    for (int i = 0; i < n; ++i) {
        f.push_back(ws.cur.atom[i].f.x);
        f.push_back(ws.cur.atom[i].f.y);
        f.push_back(ws.cur.atom[i].f.z);
    }
    r.d(tag, f);
}

void configure(GBProbe &gb, const Case &c) {
    gb.Cutoff = c.cutoff;
    gb.InnerCutoff = c.inner;
    if (c.set_vdw) { gb.VdwCutoff = c.vdw_cutoff; gb.VdwInnerCutoff = c.vdw_inner; }
    gb.DielectricSolvent = c.eps_out;
}

void run_case(FFParamSet &ffps, const Case &c, const std::string &outdir) {
    Rec r(outdir + "/" + c.name + ".bin");
    {
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ffb(ws);
        FF_Bonded bonds(ws);
        ffb.add(bonds);
        ffb.setup(); // fills cov12atom/cov13atom equilibrium distances
        Forcefield ff(ws);
        GBProbe gb(ws);
        configure(gb, c);
        ff.add(gb);
        ff.printEnergySummary(); // ensuresetup, calcNewList (shadows on), calcForces
        dump_inputs(r, ws, gb);
        std::vector<double> params;
        params.push_back(gb.Cutoff); params.push_back(gb.InnerCutoff);
        params.push_back(gb.VdwCutoff); params.push_back(gb.VdwInnerCutoff);
        params.push_back(gb.Vdw14Scaling); params.push_back(gb.Elec14Scaling);
        params.push_back(gb.DielectricSolute); params.push_back(gb.DielectricSolvent);
        params.push_back(gb.GbsaStillCutoff); params.push_back(gb.ExpApproxThreshold);
        params.push_back(gb.DielectricOffset); params.push_back(gb.BornRadiusOffset);
        for (int k = 0; k < 5; ++k) params.push_back(gb.p(k));
        params.push_back(c.padding);
        r.d("params", params);
        dump_nlist(r, ws, "nl_full");
        std::vector<double> born, deda;
        for (int i = 0; i < (int)ws.atom.size(); ++i) { born.push_back(gb.born(i)); deda.push_back(gb.deda(i)); }
        r.d("born", born);
        r.d("deda", deda);
        dump_forces(r, ws, "force");
        r.s("epot", ws.ene.epot); r.s("epot_vdw", ws.ene.epot_vdw); r.s("epot_elec", ws.ene.epot_elec);
        r.s("epot_pol", ws.ene.epot_pol); r.s("epot_pol_self", ws.ene.epot_pol_self);
        r.s("epot_pol_cross", ws.ene.epot_pol_cross);

        // energy path (calcEnergies): Born radii, FF_NonBonded, calcBornEnergy
        ff.calcEnergies();
        dump_forces(r, ws, "en_force");
        r.s("en_epot", ws.ene.epot); r.s("en_epot_vdw", ws.ene.epot_vdw); r.s("en_epot_elec", ws.ene.epot_elec);
        r.s("en_pol", gb.pol()); r.s("en_pol_self", gb.polSelf()); r.s("en_pol_cross", gb.polCross());
        r.s("en_virial", ws.ene.InternalVirial);

        // half list (shadows off), as the reference builds it inside setup
        ws.nlist().calcShadow(false);
        ws.nlist().calcNewList();
        dump_nlist(r, ws, "nl_half");
    }
    {
        // FF_NonBonded alone with its own defaults (force switching)
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ff(ws);
        FF_NonBonded nb(ws);
        ff.add(nb);
        ff.printEnergySummary();
        dump_forces(r, ws, "nb_force");
        r.s("nb_epot", ws.ene.epot); r.s("nb_epot_vdw", ws.ene.epot_vdw); r.s("nb_epot_elec", ws.ene.epot_elec);
        r.s("nb_virial", ws.ene.InternalVirial);
        std::vector<double> np;
        np.push_back(nb.Cutoff); np.push_back(nb.InnerCutoff); np.push_back(nb.VdwCutoff);
        np.push_back(nb.VdwInnerCutoff); np.push_back(nb.Vdw14Scaling); np.push_back(nb.Elec14Scaling);
        np.push_back(nb.Dielectric);
        r.d("nb_params", np);
    }
    {
        // FF_Bonded alone: term lists, then forces at the built geometry and
        // at a perturbed one (every coordinate moved by up to 0.2 A)
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        Forcefield ffb(ws);
        BondedProbe bonds(ws);
        ffb.add(bonds);
        ffb.setup();
        dump_bonded_terms(r, bonds);
        ffb.calcForces();
        dump_bonded_eval(r, ws, bonds, "bf");
        uint64_t x = 0x5EEDULL + c.name.size() * 7919ULL + (uint64_t)c.seq.size();
        std::vector<double> xyz;
        for (int i = 0; i < (int)ws.atom.size(); ++i) {
            ws.cur.atom[i].p.x += 0.2 * splitmix_unit(x);
            ws.cur.atom[i].p.y += 0.2 * splitmix_unit(x);
            ws.cur.atom[i].p.z += 0.2 * splitmix_unit(x);
            xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
        }
        r.d("bfp_xyz", xyz);
        ffb.calcForces();
        dump_bonded_eval(r, ws, bonds, "bfp");
    }
    if (c.sd_steps > 0) {
        // FF_Bonded + FF_GeneralizedBorn_Still in one Forcefield: energies at
        // the built geometry, then steepest descent
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ff(ws);
        BondedProbe bonds(ws);
        GBProbe gb(ws);
        configure(gb, c);
        ff.add(bonds);
        ff.add(gb);
        ff.printEnergySummary();
        dump_forces(r, ws, "ff_force");
        r.s("ff_epot", ws.ene.epot); r.s("ff_bond", ws.ene.epot_bond); r.s("ff_angle", ws.ene.epot_angle);
        r.s("ff_torsion", ws.ene.epot_torsion); r.s("ff_vdw", ws.ene.epot_vdw); r.s("ff_elec", ws.ene.epot_elec);
        r.s("ff_pol", ws.ene.epot_pol);
        Protocol::Minimisation mini(ff);
        mini.Algorithm = Protocol::Minimisation::SteepestDescent;
        mini.Steps = c.sd_steps;
        mini.UpdateScr = 0;
        mini.OutputLevel = Verbosity::Silent;
        int done = mini.run();
        std::vector<double> xyz;
        for (int i = 0; i < (int)ws.atom.size(); ++i) {
            xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
        }
        r.d("sdf_xyz", xyz);
        dump_forces(r, ws, "sdf_force");
        r.s("sdf_epot", ws.ene.epot);
        r.s("sdf_done", (double)done);
    }
    if (c.sd_steps > 0) {
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ffb(ws);
        FF_Bonded bonds(ws);
        ffb.add(bonds);
        ffb.setup();
        Forcefield ff(ws);
        GBProbe gb(ws);
        configure(gb, c);
        ff.add(gb);
        Protocol::Minimisation mini(ff);
        mini.Algorithm = Protocol::Minimisation::SteepestDescent;
        mini.Steps = c.sd_steps;
        mini.UpdateScr = 0;
        mini.OutputLevel = Verbosity::Silent;
        int done = mini.run();
        std::vector<double> xyz;
        for (int i = 0; i < (int)ws.atom.size(); ++i) {
            xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
        }
        r.d("sd_xyz", xyz);
        dump_forces(r, ws, "sd_force");
        r.s("sd_epot", ws.ene.epot);
        r.s("sd_steps", (double)c.sd_steps);
        r.s("sd_done", (double)done);
        r.s("sd_stepsize", mini.StepSize);
    }
    {
        // FF_SoftVDW alone (PD's defaults: Hardness 12, OlapOffset 0.6,
        // VdwCutoff 15) at the built geometry
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ff(ws);
        FF_SoftVDW soft(ws);
        ff.add(soft);
        ff.printEnergySummary();
        ff.calcForces();
        dump_forces(r, ws, "sv_force");
        r.s("sv_epot", ws.ene.epot);
        r.s("sv_vdw", ws.ene.epot_vdw);
    }
    if (c.sd_steps > 0) {
        // FF_Bonded + FF_SoftVDW steepest descent: PD's soft-core clash relief
        System sys(ffps);
        sys.add(NewProtein(ffps, pd_grammar(c.seq)));
        WorkSpace ws(sys);
        ws.nlist().setPadding(c.padding);
        Forcefield ff(ws);
        FF_Bonded bonds(ws);
        FF_SoftVDW soft(ws);
        ff.add(bonds);
        ff.add(soft);
        ff.printEnergySummary();
        const int steps = c.sd_steps < 50 ? c.sd_steps : 50;
        Protocol::Minimisation mini(ff);
        mini.Algorithm = Protocol::Minimisation::SteepestDescent;
        mini.Steps = steps;
        mini.UpdateScr = 0;
        mini.OutputLevel = Verbosity::Silent;
        int done = mini.run();
        std::vector<double> xyz;
        for (int i = 0; i < (int)ws.atom.size(); ++i) {
            xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
        }
        r.d("svsd_xyz", xyz);
        dump_forces(r, ws, "svsd_force");
        r.s("svsd_epot", ws.ene.epot);
        r.s("svsd_vdw", ws.ene.epot_vdw);
        r.s("svsd_steps", (double)steps);
        r.s("svsd_done", (double)done);
    }
    if (c.md_steps > 0) {
        // PD's MolecularDynamics on FF_Bonded + GB, after 100 steepest-descent
        // steps (the built chain can hold clashes), from velocities drawn from
        // a SplitMix64 stream (uniform in +-sqrt(3 kB T / m) per component,
        // T = 300 K). One run per integrator from the same start: md_*
        // VelocityVerlet, lg_* Langevin (PD's defaults: 10/ps friction, 300 K,
        // LangevinOnHydrogens; libc rand() after srand(lg_seed)), bm_* Beeman
        // (the workspace's old forces zeroed first), lx_* Langevin at 100/ps
        // (gamma dt = 0.1: the exponential form of the coefficients), lh_*
        // Langevin with LangevinOnHydrogens = false (hydrogens on plain
        // kicks; their flags as is_h).
        // vb_* / bb_*: VelocityVerlet / Beeman with PD's Berendsen thermostat
        // (300 K, BerendsenTau 100 fs; fewer than UpdateRemoveTotalMomentum
        // = 200 steps, so no momentum removal happens).
        for (int integ = 0; integ < 7; ++integ) {
            System sys(ffps);
            sys.add(NewProtein(ffps, pd_grammar(c.seq)));
            WorkSpace ws(sys);
            ws.nlist().setPadding(c.padding);
            Forcefield ff(ws);
            FF_Bonded bonds(ws);
            GBProbe gb(ws);
            configure(gb, c);
            ff.add(bonds);
            ff.add(gb);
            ff.printEnergySummary();
            {
                Protocol::Minimisation mini(ff);
                mini.Algorithm = Protocol::Minimisation::SteepestDescent;
                mini.Steps = 100;
                mini.UpdateScr = 0;
                mini.OutputLevel = Verbosity::Silent;
                mini.run();
            }
            const int n = (int)ws.atom.size();
            std::vector<double> v0, x0;
            uint64_t z = 0x5EEDULL;
            for (int i = 0; i < n; ++i) {
                const double vmax = sqrt(3.0 * PhysicsConst::kB * 300.0 / ws.atom[i].mass);
                double u[3];
                for (int k = 0; k < 3; ++k) {
                    z += 0x9E3779B97F4A7C15ULL;
                    uint64_t w = z;
                    w = (w ^ (w >> 30)) * 0xBF58476D1CE4E5B9ULL;
                    w = (w ^ (w >> 27)) * 0x94D049BB133111EBULL;
                    w ^= w >> 31;
                    u[k] = (2.0 * ((double)(w >> 11) * (1.0 / 9007199254740992.0)) - 1.0) * vmax;
                }
                ws.cur.atom[i].v.setTo(u[0], u[1], u[2]);
                v0.push_back(u[0]); v0.push_back(u[1]); v0.push_back(u[2]);
                x0.push_back(ws.cur.atom[i].p.x); x0.push_back(ws.cur.atom[i].p.y); x0.push_back(ws.cur.atom[i].p.z);
            }
            Protocol::MolecularDynamics md(ff);
            const char *tag = "md";
            md.Integrator = Protocol::MolecularDynamics::VelocityVerlet;
            const unsigned lg_seed = integ == 3 ? 4321u : integ == 4 ? 999u : 1234u;
            if (integ == 1 || integ == 3 || integ == 4) {
                tag = integ == 1 ? "lg" : integ == 3 ? "lx" : "lh";
                md.Integrator = Protocol::MolecularDynamics::Langevin;
                if (integ == 3) md.FricCoeff = 1e14;
                if (integ == 4) md.LangevinOnHydrogens = false;
            } else if (integ == 2 || integ == 6) {
                tag = integ == 2 ? "bm" : "bb";
                md.Integrator = Protocol::MolecularDynamics::Beeman;
                for (int i = 0; i < n; ++i) ws.old.atom[i].f.setTo(0, 0, 0);
            } else if (integ == 5) {
                tag = "vb";
            }
            if (integ == 5 || integ == 6) md.Thermostat = Protocol::MolecularDynamics::Berendsen;
            md.Steps = c.md_steps;
            md.Timestep = 1e-15;
            md.RandVel = false;
            md.UpdateNList = 10;
            md.UpdateScr = 0;
            md.UpdateTra = 0;
            md.OutputLevel = Verbosity::Silent;
            if (integ == 1 || integ == 3 || integ == 4) srand(lg_seed);
            // the VelocityVerlet run also writes a NAMD trajectory every 10
            // steps through the workspace's trajectory container (md.cpp:247)
            IO::OutTra_NAMD tra_md(outdir + "/" + c.name + "_mdtra", ws);
            if (integ == 0) {
                ws.outtra.add(tra_md);
                md.UpdateTra = 10;
            }
            md.run();
            if (integ == 0) r.b("md_dcd", slurp(outdir + "/" + c.name + "_mdtra.dcd"));
            std::vector<double> xyz, vel;
            for (int i = 0; i < n; ++i) {
                xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
                vel.push_back(ws.cur.atom[i].v.x); vel.push_back(ws.cur.atom[i].v.y); vel.push_back(ws.cur.atom[i].v.z);
            }
            const std::string t(tag);
            if (integ == 0) {
                r.d("md_xyz0", x0);
                r.d("md_v0", v0);
                r.s("md_steps", (double)c.md_steps);
                r.s("md_dt", md.Timestep);
                r.s("md_update_nlist", (double)md.UpdateNList);
            } else {
                // the same start as the VelocityVerlet run (checked by the tests)
                r.d(t + "_xyz0", x0);
            }
            r.d(t + "_xyz", xyz);
            r.d(t + "_v", vel);
            r.s(t + "_epot", ws.ene.epot);
            r.s(t + "_ekin", ws.ene.ekin);
            if (integ == 5 || integ == 6) r.s(t + "_tau", md.BerendsenTau);
            if (integ == 4) {
                std::vector<int8_t> h;
                for (int i = 0; i < n; ++i) h.push_back(ws.atom[i].isHydrogen() ? 1 : 0);
                r.b("is_h", h);
            }
            if (integ == 1 || integ == 3 || integ == 4) {
                r.s(t + "_seed", (double)lg_seed);
                r.s(t + "_temp", 300.0);
                r.s(t + "_fric", md.FricCoeff);
            }
            if (integ == 0) {
                // PD's result hand-off formats on the final VelocityVerlet
                // state: WorkSpace::save() as SnapShot MIME text
                // (workspace/snapshot.cpp:316-348) after a force evaluation,
                // and a NAMD DCD trajectory (fileio/psfdcd.cpp) of three
                // frames: the start, the end, and the end shifted by 0.1 A
                ff.calcForces();
                const std::string stem = outdir + "/" + c.name + "_io";
                SnapShot snap = ws.save();
                snap.writeMIME(stem + ".mime", c.name, "w");
                r.b("io_mime", slurp(stem + ".mime"));
                dump_forces(r, ws, "io_force");
                r.s("io_epot", ws.ene.epot);
                std::vector<double> fx;
                IO::OutTra_NAMD tra(stem, ws);
                for (int frame = 0; frame < 3; ++frame) {
                    for (int i = 0; i < n; ++i) {
                        const double s0 = frame == 2 ? 0.1 : 0.0;
                        if (frame == 0) ws.cur.atom[i].p.setTo(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
                        else ws.cur.atom[i].p.setTo(xyz[3 * i] + s0, xyz[3 * i + 1] - s0, xyz[3 * i + 2] + 2 * s0);
                        fx.push_back(ws.cur.atom[i].p.x); fx.push_back(ws.cur.atom[i].p.y); fx.push_back(ws.cur.atom[i].p.z);
                    }
                    tra.append();
                }
                r.d("io_dcd_xyz", fx);
                r.b("io_dcd", slurp(stem + ".dcd"));
                // the same three frames as a Bristol trajectory with every
                // include (OutTra_BTF, fileio/tra.cpp:28-180), and what its
                // header and records take from the workspace
                {
                    IO::OutTra_BTF btf(stem, ws);
                    btf.create(IO::AllIncludes);
                    for (int frame = 0; frame < 3; ++frame) {
                        for (int i = 0; i < n; ++i) {
                            const double s0 = frame == 2 ? 0.1 : 0.0;
                            if (frame == 0) ws.cur.atom[i].p.setTo(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]);
                            else ws.cur.atom[i].p.setTo(xyz[3 * i] + s0, xyz[3 * i + 1] - s0, xyz[3 * i + 2] + 2 * s0);
                        }
                        btf.append();
                    }
                    r.b("io_btf", slurp(stem + ".tra"));
                    std::vector<int8_t> names;
                    std::vector<int32_t> cov, ncov, known, bb;
                    std::vector<double> ref;
                    for (int i = 0; i < n; ++i) {
                        const Particle &a = ws.atom[i];
                        const std::string f[4] = {a.pdbname, a.type_name, a.rawname, a.parentname};
                        for (int q = 0; q < 4; ++q)
                            for (int k = 0; k < 16; ++k) names.push_back(k < (int)f[q].size() ? f[q][k] : 0);
                        ncov.push_back((int32_t)a.cov12atom.size());
                        for (int k = 0; k < 6; ++k) cov.push_back(k < (int)a.cov12atom.size() ? a.cov12atom[k].i : -1);
                        known.push_back(a.isKnownStructure() ? 1 : 0);
                        ref.push_back(a.posRef().x); ref.push_back(a.posRef().y); ref.push_back(a.posRef().z);
                    }
                    for (size_t k = 0; k < ws.res.size(); ++k) {
                        bb.push_back(ws.res[k].iN); bb.push_back(ws.res[k].iCA); bb.push_back(ws.res[k].iC);
                    }
                    r.b("io_names", names);
                    r.i("io_cov12", cov); r.i("io_ncov12", ncov); r.i("io_known", known); r.d("io_ref", ref);
                    r.i("io_backbone", bb);
                    std::vector<double> st;
                    st.push_back((double)ws.Step); st.push_back(ws.ene.cRMS); st.push_back(ws.ene.dRMS);
                    st.push_back(ws.ene.ekin); st.push_back(ws.ene.etot);
                    st.push_back(ws.ene.data1); st.push_back(ws.ene.data2); st.push_back(ws.ene.data3);
                    st.push_back(ws.ene.data4); st.push_back(ws.ene.data5); st.push_back(ws.ene.data6);
                    st.push_back(ws.ene.data7); st.push_back(ws.ene.data8); st.push_back(ws.ene.data9);
                    st.push_back(ws.ene.data10);
                    r.d("io_state", st);
                }
            }
        }
    }
}

} // namespace

int main(int argc, char **argv) {
    if (argc < 3) { fprintf(stderr, "usage: ref_dump <param.ff> <outdir>\n"); return 2; }
    FFParamSet ffps;
    ffps.readLib(argv[1]);
    std::vector<Case> cases;
    // name, seq, cutoff, inner, vdwcut, vdwinner, padding, eps_out, set_vdw, sd_steps, md_steps
    Case c1 = {"tri_ala", "AAA", 12, 9, 12, 9, 2.0, 78.5, true, 20, 40}; cases.push_back(c1);
    Case c2 = {"gly_ala", "GA", 12, 9, 12, 9, 2.0, 78.5, true, 0}; cases.push_back(c2);
    for (int seed = 1; seed <= 3; ++seed) {
        Case c = {"c1_seed" + std::string(1, char('0' + seed)), random_sequence(seed, 20),
                  12, 9, 12, 9, 2.0, 78.5, true, seed == 1 ? 200 : 30, seed == 2 ? 60 : 0};
        cases.push_back(c);
    }
    Case c60 = {"res60_seed4", random_sequence(4, 60), 12, 9, 12, 9, 2.0, 78.5, true, 0}; cases.push_back(c60);
    Case cdef = {"pd_defaults_seed5", random_sequence(5, 12), 15, 12, 15, 12, 0.5, 80.0, false, 0}; cases.push_back(cdef);
    Case cq6 = {"vdw_default_seed6", random_sequence(6, 12), 12, 9, 0, 0, 0.5, 78.5, false, 0}; cases.push_back(cq6);
    Case c10 = {"cut10_seed7", random_sequence(7, 30), 10, 7, 10, 7, 2.0, 78.5, true, 0}; cases.push_back(c10);
    for (size_t k = 0; k < cases.size(); ++k) {
        fprintf(stderr, "case %s seq %s\n", cases[k].name.c_str(), cases[k].seq.c_str());
        run_case(ffps, cases[k], argv[2]);
    }
    return 0;
}
