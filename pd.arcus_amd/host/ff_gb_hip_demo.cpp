// ff_gb_hip_demo.cpp -- drives FF_GeneralizedBorn_Still_HIP the way PD's
// protocols drive a forcefield component: setup, list refresh, calcForces,
// calcEnergies. Input and output are flat little-endian binary files:
//   in : int32 n, n12, n13, n14; int32 precision; double knobs[9]
//        (Cutoff, InnerCutoff, VdwCutoff, VdwInnerCutoff, Vdw14, Elec14,
//         DielectricSolvent, Padding, ExpApproxThreshold);
//        per atom: double x,y,z, charge, radius, epsilon, gb_radius, mass; int32 ir;
//        12 pairs (int32 i, int32 j, double d); 13 pairs (i, j, d); 14 pairs (i, j)
//        optionally FF_Bonded term lists: int32 nb, na, nt, ni; nb x (int32 i, j;
//        double l, k); na x (int32 i, a, j; double theta0, k); nt then ni x
//        (int32 i, a, b, j, terms; double vn[4], n[4], gamma[4])
//   out: double epot, vdw, elec, pol, pol_self, pol_cross; forces[3n];
//        then the same for calcEnergies (plus InternalVirial); with term
//        lists, then FF_Bonded + GB in one forcefield: double epot, bond,
//        angle, torsion, vdw, elec, pol, InternalVirial; forces[3n]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ff_gb_hip.h"

template <typename T> static T rd(FILE *f) {
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) { fprintf(stderr, "short input\n"); exit(2); }
    return v;
}

int main(int argc, char **argv) {
    if (argc < 3) { fprintf(stderr, "usage: ff_gb_hip_demo in.bin out.bin\n"); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    const int n = rd<int32_t>(f), n12 = rd<int32_t>(f), n13 = rd<int32_t>(f), n14 = rd<int32_t>(f);
    const int precision = rd<int32_t>(f);
    double k[9];
    for (double &x : k) x = rd<double>(f);
    pdarcus::WorkSpaceView ws;
    ws.pos.resize(3 * n);
    ws.atom.resize(n);
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) ws.pos[3 * i + c] = rd<double>(f);
        auto &a = ws.atom[i];
        a.charge = rd<double>(f); a.radius = rd<double>(f); a.epsilon = rd<double>(f);
        a.gb_radius = rd<double>(f); a.mass = rd<double>(f); a.ir = rd<int32_t>(f);
    }
    for (int q = 0; q < n12; ++q) { int a = rd<int32_t>(f), b = rd<int32_t>(f); ws.cov12.push_back({a, b}); ws.d12.push_back(rd<double>(f)); }
    for (int q = 0; q < n13; ++q) { int a = rd<int32_t>(f), b = rd<int32_t>(f); ws.cov13.push_back({a, b}); ws.d13.push_back(rd<double>(f)); }
    for (int q = 0; q < n14; ++q) { int a = rd<int32_t>(f), b = rd<int32_t>(f); ws.cov14.push_back({a, b}); }
    pdarcus::BondedTerms bt;
    int32_t cnt[4];
    const bool bonded = fread(cnt, 4, 4, f) == 4;
    if (bonded) {
        for (int q = 0; q < cnt[0]; ++q) {
            for (int c = 0; c < 2; ++c) bt.bond_ij.push_back(rd<int32_t>(f));
            for (int c = 0; c < 2; ++c) bt.bond_lk.push_back(rd<double>(f));
        }
        for (int q = 0; q < cnt[1]; ++q) {
            for (int c = 0; c < 3; ++c) bt.angle_iaj.push_back(rd<int32_t>(f));
            for (int c = 0; c < 2; ++c) bt.angle_tk.push_back(rd<double>(f));
        }
        for (int kind = 0; kind < 2; ++kind) {
            auto &idx = kind ? bt.improper_iabj : bt.torsion_iabj;
            auto &terms = kind ? bt.improper_terms : bt.torsion_terms;
            auto &vn = kind ? bt.improper_vn : bt.torsion_vn;
            auto &nn = kind ? bt.improper_n : bt.torsion_n;
            auto &gm = kind ? bt.improper_gamma : bt.torsion_gamma;
            for (int q = 0; q < cnt[2 + kind]; ++q) {
                for (int c = 0; c < 4; ++c) idx.push_back(rd<int32_t>(f));
                terms.push_back(rd<int32_t>(f));
                for (int c = 0; c < 4; ++c) vn.push_back(rd<double>(f));
                for (int c = 0; c < 4; ++c) nn.push_back(rd<double>(f));
                for (int c = 0; c < 4; ++c) gm.push_back(rd<double>(f));
            }
        }
    }
    fclose(f);

    pdarcus::FF_GeneralizedBorn_Still_HIP gb(ws, precision);
    gb.Cutoff = k[0]; gb.InnerCutoff = k[1]; gb.VdwCutoff = k[2]; gb.VdwInnerCutoff = k[3];
    gb.setScaling14(k[4], k[5]);
    gb.DielectricSolvent = k[6]; gb.Padding = k[7]; gb.ExpApproxThreshold = k[8];
    FILE *o = fopen(argv[2], "wb");
    try {
        gb.ensuresetup();
        gb.calcNewList();          // Forcefield::printEnergySummary refreshes the list first
        ws.zeroForces();           // Forcefield::calcForces zeroes, then calls components
        gb.calcForces();
        double e1[6] = {ws.ene.epot, ws.ene.epot_vdw, ws.ene.epot_elec, ws.ene.epot_pol, ws.ene.epot_pol_self,
                        ws.ene.epot_pol_cross};
        fwrite(e1, 8, 6, o);
        fwrite(ws.force.data(), 8, ws.force.size(), o);
        ws.zeroForces();
        gb.calcEnergies();
        double e2[7] = {ws.ene.epot, ws.ene.epot_vdw, ws.ene.epot_elec, gb.epot_pol, gb.epot_pol_self,
                        gb.epot_pol_cross, ws.ene.InternalVirial};
        fwrite(e2, 8, 7, o);
        fwrite(ws.force.data(), 8, ws.force.size(), o);
        if (bonded) {
            // Forcefield with FF_Bonded added first, then GB (forcefield.cpp:220-227)
            pdarcus::FF_Bonded_HIP bo(ws, bt, precision);
            bo.ensuresetup();
            gb.calcNewList();
            ws.zeroForces();
            bo.calcForces();
            gb.calcForces();
            double e3[8] = {ws.ene.epot, ws.ene.epot_bond, ws.ene.epot_angle, ws.ene.epot_torsion,
                            ws.ene.epot_vdw, ws.ene.epot_elec, ws.ene.epot_pol, ws.ene.InternalVirial};
            fwrite(e3, 8, 8, o);
            fwrite(ws.force.data(), 8, ws.force.size(), o);
            bo.infoLineHeader();
        }
        gb.infoLineHeader();
        printf("\n");
        gb.infoLine();
        printf("\n");
    } catch (const std::exception &ex) {
        fprintf(stderr, "PD Exception: %s\n", ex.what());
        fclose(o);
        return 1;
    }
    fclose(o);
    return 0;
}
