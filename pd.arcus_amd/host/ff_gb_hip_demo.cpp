// ff_gb_hip_demo.cpp -- drives FF_GeneralizedBorn_Still_HIP the way PD's
// protocols drive a forcefield component: setup, list refresh, calcForces,
// calcEnergies. Input and output are flat little-endian binary files:
//   in : int32 n, n12, n13, n14; int32 precision; double knobs[9]
//        (Cutoff, InnerCutoff, VdwCutoff, VdwInnerCutoff, Vdw14, Elec14,
//         DielectricSolvent, Padding, ExpApproxThreshold);
//        per atom: double x,y,z, charge, radius, epsilon, gb_radius, mass; int32 ir;
//        12 pairs (int32 i, int32 j, double d); 13 pairs (i, j, d); 14 pairs (i, j)
//   out: double epot, vdw, elec, pol, pol_self, pol_cross; forces[3n];
//        then the same for calcEnergies (plus InternalVirial)
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
