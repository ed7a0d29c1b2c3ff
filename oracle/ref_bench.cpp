// ref_bench.cpp -- times the compiled reference's own CPU path on the batch
// workload's molecules. Test infrastructure: runs only where the reference
// tree and its parameter file exist.
//
// Usage: ref_bench <param.ff> <outdir|-> <first> <count> <budget_s> [lengths]
// Molecule k (first <= k < first+count) is a peptide with the random
// sequence of seed 1000 + k (SplitMix64 over the 20 standard residues, as
// the golden cases) and the length on line k of the lengths file (default
// 10 + (k*37 mod 141) residues), built extended by PD's NewProtein. For each: FF_Bonded setup (the equilibrium 1-2/1-3
// distances GB's fixed terms read), then FF_GeneralizedBorn_Still with the
// batch knobs (Cutoff 10, InnerCutoff 7, VdwCutoff 10, VdwInnerCutoff 7,
// eps_out 78.5, padding 2). Timed per molecule: NeighbourList::calcNewList
// + Forcefield::calcForces (Born radii, fused pair pass, chain rule). Stops
// after budget_s seconds of timed work. Prints one line per molecule
// (index, atoms, unordered pairs within Cutoff, seconds, epot) and, with an
// outdir, writes the molecule's inputs as <outdir>/bench_<k>.bin for the
// oracle port to be timed on the same molecules.
#include "global.h"
#include "system/system.h"
#include "system/genpolymer.h"
#include "workspace/workspace.h"
#include "workspace/neighbourlist.h"
#include "workspace/bondorder.h"
#include "forcefields/ffparam.h"
#include "forcefields/forcefield.h"
#include "forcefields/ffbonded.h"
#include "forcefields/gbff.h"

#include <stdint.h>
#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace Physics;

namespace {

double now() {
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

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
    void b(const std::string &n, const std::vector<int8_t> &v) { put(n, 'b', v.empty() ? 0 : &v[0], v.size(), 1); }
};

class GBProbe : public FF_GeneralizedBorn_Still {
public:
    explicit GBProbe(WorkSpace &ws) : FF_GeneralizedBorn_Still(ws) {}
    double typeRadius(int fftype) const { return GBtype[fftype].radius; }
    double vj(int i) const { return GB_atom_param_still[i].Vj; }
    double terms123(int i) const { return GB_atom_param_still[i].terms123; }
};

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

std::string grammar(const std::string &seq) {
    std::string mid;
    for (size_t k = 1; k + 1 < seq.size(); ++k) mid += seq[k];
    return "*" + seq.substr(0, 1) + "-(" + mid + ")-" + seq.substr(seq.size() - 1) + "*";
}

long long pairs_within(WorkSpace &ws, double cutoff) {
    const NeighbourData *nb = ws.nlist().getData();
    long long n = 0;
    for (int i = 0; i < (int)ws.atom.size(); ++i)
        for (int k = 0; k < nb[i].n; ++k) {
            const int j = nb[i].i[k] & 0xFFFFFF;
            if (j <= i) continue;
            const double dx = ws.cur.atom[j].p.x - ws.cur.atom[i].p.x;
            const double dy = ws.cur.atom[j].p.y - ws.cur.atom[i].p.y;
            const double dz = ws.cur.atom[j].p.z - ws.cur.atom[i].p.z;
            if (dx * dx + dy * dy + dz * dz <= cutoff * cutoff) ++n;
        }
    return n;
}

void dump(const std::string &path, WorkSpace &ws, GBProbe &gb) {
    Rec r(path);
    const int n = (int)ws.atom.size();
    std::vector<double> xyz, q, rad, eps, gbr, vj, t123, mass;
    for (int i = 0; i < n; ++i) {
        xyz.push_back(ws.cur.atom[i].p.x); xyz.push_back(ws.cur.atom[i].p.y); xyz.push_back(ws.cur.atom[i].p.z);
        q.push_back(ws.atom[i].charge); rad.push_back(ws.atom[i].radius); eps.push_back(ws.atom[i].epsilon);
        gbr.push_back(gb.typeRadius(ws.atom[i].FFType)); vj.push_back(gb.vj(i)); t123.push_back(gb.terms123(i));
        mass.push_back(ws.atom[i].mass);
    }
    r.d("xyz", xyz); r.d("charge", q); r.d("radius", rad); r.d("epsilon", eps);
    r.d("gb_radius", gbr); r.d("vj", vj); r.d("terms123", t123); r.d("mass", mass);
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

} // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: ref_bench <param.ff> <outdir|-> <first> <count> <budget_s> [lengths]\n");
        return 2;
    }
    FFParamSet ffps;
    ffps.readLib(argv[1]);
    const std::string outdir = argv[2];
    const int first = atoi(argv[3]), count = atoi(argv[4]);
    const double budget = atof(argv[5]);
    std::vector<int> lengths;
    if (argc > 6) {
        FILE *lf = fopen(argv[6], "r");
        if (!lf) { fprintf(stderr, "cannot read %s\n", argv[6]); return 2; }
        int v;
        while (fscanf(lf, "%d", &v) == 1) lengths.push_back(v);
        fclose(lf);
    }
    double busy = 0;
    for (int k = first; k < first + count && busy < budget; ++k) {
        const int len = k < (int)lengths.size() ? lengths[k] : 10 + (k * 37) % 141;
        System sys(ffps);
        sys.add(NewProtein(ffps, grammar(random_sequence(1000 + k, len))));
        WorkSpace ws(sys);
        ws.nlist().setPadding(2.0);
        Forcefield ffb(ws);
        FF_Bonded bonds(ws);
        ffb.add(bonds);
        ffb.setup();
        Forcefield ff(ws);
        GBProbe gb(ws);
        gb.Cutoff = 10; gb.InnerCutoff = 7; gb.VdwCutoff = 10; gb.VdwInnerCutoff = 7;
        gb.DielectricSolvent = 78.5;
        ff.add(gb);
        ff.printEnergySummary(); // setup (shadow list on), first list and evaluation, untimed
        const double t0 = now();
        ws.nlist().calcNewList();
        ff.calcForces();
        const double dt = now() - t0;
        busy += dt;
        printf("%d %d %lld %.9f %.12e\n", k, (int)ws.atom.size(), pairs_within(ws, gb.Cutoff), dt, ws.ene.epot);
        fflush(stdout);
        if (outdir != "-") {
            char name[64];
            snprintf(name, sizeof(name), "/bench_%d.bin", k);
            dump(outdir + name, ws, gb);
        }
    }
    return 0;
}
