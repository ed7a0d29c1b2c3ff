// pdgpu_setup.cpp -- host-side setup work of the component: option
// defaults, PD's BondOrder band (A2) and the Still fixed Born terms (A3).
// Both run once per molecule and are O(bonds); they live on the host.
#include "../../include/pdgpu.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

extern "C" void pdgpu_options_default(pdgpu_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = PDGPU_FP64;
    o->device = 0;
    o->fast_mode = 1;
    // FF_NonBonded::settodefault (forcefields/nonbonded.cpp:968-1004) under
    // the GB constructor's overrides (forcefields/gbff.cpp:17-37)
    o->cutoff = 15.0;
    o->inner_cutoff = 12.0;
    o->vdw_cutoff = 15.0;
    o->vdw_inner_cutoff = 12.0;
    o->vdw14_scaling = 1.0;
    o->elec14_scaling = 1.0;
    o->dielectric_solute = 1.0;
    o->dielectric_solvent = 80.0;
    o->gbsa_still_cutoff = 10.0;
    o->exp_approx_threshold = -4.0;
    o->dielectric_offset = -0.090;
    o->born_radius_offset = 0.0;
    o->p1 = 0.073;
    o->p2 = 0.921;
    o->p3 = 6.211;
    o->p4 = 15.236;
    o->p5 = 1.254;
    o->nlist_padding = 0.5; // NeighbourListBase default (workspace/neighbourlist.cpp:18)
    o->elec_mode = 3;       // EnergySwitch
    o->vdw_mode = 2;
    o->device_memory_budget = 0;
    o->max_instances = 0;
    o->tiles_per_wave = 0;
    o->components = PDGPU_FF_ALL;
    o->langevin_on_hydrogens = 1; // MolecularDynamics::settodefault (protocols/md.cpp:59)
    o->use_graphs = 0; // measured no faster: graph replay leaves ~10 us between kernel nodes
    // FF_SoftVDW::settodefault (forcefields/ffsoftvdw.h:39-47)
    o->softvdw_hardness = 12.0;
    o->softvdw_olap_offset = 0.6;
    o->softvdw_cutoff = 15.0;
}

extern "C" void pdgpu_options_default_nonbonded(pdgpu_options *o) {
    if (!o) return;
    pdgpu_options_default(o);
    // FF_NonBonded::settodefault (forcefields/nonbonded.cpp:968-1004):
    // ForceSwitch = true, EnergySwitch = false
    o->elec_mode = 4;
    o->vdw_mode = 2;
    o->components = PDGPU_FF_NONBONDED | PDGPU_FF_BONDED;
}

// BondOrder::reinit (workspace/bondorder.cpp:77-195). The band holds the
// lowest topological order per pair for index gaps up to max_index_delta;
// wider pairs are returned in the sparse arrays (i<j), each pair once.
extern "C" int pdgpu_bondorder_build(int32_t n, const int32_t *resid,
                                     int32_t n12, const int32_t *a12, const int32_t *b12,
                                     int32_t n13, const int32_t *a13, const int32_t *b13,
                                     int32_t n14, const int32_t *a14, const int32_t *b14,
                                     int32_t *mid_out, int8_t *band, int32_t *nsparse,
                                     int32_t *si, int32_t *sj, int32_t *so) {
    if (n <= 0 || !resid || !mid_out) return PDGPU_ERR_INVALID;
    const int32_t *A[3] = {a12, a13, a14};
    const int32_t *B[3] = {b12, b13, b14};
    const int32_t cnt[3] = {n12, n13, n14};
    int mid = 0;
    for (int l = 0; l < 3; ++l) {
        if (cnt[l] > 0 && (!A[l] || !B[l])) return PDGPU_ERR_INVALID;
        for (int k = 0; k < cnt[l]; ++k) {
            const int a = A[l][k], b = B[l][k];
            if (a < 0 || b < 0 || a >= n || b >= n) return PDGPU_ERR_INVALID;
            // only partners within neighbouring residues widen the band
            if (std::abs(resid[a] - resid[b]) > 1) continue;
            mid = std::max(mid, std::abs(a - b));
        }
    }
    mid += 1;
    *mid_out = mid;
    if (!band) return PDGPU_OK;
    if (n > 1) std::memset(band, 7, (size_t)mid * (size_t)(n - 1));
    std::map<std::pair<int, int>, int> sparse; // pair -> lowest order
    std::vector<std::pair<int, int>> order_seen;
    for (int l = 0; l < 3; ++l) {
        const int order = l + 1;
        for (int k = 0; k < cnt[l]; ++k) {
            int i = A[l][k], j = B[l][k];
            if (i == j) continue;
            if (i > j) std::swap(i, j);
            const int d = j - i;
            if (d > mid) {
                auto key = std::make_pair(i, j);
                auto it = sparse.find(key);
                if (it == sparse.end()) {
                    sparse[key] = order;
                    order_seen.push_back(key);
                } else if (order < it->second) {
                    it->second = order;
                }
            } else {
                int8_t &slot = band[(size_t)i * mid + (d - 1)];
                if (order < slot) slot = (int8_t)order;
            }
        }
    }
    if (nsparse) *nsparse = (int32_t)order_seen.size();
    if (si && sj && so) {
        for (size_t k = 0; k < order_seen.size(); ++k) {
            si[k] = order_seen[k].first;
            sj[k] = order_seen[k].second;
            so[k] = sparse[order_seen[k]];
        }
    }
    return PDGPU_OK;
}

// calcFixedBornRadiiTerms (forcefields/gbff.cpp:307-381).
extern "C" int pdgpu_fixed_terms(int32_t n, const double *gbr,
                                 int32_t n12, const int32_t *a12, const int32_t *b12, const double *d12,
                                 int32_t n13, const int32_t *a13, const int32_t *b13, const double *d13,
                                 const pdgpu_options *o, double *vj, double *t123) {
    if (n <= 0 || !gbr || !o || !vj || !t123) return PDGPU_ERR_INVALID;
    const double pi = 3.14159265358979323846;
    // hydrogen rule of Qiu et al.: type radius <= 0.01 becomes 1.15
    auto radj = [](double r) { return r <= 0.01 ? 1.15 : r; };
    for (int j = 0; j < n; ++j) {
        const double R = radj(gbr[j]);
        vj[j] = 4.0 * pi * R * R * R / 3;
    }
    for (int k = 0; k < n12; ++k) {
        const int a = a12[k], b = b12[k];
        if (a < 0 || b < 0 || a >= n || b >= n) return PDGPU_ERR_INVALID;
        for (int side = 0; side < 2; ++side) {
            const int j = side ? b : a, o2 = side ? a : b;
            const double Rj = radj(gbr[j]), Rk = radj(gbr[o2]);
            const double r = 1.01 * d12[k]; // overlap at 101% of the bond length
            const double h = Rj * (1.0 + (Rk * Rk - Rj * Rj - r * r) / (2.0 * Rj * r));
            vj[j] -= pi * h * h * (3 * Rj - h) / 3;
        }
    }
    for (int j = 0; j < n; ++j)
        if (vj[j] < 0) vj[j] = 0;
    std::vector<double> s12(n, 0.0), s13(n, 0.0);
    for (int k = 0; k < n12; ++k) {
        const double r4 = d12[k] * d12[k] * d12[k] * d12[k];
        s12[a12[k]] += o->p2 * vj[b12[k]] / r4;
        s12[b12[k]] += o->p2 * vj[a12[k]] / r4;
    }
    for (int k = 0; k < n13; ++k) {
        const int a = a13[k], b = b13[k];
        if (a < 0 || b < 0 || a >= n || b >= n) return PDGPU_ERR_INVALID;
        const double r4 = d13[k] * d13[k] * d13[k] * d13[k];
        s13[a] += o->p3 * vj[b] / r4;
        s13[b] += o->p3 * vj[a] / r4;
    }
    for (int i = 0; i < n; ++i) {
        const double R = radj(gbr[i]);
        t123[i] = -166.0 / (R + o->dielectric_offset + o->p1) + s12[i] + s13[i];
    }
    return PDGPU_OK;
}
