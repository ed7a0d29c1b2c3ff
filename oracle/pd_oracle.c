/*
 * pd_oracle.c -- CPU FP64 restatement of PD's per-step forcefield hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Written from the reference's behaviour (JonnyWideFoot/pd.arcus, read as
 * text). Each function names the reference lines it restates. How each
 * function is pinned (tests/test_oracle_*.py):
 *   orc_bondorder        golden vectors from the compiled reference (band + sparse,
 *                        amber99 peptides) and hand-built chains/rings
 *   orc_nlist_build      golden vectors from the compiled reference (exact packed
 *                        entries) and brute force on tiny inputs
 *   orc_fixed_terms      golden V_j/terms123 from the compiled reference; closed
 *                        form for an isolated atom and a diatomic
 *   orc_born_radii       golden Born radii from the compiled reference; P4=0
 *                        closed form; two-atom closed form
 *   orc_gb_calc_forces   golden energies/forces from the compiled reference;
 *                        two-atom closed forms; Newton's third law; finite
 *                        differences with the chain-rule constant set to halfeconv
 *   orc_nb_calc_forces   golden FF_NonBonded energies/forces from the compiled
 *                        reference; finite differences (energy switch)
 *   orc_gb_calc_energies golden calcEnergies totals from the compiled reference
 *   orc_minimise_sd      golden per-step energies of the compiled reference's
 *                        Minimisation (SteepestDescent)
 *   orc_md_vv            energy conservation of a two-atom system (textbook VV is
 *                        a deliberate choice; PD's own VV update differs, see
 *                        protocols/md.cpp:788-838)
 *   orc_md_vv_pd         PD's own MolecularDynamics runs with the VelocityVerlet
 *                        integrator (compiled reference, tests/golden md_*)
 *   orc_langevin_coef    exp-form and Taylor-form coefficients agree where both
 *                        hold; gamma*dt -> 0 limits (c0, c1, c2 -> 1, 1, 1/2)
 *   orc_langevin_noise   PD's nrand over libc rand() in the reference's call
 *                        order; the Langevin golden run uses the same stream
 *   orc_md_langevin_pd   PD's own MolecularDynamics runs with the Langevin
 *                        integrator (compiled reference, tests/golden lg_*)
 *   orc_md_beeman_pd     PD's own MolecularDynamics runs with the Beeman
 *                        integrator (compiled reference, tests/golden bm_*)
 */
#include "pd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants: forcefields/forcefield.cpp:18-38, maths/maths.cpp:6-8 ---- */
static const double kKcal2J = 4184.0;
static const double kNa = 6.0221418E23;
static const double kAngstrom = 1E-10;
static const double kECharge = 1.60217645E-19;
static const double kE0 = 8.854187817E-12;
static const double kPI = 3.14159265358979323846;
static const double kKB = 1.38065E-23;
static const double kAmu = 1.66053878E-27;

static double econv_joule(void) {
    const double four_pi_e0 = (kPI * 4.0) * kE0;
    return (kECharge * kECharge) / (four_pi_e0 * kAngstrom);
}
static double econv_kcal(void) {
    const double four_pi_e0 = (kPI * 4.0) * kE0;
    return ((1.0 / kKcal2J) * kNa * (kECharge * kECharge)) / (four_pi_e0 * kAngstrom);
}

double orc_const(int which) {
    switch (which) {
    case ORC_KCAL2J: return kKcal2J;
    case ORC_NA: return kNa;
    case ORC_ANGSTROM: return kAngstrom;
    case ORC_ECONV_J: return econv_joule();
    case ORC_HALFECONV_J: return econv_joule() * 0.5;
    case ORC_ECONV: return econv_kcal();
    case ORC_HALFECONV: return econv_kcal() * 0.5;
    case ORC_KB: return kKB;
    case ORC_AMU: return kAmu;
    }
    return NAN;
}

static inline double sqr(double x) { return x * x; }

/* Defaults of a freshly constructed FF_GeneralizedBorn_Still
 * (forcefields/gbff.cpp:17-37 over forcefields/nonbonded.cpp:968-1004).
 * 1-4 scalings are 1.0 here; PD's AutoScaling14 copies them from the
 * parameter file at setup (forcefields/nonbonded.cpp:1014-1018). */
void orc_default_params(orc_params *p) {
    memset(p, 0, sizeof(*p));
    p->cutoff = 15.0;
    p->inner_cutoff = 12.0;
    p->vdw_cutoff = 15.0;
    p->vdw_inner_cutoff = 12.0;
    p->vdw14_scaling = 1.0;
    p->elec14_scaling = 1.0;
    p->dielectric_solute = 1.0;
    p->dielectric_solvent = 80.0;
    p->still_cutoff = 10.0;
    p->exp_approx_threshold = -4.0;
    p->dielectric_offset = -0.090;
    p->born_radius_offset = 0.0;
    p->P1 = 0.073; p->P2 = 0.921; p->P3 = 6.211; p->P4 = 15.236; p->P5 = 1.254;
    p->padding = 0.5;
    p->elec_mode = 3; /* GB constructor: EnergySwitch=true, ForceSwitch=false */
    p->vdw_mode = 2;
}

/* ---------------- A2: workspace/bondorder.cpp:77-195 ---------------- */
int orc_bondorder(int n, const int32_t *resid,
                  int n12, const int32_t *a12, const int32_t *b12,
                  int n13, const int32_t *a13, const int32_t *b13,
                  int n14, const int32_t *a14, const int32_t *b14,
                  int32_t *mid_out, int8_t *band,
                  int32_t *sp_i, int32_t *sp_j, int32_t *sp_o) {
    /* calcMaxIndexDelta: widest index gap among 1-2/1-3/1-4 partners whose
     * residues differ by at most one, plus one (bondorder.cpp:77-127). */
    int mid = 0;
    const int32_t *A[3] = {a12, a13, a14}, *B[3] = {b12, b13, b14};
    const int cnt[3] = {n12, n13, n14};
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < cnt[l]; ++k) {
            int a = A[l][k], b = B[l][k];
            if (abs(resid[a] - resid[b]) > 1) continue;
            int d = abs(a - b);
            if (d > mid) mid = d;
        }
    mid += 1;
    *mid_out = mid;
    if (!band) return 0;
    /* reinit: band initialised to MaxBondOrder 7, then 1-4, 1-3, 1-2 written
     * in that order so the lowest order wins (bondorder.cpp:170-195). Pairs
     * with index gap > mid go to the sparse list, where the reference's
     * later patches (correctOffDiagonalBondOrders) also let the lowest order
     * win; here each sparse pair is stored once with that final order. */
    if (n > 1) memset(band, 7, (size_t)mid * (size_t)(n - 1));
    int nsp = 0;
    for (int l = 2; l >= 0; --l) {
        const int order = l + 1;
        for (int k = 0; k < cnt[l]; ++k) {
            int i = A[l][k], j = B[l][k];
            if (i == j) continue;
            if (i > j) { int t = i; i = j; j = t; }
            int d = j - i;
            if (d > mid) {
                int found = -1;
                for (int q = 0; q < nsp; ++q)
                    if (sp_i[q] == i && sp_j[q] == j) { found = q; break; }
                if (found < 0) { sp_i[nsp] = i; sp_j[nsp] = j; sp_o[nsp] = order; ++nsp; }
                else sp_o[found] = order;
            } else {
                band[(size_t)i * mid + (d - 1)] = (int8_t)order;
            }
        }
    }
    return nsp;
}

/* Bond order as recorded by calcNewList_InfiniteSpace + the off-diagonal
 * patch (workspace/neighbourlist.cpp:281-289, 162-203): band lookup only for
 * |i-j| < mid, the sparse list otherwise, 7 when absent. */
int orc_pair_bondorder(const orc_system *s, int i, int j) {
    int d = abs(j - i);
    if (d < s->mid) {
        if (d == 0) return 0;
        int lo = i < j ? i : j;
        return s->band[(size_t)lo * s->mid + (d - 1)] & 7;
    }
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int bo = 7;
    for (int q = 0; q < s->nsparse; ++q)
        if (s->sp_i[q] == lo && s->sp_j[q] == hi) bo = s->sp_o[q] & 7;
    return bo;
}

/* ---------------- A1: workspace/neighbourlist.cpp:205-316 ---------------- */
int64_t orc_nlist_build(const orc_system *s, const double *xyz, double list_cutoff,
                        double padding, int shadow, orc_nlist *nl) {
    const int n = s->n;
    const double lim = sqr(list_cutoff + padding);
    int64_t cnt = 0;
    for (int i = 0; i < n; ++i) {
        nl->row_start[i] = (int32_t)cnt;
        const int limit = shadow ? n : i;
        for (int j = 0; j < limit; ++j) {
            double dx = xyz[3 * j] - xyz[3 * i];
            double dy = xyz[3 * j + 1] - xyz[3 * i + 1];
            double dz = xyz[3 * j + 2] - xyz[3 * i + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > lim) continue;
            if (j == i) continue;
            int bo = orc_pair_bondorder(s, i, j);
            if (cnt < nl->cap)
                nl->entries[cnt] = ((uint32_t)j & 0x00FFFFFFu) | ((uint32_t)(bo & 7) << 24);
            ++cnt;
        }
    }
    nl->row_start[n] = (int32_t)cnt;
    nl->count = cnt;
    return cnt > nl->cap ? cnt : 0;
}

static inline int nl_j(uint32_t e) { return (int)(e & 0x00FFFFFFu); }
static inline int nl_bo(uint32_t e) { return (int)((e >> 24) & 7u); }

/* ---------------- A3: forcefields/gbff.cpp:307-381 ---------------- */
void orc_fixed_terms(int n, const double *gb_radius,
                     int n12, const int32_t *a12, const int32_t *b12, const double *d12,
                     int n13, const int32_t *a13, const int32_t *b13, const double *d13,
                     const orc_params *p, double *vj, double *terms123) {
    /* hydrogen rule: a type radius <= 0.01 is replaced by 1.15 (gbff.cpp:326-327) */
#define ORC_RADJ(r) ((r) <= 0.01 ? 1.15 : (r))
    const double four_pi = 4.0 * kPI;
    for (int j = 0; j < n; ++j) {
        double Rj = ORC_RADJ(gb_radius[j]);
        vj[j] = four_pi * Rj * Rj * Rj / 3;
    }
    /* subtract half-overlap caps of each 1-2 partner, both directions */
    for (int k = 0; k < n12; ++k) {
        for (int side = 0; side < 2; ++side) {
            int j = side ? b12[k] : a12[k];
            int o = side ? a12[k] : b12[k];
            double Rj = ORC_RADJ(gb_radius[j]);
            double Rk = ORC_RADJ(gb_radius[o]);
            double rjk = 1.01 * d12[k];
            double h = Rj * (1.0 + (sqr(Rk) - sqr(Rj) - sqr(rjk)) / (2.0 * Rj * rjk));
            vj[j] -= kPI * sqr(h) * (3 * Rj - h) / 3;
        }
    }
    for (int j = 0; j < n; ++j)
        if (vj[j] < 0) vj[j] = 0;
    for (int i = 0; i < n; ++i) {
        double Ri = ORC_RADJ(gb_radius[i]);
        terms123[i] = -166.0 / (Ri + p->dielectric_offset + p->P1);
    }
    if (n <= 0) return;
    double *t2 = (double *)calloc((size_t)n, sizeof(double));
    double *t3 = (double *)calloc((size_t)n, sizeof(double));
    for (int k = 0; k < n12; ++k) {
        double r4 = sqr(d12[k]) * sqr(d12[k]);
        t2[a12[k]] += p->P2 * vj[b12[k]] / r4;
        t2[b12[k]] += p->P2 * vj[a12[k]] / r4;
    }
    for (int k = 0; k < n13; ++k) {
        double r4 = sqr(d13[k]) * sqr(d13[k]);
        t3[a13[k]] += p->P3 * vj[b13[k]] / r4;
        t3[b13[k]] += p->P3 * vj[a13[k]] / r4;
    }
    for (int i = 0; i < n; ++i) terms123[i] = terms123[i] + t2[i] + t3[i];
    free(t2);
    free(t3);
#undef ORC_RADJ
}

/* ---------------- A4: forcefields/gbff.cpp:415-466 ---------------- */
void orc_born_radii(const orc_system *s, const double *xyz, const orc_nlist *nl,
                    const orc_params *p, double *born) {
    const double halfeconv = econv_kcal() * 0.5;
    const double invP5 = 1 / p->P5;
    const double still2 = sqr(p->still_cutoff);
    for (int i = 0; i < s->n; ++i) {
        const double Ri = s->gb_radius[i];
        double Gpol = s->terms123[i];
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const uint32_t e = nl->entries[k];
            const int j = nl_j(e);
            double g = xyz[3 * i] - xyz[3 * j];
            double h = xyz[3 * i + 1] - xyz[3 * j + 1];
            double q = xyz[3 * i + 2] - xyz[3 * j + 2];
            double d2 = g * g + h * h + q * q;
            if (d2 > still2) continue;
            if (nl_bo(e) <= 2) continue; /* 1-2 and 1-3 pairs excluded */
            const double Rj = s->gb_radius[j];
            double rdivR = d2 / sqr(Ri + Rj);
            double CCF = rdivR > invP5 ? 1.0 : sqr(0.5 - 0.5 * cos(rdivR * p->P5 * kPI));
            Gpol += p->P4 * s->vj[j] * CCF / sqr(d2);
        }
        born[i] = -halfeconv / Gpol + p->born_radius_offset;
    }
}

/* ---------------- A5 + A6: forcefields/gbff.cpp:1449-1818 ---------------- */
void orc_gb_calc_forces(const orc_system *s, const double *xyz, const orc_nlist *nl,
                        const orc_params *p, double *force, orc_energies *e,
                        double *born_out, double *deda_out) {
    const int n = s->n;
    double *a = (double *)malloc(sizeof(double) * (size_t)n);
    double *deda = (double *)calloc((size_t)n, sizeof(double));
    double *dedi = (double *)calloc((size_t)n * 3, sizeof(double));
    orc_born_radii(s, xyz, nl, p, a);

    const double econvJ = econv_joule();
    const double premul = -(econvJ * 0.5) * (1 / p->dielectric_solute - 1 / p->dielectric_solvent);
    const double sqrcutoff = sqr(p->cutoff);
    const double Swidth = p->cutoff - p->inner_cutoff;
    const double invSwidth = 1.0 / Swidth;
    const double vdwSwidth = p->vdw_cutoff - p->vdw_inner_cutoff;
    const double vdwinvSwidth = 1.0 / vdwSwidth;
    const double invdiel = 1 / p->dielectric_solute;
    const double invsolv = 1 / p->dielectric_solvent;

    double epot_vdw = 0, epot_elec = 0, epot = 0;
    double epot_pol = 0, epot_pol_self = 0, epot_pol_cross = 0;

    for (int i = 0; i < n; ++i) {
        const double ai = a[i];
        const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        const double Ri = s->radius[i], ei = s->epsilon[i], qi = s->charge[i];
        double dedai = 0;
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const uint32_t ent = nl->entries[k];
            const int j = nl_j(ent);
            if (j > i) break; /* shadow entries are not part of the pair pass */
            const int bo = nl_bo(ent);
            const double aj = a[j];
            const double qiqj = qi * s->charge[j];
            const double dx = xi - xyz[3 * j];
            const double dy = yi - xyz[3 * j + 1];
            const double dz = zi - xyz[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > sqrcutoff) continue;
            const double r = sqrt(r2);
            const double invr = 1 / r;
            const double aiaj = ai * aj;
            const double u = -r2 * 0.25 / aiaj;
            const double s14e = (bo == 3) ? p->elec14_scaling : 1.0;
            double elec_pot = 0, elec_force = 0;
            if (u > p->exp_approx_threshold) {
                const double t = exp(u);
                const double L = 1 / (r2 + t * aiaj);
                double epol = premul * qiqj * sqrt(L);
                double dedr = -epol * (r - 0.25 * r * t) * L;
                double dedalpha = -epol * t * (1.0 - u) * L;
                if (bo >= 3) {
                    elec_pot = econvJ * invdiel * s14e * qiqj * invr;
                    elec_force = -elec_pot * invr;
                }
                if (r >= p->inner_cutoff) {
                    double S = 1.0 - sqr(invSwidth * (r - p->inner_cutoff));
                    double dSdd = -4.0 * sqr(invSwidth) * (r - p->inner_cutoff) * S;
                    S = sqr(S);
                    elec_force = S * elec_force + elec_pot * dSdd;
                    elec_pot *= S;
                    dedr = S * dedr + dSdd * epol;
                    dedalpha *= S;
                    epol *= S;
                }
                dedr *= 2 * invr;
                epot_pol += epol;
                epot_pol += epol;
                epot_pol_cross += epol;
                epot_pol_cross += epol;
                dedi[3 * i] += dedr * dx; dedi[3 * i + 1] += dedr * dy; dedi[3 * i + 2] += dedr * dz;
                dedi[3 * j] -= dedr * dx; dedi[3 * j + 1] -= dedr * dy; dedi[3 * j + 2] -= dedr * dz;
                dedai += dedalpha * aj;
                deda[j] += dedalpha * ai;
            } else if (bo >= 3) {
                /* far pair: GB dropped, solvent-dielectric Coulomb instead */
                elec_pot = econvJ * invsolv * s14e * qiqj * invr;
                elec_force = -elec_pot * invr;
                if (r >= p->inner_cutoff) {
                    double S = 1.0 - sqr(invSwidth * (r - p->inner_cutoff));
                    double dSdd = -4.0 * sqr(invSwidth) * (r - p->inner_cutoff) * S;
                    S = sqr(S);
                    elec_force = S * elec_force + elec_pot * dSdd;
                    elec_pot *= S;
                }
            }
            if (bo >= 3) {
                const double s14v = (bo == 3) ? p->vdw14_scaling : 1.0;
                double vdw_force = 0;
                if (r < p->vdw_cutoff) {
                    const double rr = sqr(invr * (Ri + s->radius[j]));
                    const double eps = s14v * ei * s->epsilon[j];
                    const double B = rr * rr * rr;
                    const double A = sqr(B);
                    double vdw_pot = (2.0 * eps) * (0.5 * A - B);
                    vdw_force = -(12.0 * eps * invr) * (A - B);
                    if (r > p->vdw_inner_cutoff) {
                        double vS = 1.0 - sqr(vdwinvSwidth * (r - p->vdw_inner_cutoff));
                        double vdSdd = -4.0 * sqr(vdwinvSwidth) * (r - p->vdw_inner_cutoff) * vS;
                        vS = sqr(vS);
                        vdw_force = vS * vdw_force + vdw_pot * vdSdd;
                        vdw_pot *= vS;
                    }
                    epot += vdw_pot;
                    epot_vdw += vdw_pot;
                }
                epot += elec_pot;
                epot_elec += elec_pot;
                const double fm = (vdw_force + elec_force) * invr;
                dedi[3 * i] += fm * dx; dedi[3 * i + 1] += fm * dy; dedi[3 * i + 2] += fm * dz;
                dedi[3 * j] -= fm * dx; dedi[3 * j + 1] -= fm * dy; dedi[3 * j + 2] -= fm * dz;
            }
        }
        deda[i] += dedai;
        const double epol = premul * sqr(qi) / ai; /* self term */
        epot_pol += epol;
        epot_pol_self += epol;
        deda[i] += -epol / ai;
    }

    /* chain rule through the Born radii over the full list (gbff.cpp:1751-1808).
     * The literal 166.0 here (vs halfeconv in the radii) is PD's. */
    const double invP5 = 1 / p->P5;
    const double still2 = sqr(p->still_cutoff);
    for (int i = 0; i < n; ++i) {
        const double Ri = s->gb_radius[i];
        const double gpi = -sqr(a[i]) / -166.0;
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const uint32_t ent = nl->entries[k];
            const int j = nl_j(ent);
            const double dx = xyz[3 * j] - xyz[3 * i];
            const double dy = xyz[3 * j + 1] - xyz[3 * i + 1];
            const double dz = xyz[3 * j + 2] - xyz[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > still2) continue;
            if (nl_bo(ent) <= 2) continue;
            const double Rj = s->gb_radius[j];
            const double r6 = r2 * r2 * r2;
            const double rdivR = r2 / sqr(Ri + Rj);
            double CCF, dCCF;
            if (rdivR > invP5) {
                CCF = 1.0;
                dCCF = 0.0;
            } else {
                const double th = rdivR * p->P5 * kPI;
                const double sq = 0.5 - 0.5 * cos(th);
                CCF = sqr(sq);
                dCCF = 2.0 * sq * sin(th) * th;
            }
            const double dedr = gpi * deda[i] * p->P4 * s->vj[j] * (4.0 * CCF - dCCF) / r6;
            dedi[3 * i] += dedr * dx; dedi[3 * i + 1] += dedr * dy; dedi[3 * i + 2] += dedr * dz;
            dedi[3 * j] -= dedr * dx; dedi[3 * j + 1] -= dedr * dy; dedi[3 * j + 2] -= dedr * dz;
        }
    }

    if (force) {
        const double conv = -1.0 / kAngstrom;
        for (int i = 0; i < 3 * n; ++i) force[i] += dedi[i] * conv;
    }
    if (e) {
        /* the reference overwrites epot_vdw/epot_elec, adds the rest */
        e->epot_vdw = epot_vdw;
        e->epot_elec = epot_elec;
        e->epot += epot + epot_pol;
        e->epot_pol += epot_pol;
        e->epot_pol_cross += epot_pol_cross;
        e->epot_pol_self += epot_pol_self;
    }
    if (born_out) memcpy(born_out, a, sizeof(double) * (size_t)n);
    if (deda_out) memcpy(deda_out, deda, sizeof(double) * (size_t)n);
    free(a);
    free(deda);
    free(dedi);
}

/* ---------------- A8: forcefields/nonbonded.cpp:542-893 ---------------- */
/* FF_SoftVDW::calcForces_Soft, forcefields/ffsoftvdw.cpp:71-145: over the
 * list pairs j < i of bond order above 1-4 with r^2 < (R_i+R_j-olap)^2 and
 * r^2 < vdw_cutoff^2, E = H (1 - A^2 r^2)^2 with A = 1/(R_i+R_j-olap) and H
 * the hardness in J/molecule; the force is 4 H/A (A^2 r^2 A^2 r - A^2 r)
 * along i-j, subtracted from atom i and added to atom j. */
void orc_softvdw_forces(const orc_system *s, const double *xyz, const orc_nlist *nl, double hardness,
                        double olap, double vdw_cutoff, double *force, orc_energies *e) {
    const double localHardness = hardness * (kKcal2J / kNa);
    const double fourInvAnglocalHardness = 4.0 * (1.0 / kAngstrom) * localHardness;
    const double sqrVdwCutoff = vdw_cutoff * vdw_cutoff;
    double esum = 0;
    for (int i = 0; i < s->n; ++i) {
        const double ri = s->radius[i];
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const uint32_t ent = nl->entries[k];
            const int j = nl_j(ent);
            if (j > i) break;
            if (nl_bo(ent) <= 3) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                         dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double sqrdist = dx * dx + dy * dy + dz * dz;
            double vsum = ri + s->radius[j];
            vsum -= olap;
            const double sqrvsum = vsum * vsum;
            if (sqrdist < sqrvsum && sqrdist < sqrVdwCutoff) {
                const double dist = sqrt(sqrdist);
                const double A = 1.0 / vsum;
                const double A2X = A * A * dist;
                const double A2X2 = A2X * dist;
                double pot = 1.0 - A2X2;
                pot *= pot;
                pot *= localHardness;
                const double fmag = fourInvAnglocalHardness * ((A2X2 * A2X) - A2X);
                esum += pot;
                const double sc = fmag / dist;
                force[3 * i] -= dx * sc; force[3 * i + 1] -= dy * sc; force[3 * i + 2] -= dz * sc;
                force[3 * j] += dx * sc; force[3 * j + 1] += dy * sc; force[3 * j + 2] += dz * sc;
            }
        }
    }
    e->epot += esum;
    e->epot_vdw = esum;
}

void orc_nb_calc_forces(const orc_system *s, const double *xyz, const orc_nlist *nl,
                        const orc_params *p, double *force, orc_energies *e) {
    const int n = s->n;
    const double econvJ = econv_joule();
    const double sqrcutoff = sqr(p->cutoff);
    const double sqrinner = sqr(p->inner_cutoff);
    const double invdiel = 1.0 / p->dielectric_solute;
    const double invSw = 1.0 / (p->cutoff - p->inner_cutoff);
    const double vinvSw = 1.0 / (p->vdw_cutoff - p->vdw_inner_cutoff);
    /* Steinbach-Brooks force-switch constants (nonbonded.cpp:601-616) */
    const double rc = p->cutoff, ri = p->inner_cutoff;
    const double rc2 = sqr(rc), ri2 = sqr(ri);
    const double sA = 1.0 / ((rc2 - ri2) * (rc2 - ri2) * (rc2 - ri2));
    const double sB = -(rc2 * rc2 * rc2 - 3.0 * rc2 * rc2 * ri2);
    const double sC = 6.0 * rc2 * ri2;
    const double sD = -(rc2 + ri2);
    const double sE = 2.0 / 5.0;
    const double fsw_in = sA * (sB * (1.0 / ri) + sC * ri + sD * ri * ri * ri + sE * ri * ri * ri * ri2);
    const double fsw_cut = sA * (sB * (1.0 / rc) + sC * rc + sD * rc * rc * rc + sE * rc * rc * rc * rc2);
    const double eshift = (ri < rc) ? 1 / ri + (fsw_in - fsw_cut) : 0;
    const double tv[8] = {0, 0, 0, p->vdw14_scaling, 1, 1, 1, 1};
    const double te[8] = {0, 0, 0, p->elec14_scaling, 1, 1, 1, 1};

    double ev = 0, ee = 0, ep = 0, vir = 0;
    for (int i = 0; i < n; ++i) {
        const double Ri = s->radius[i], ei = s->epsilon[i], qi = s->charge[i];
        double fix = 0, fiy = 0, fiz = 0;
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const uint32_t ent = nl->entries[k];
            const int j = nl_j(ent);
            if (j >= i) break;
            const int bo = nl_bo(ent);
            double fx = xyz[3 * j] - xyz[3 * i];
            double fy = xyz[3 * j + 1] - xyz[3 * i + 1];
            double fz = xyz[3 * j + 2] - xyz[3 * i + 2];
            const double d2 = fx * fx + fy * fy + fz * fz;
            if (d2 > sqrcutoff) continue;
            const double r = sqrt(d2);
            const double inv = 1.0 / r;
            double vdw_pot = 0, vdw_force = 0;
            if (p->vdw_mode > 0 && r < p->vdw_cutoff) {
                const double rij = Ri + s->radius[j];
                const double eps = tv[bo] * ei * s->epsilon[j];
                double B = sqr(rij * inv);
                B *= B * B;
                const double A = sqr(B);
                vdw_pot = (2.0 * eps) * (0.5 * A - B);
                vdw_force = -(12.0E10 * eps * inv) * (A - B);
                if (p->vdw_mode == 2 && r > p->vdw_inner_cutoff) {
                    double vS = 1.0 - sqr(vinvSw * (r - p->vdw_inner_cutoff));
                    double vd = -4.0E10 * sqr(vinvSw) * (r - p->vdw_inner_cutoff) * vS;
                    vS = sqr(vS);
                    vdw_force = vS * vdw_force + vdw_pot * vd;
                    vdw_pot *= vS;
                }
            }
            double elec_pot = 0, elec_force = 0;
            const double qiqj = qi * s->charge[j];
            if (p->elec_mode == 1 || p->elec_mode == 3) {
                elec_pot = econvJ * invdiel * te[bo] * qiqj * inv;
                elec_force = -elec_pot * inv * 1E10;
                if (p->elec_mode == 3 && r > p->inner_cutoff) {
                    double S = 1.0 - sqr(invSw * (r - p->inner_cutoff));
                    double dS = -4.0E10 * sqr(invSw) * (r - p->inner_cutoff) * S;
                    S = sqr(S);
                    elec_force = S * elec_force + elec_pot * dS;
                    elec_pot *= S;
                }
            } else if (p->elec_mode == 4) {
                elec_pot = econvJ * invdiel * te[bo] * qiqj;
                if (r > p->inner_cutoff) {
                    elec_force = -1E10 * elec_pot * sA * sqr(inv) * sqr(sqrcutoff - d2) *
                                 (sqrcutoff - 3.0 * sqrinner + 2.0 * d2);
                    elec_pot *= -(sA * r * (sB * sqr(inv) + sC + sD * d2 + sE * sqr(d2)) - fsw_cut);
                } else {
                    elec_force = -1E10 * elec_pot * sqr(inv);
                    elec_pot *= (inv - eshift);
                }
            }
            ep += vdw_pot + elec_pot;
            ev += vdw_pot;
            ee += elec_pot;
            const double fm = vdw_force + elec_force;
            vir += r * fm * kAngstrom;
            const double sc = inv * fm;
            fx *= sc; fy *= sc; fz *= sc;
            fix += fx; fiy += fy; fiz += fz;
            if (force) { force[3 * j] -= fx; force[3 * j + 1] -= fy; force[3 * j + 2] -= fz; }
        }
        if (force) { force[3 * i] += fix; force[3 * i + 1] += fiy; force[3 * i + 2] += fiz; }
    }
    if (e) {
        e->epot_vdw = ev;
        e->epot_elec = ee;
        e->epot += ep;
        e->virial += vir;
    }
}

/* ---------------- A7 energy path: gbff.cpp:161-166, 520-579 ---------------- */
void orc_gb_calc_energies(const orc_system *s, const double *xyz, const orc_nlist *nl,
                          const orc_params *p, double *force, orc_energies *e,
                          double *born_out) {
    const int n = s->n;
    double *a = (double *)malloc(sizeof(double) * (size_t)n);
    orc_born_radii(s, xyz, nl, p, a);
    orc_nb_calc_forces(s, xyz, nl, p, force, e);
    const double premul = -(econv_joule() * 0.5) * (1 / p->dielectric_solute - 1 / p->dielectric_solvent);
    const double invSw = 1 / (p->cutoff - p->inner_cutoff);
    double pol = 0, self = 0, cross = 0;
    for (int i = 0; i < n; ++i) {
        const double qi = s->charge[i], ai = a[i];
        pol -= sqr(qi) / ai;
        self -= sqr(qi) / ai;
        for (int k = nl->row_start[i]; k < nl->row_start[i + 1]; ++k) {
            const int j = nl_j(nl->entries[k]);
            if (i > j) continue; /* only entries with j > i (shadow half) */
            const double g = xyz[3 * i] - xyz[3 * j];
            const double h = xyz[3 * i + 1] - xyz[3 * j + 1];
            const double q = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double r = sqrt(g * g + h * h + q * q);
            if (r > p->cutoff) continue;
            double S;
            if (r < p->inner_cutoff) S = 1.0;
            else { S = 1.0 - sqr(invSw * (r - p->inner_cutoff)); S = sqr(S); }
            const double r2 = sqr(r);
            const double aiaj = ai * a[j];
            const double epol = 2.0 * S * qi * s->charge[j] / sqrt(r2 + aiaj * exp(-r2 * 0.25 / aiaj));
            pol -= epol;
            cross -= epol;
        }
    }
    pol *= -premul;
    cross *= -premul;
    self *= -premul;
    if (e) {
        e->epot += pol;
        e->epot_pol += pol;
        e->epot_pol_cross += cross;
        e->epot_pol_self += self;
    }
    if (born_out) memcpy(born_out, a, sizeof(double) * (size_t)n);
    free(a);
}

/* ---------------- FF_Bonded: forcefields/ffbonded.cpp:809-1590 ---------------- */
typedef struct { double x, y, z; } v3;
static inline v3 v3_sub(const double *a, const double *b) { v3 r = {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; return r; }
static inline v3 v3_cross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline double v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_scale(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static inline v3 v3_axpy(v3 a, double s, v3 b) { v3 r = {a.x * s - b.x, a.y * s - b.y, a.z * s - b.z}; return r; }
static inline void f_add(double *f, int i, v3 v, double s) {
    if (!f) return;
    f[3 * i] += s * v.x; f[3 * i + 1] += s * v.y; f[3 * i + 2] += s * v.z;
}

/* calcBondForces (:885-940): E = k(d-l)^2, force 2k(d-l)/Angstrom along the
 * bond, virial d * force * Angstrom. */
static void bonded_bonds(const orc_bonded *b, const double *xyz, double *f, orc_energies *e) {
    for (int t = 0; t < b->nbonds; ++t) {
        const int i = b->bond_ij[2 * t], j = b->bond_ij[2 * t + 1];
        const double l = b->bond_lk[2 * t], k = b->bond_lk[2 * t + 1];
        const v3 v = v3_sub(xyz + 3 * i, xyz + 3 * j);
        const double d = sqrt(v3_dot(v, v));
        const double fm = 2.0 * k * (d - l) / kAngstrom;
        e->virial += d * fm * kAngstrom;
        e->epot_bond += k * sqr(d - l);
        f_add(f, i, v, -fm / d);
        f_add(f, j, v, fm / d);
    }
}

/* calcAngleForces (:994-1120): theta from the unit bond vectors at the apex,
 * E = k(theta-theta0)^2; terms with k == 0 are skipped. */
static void bonded_angles(const orc_bonded *b, const double *xyz, double *f, orc_energies *e) {
    for (int t = 0; t < b->nangles; ++t) {
        const double k = b->angle_tk[2 * t + 1];
        if (k == 0) continue;
        const int i = b->angle_iaj[3 * t], a = b->angle_iaj[3 * t + 1], j = b->angle_iaj[3 * t + 2];
        const v3 vi = v3_sub(xyz + 3 * i, xyz + 3 * a), vj = v3_sub(xyz + 3 * j, xyz + 3 * a);
        const double ri = 1.0 / sqrt(v3_dot(vi, vi)), rj = 1.0 / sqrt(v3_dot(vj, vj));
        const v3 ui = v3_scale(vi, ri), uj = v3_scale(vj, rj);
        const double c = v3_dot(ui, uj);
        const double theta = acos(c), sn = sin(theta);
        const double dth = theta - b->angle_tk[2 * t];
        e->epot_angle += k * dth * dth;
        const double fm = 2.0 * k * dth / kAngstrom;
        /* dtheta/dx_i = (u_i cos - u_j) / (|v_i| sin); the apex takes the rest */
        const v3 gi = v3_scale(v3_axpy(ui, c, uj), fm * ri / sn);
        const v3 gj = v3_scale(v3_axpy(uj, c, ui), fm * rj / sn);
        f_add(f, i, gi, -1.0);
        f_add(f, a, gi, 1.0);
        f_add(f, j, gj, -1.0);
        f_add(f, a, gj, 1.0);
    }
}

/* calcDihedralForces (:1199-1440). With v1 = i-a, v2 = a-b, v3 = b-j,
 * m = v1 x v2, n = v2 x v3, o = v2 x m: cos phi from m.n, sin from o.n,
 * phi = -atan2(sin, cos) (the IUPAC dihedral). The gradient goes through
 * d(cos)/d(m,n) when |sin phi| > 0.1 and through d(sin)/d(o,n) otherwise,
 * as the reference does. Returns the energy of the dihedral. */
static double bonded_dihedral(const double *xyz, const int32_t *q, int terms, const double *vn,
                              const double *nn, const double *gam, double *f) {
    const int i = q[0], a = q[1], bb = q[2], j = q[3];
    const v3 v1 = v3_sub(xyz + 3 * i, xyz + 3 * a);
    const v3 v2 = v3_sub(xyz + 3 * a, xyz + 3 * bb);
    const v3 v3_ = v3_sub(xyz + 3 * bb, xyz + 3 * j);
    v3 m = v3_cross(v1, v2), n = v3_cross(v2, v3_), o = v3_cross(v2, m);
    const double im = 1.0 / sqrt(v3_dot(m, m)), in = 1.0 / sqrt(v3_dot(n, n)), io = 1.0 / sqrt(v3_dot(o, o));
    const double cph = v3_dot(m, n) * im * in;
    const double sph = v3_dot(o, n) * io * in;
    n = v3_scale(n, in);
    const double phi = -atan2(sph, cph);
    const int use_cos = fabs(sph) > 0.1;
    v3 g1, g2; /* d(cos)/dm, d(cos)/dn   or   d(sin)/do, d(sin)/dn */
    if (use_cos) {
        m = v3_scale(m, im);
        g1 = v3_scale(v3_axpy(m, cph, n), im);
        g2 = v3_scale(v3_axpy(n, cph, m), in);
    } else {
        o = v3_scale(o, io);
        g1 = v3_scale(v3_axpy(o, sph, n), io);
        g2 = v3_scale(v3_axpy(n, sph, o), in);
    }
    double etot = 0;
    v3 fi = {0, 0, 0}, fab = {0, 0, 0}, fj = {0, 0, 0};
    for (int t = 0; t < terms; ++t) {
        double en, dedphi;
        if (nn[t] > 0) {
            en = vn[t] * (1.0 + cos(nn[t] * phi + gam[t]));
            dedphi = -nn[t] * vn[t] * sin(nn[t] * phi + gam[t]);
        } else {
            double d = phi - gam[t];
            if (d < -kPI) d += 2.0 * kPI;
            else if (d > kPI) d -= 2.0 * kPI;
            en = vn[t] * d * d;
            dedphi = 2.0 * vn[t] * d;
        }
        etot += en;
        if (use_cos) {
            const double s = dedphi / sph;
            /* m = v1 x v2 and n = v2 x v3 carried back to the atoms */
            const v3 ci = v3_cross(v2, g1);
            const v3 cj = v3_cross(g2, v2);
            const v3 cab1 = v3_cross(g1, v1), cab2 = v3_cross(v3_, g2);
            fi.x += s * ci.x; fi.y += s * ci.y; fi.z += s * ci.z;
            fj.x += s * cj.x; fj.y += s * cj.y; fj.z += s * cj.z;
            fab.x += s * (cab1.x + cab2.x); fab.y += s * (cab1.y + cab2.y); fab.z += s * (cab1.z + cab2.z);
        } else {
            const double s = -dedphi / cph;
            /* o = v2 x (v1 x v2): d(o)/d(v1) = |v2|^2 I - v2 v2^T, and the
             * v2-dependence of o, m and n for the middle atoms */
            const double b2 = v3_dot(v2, v2);
            const double gv2 = v3_dot(g1, v2);
            const v3 ti = {b2 * g1.x - v2.x * gv2, b2 * g1.y - v2.y * gv2, b2 * g1.z - v2.z * gv2};
            const v3 cj = v3_cross(g2, v2);
            /* d(o)/d(v2) acting on g1: 2 v2 (v1.g1) ... written per component */
            const double v1g = v3_dot(v1, g1), v1v2 = v3_dot(v1, v2);
            const v3 tab = {
                -v1v2 * g1.x + 2.0 * v2.x * v1g - v1.x * gv2,
                -v1v2 * g1.y + 2.0 * v2.y * v1g - v1.y * gv2,
                -v1v2 * g1.z + 2.0 * v2.z * v1g - v1.z * gv2};
            const v3 cab2 = v3_cross(v3_, g2);
            fi.x += s * ti.x; fi.y += s * ti.y; fi.z += s * ti.z;
            fj.x += s * cj.x; fj.y += s * cj.y; fj.z += s * cj.z;
            fab.x += s * (tab.x + cab2.x); fab.y += s * (tab.y + cab2.y); fab.z += s * (tab.z + cab2.z);
        }
    }
    if (f) {
        const double u = 1.0 / kAngstrom;
        f[3 * i] += fi.x * u; f[3 * i + 1] += fi.y * u; f[3 * i + 2] += fi.z * u;
        f[3 * a] += (fab.x - fi.x) * u; f[3 * a + 1] += (fab.y - fi.y) * u; f[3 * a + 2] += (fab.z - fi.z) * u;
        f[3 * bb] += (fj.x - fab.x) * u; f[3 * bb + 1] += (fj.y - fab.y) * u; f[3 * bb + 2] += (fj.z - fab.z) * u;
        f[3 * j] -= fj.x * u; f[3 * j + 1] -= fj.y * u; f[3 * j + 2] -= fj.z * u;
    }
    return etot;
}

void orc_bonded_forces(const orc_bonded *b, const double *xyz, double *force, orc_energies *e) {
    orc_energies z;
    memset(&z, 0, sizeof(z));
    bonded_bonds(b, xyz, force, &z);
    bonded_angles(b, xyz, force, &z);
    for (int t = 0; t < b->ntorsions; ++t)
        z.epot_torsion += bonded_dihedral(xyz, b->tors_iabj + 4 * t, b->tors_terms[t], b->tors_vn + 4 * t,
                                          b->tors_n + 4 * t, b->tors_gamma + 4 * t, force);
    for (int t = 0; t < b->nimpropers; ++t)
        z.epot_improper += bonded_dihedral(xyz, b->impr_iabj + 4 * t, b->impr_terms[t], b->impr_vn + 4 * t,
                                           b->impr_n + 4 * t, b->impr_gamma + 4 * t, force);
    /* calcImproperForces adds impropers into the Hamiltonian's epot_torsion (:1556) */
    e->epot_bond += z.epot_bond;
    e->epot_angle += z.epot_angle;
    e->epot_torsion += z.epot_torsion + z.epot_improper;
    e->epot_improper += z.epot_improper;
    e->virial += z.virial;
    e->epot += z.epot_bond + z.epot_angle + z.epot_torsion + z.epot_improper;
}

/* ---------------- A11: protocols/minimise.cpp:99-235 ---------------- */
static int build_list(const orc_system *s, const double *xyz, const orc_params *p, orc_nlist *nl) {
    const double lc = p->cutoff > p->vdw_cutoff ? p->cutoff : p->vdw_cutoff;
    for (;;) {
        int64_t need = orc_nlist_build(s, xyz, lc, p->padding, 1, nl);
        if (need == 0) return 0;
        free(nl->entries);
        nl->cap = need + need / 2;
        nl->entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl->cap);
        if (!nl->entries) return -1;
    }
}

static void eval_forces(const orc_system *s, const double *xyz, const orc_nlist *nl,
                        const orc_params *p, double *f, orc_energies *e) {
    memset(e, 0, sizeof(*e));
    memset(f, 0, sizeof(double) * 3 * (size_t)s->n);
    /* Forcefield::calcForces runs its components in the order they were
     * added (forcefields/forcefield.cpp:220-227): FF_Bonded, then GB */
    if (s->bonded) orc_bonded_forces(s->bonded, xyz, f, e);
    orc_gb_calc_forces(s, xyz, nl, p, f, e, NULL, NULL);
}

int orc_minimise_sd(const orc_system *s, double *xyz, const orc_params *p,
                    int steps, double step_size, int update_nlist,
                    double *trace_e, orc_energies *e_final, double *force_final) {
    const int n = s->n;
    orc_nlist nl;
    nl.row_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    nl.cap = (int64_t)n * 64 + 1024;
    nl.entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl.cap);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *op = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_energies e;
    int rc = 0;
    build_list(s, xyz, p, &nl);
    memcpy(op, xyz, sizeof(double) * 3 * (size_t)n); /* wspace.old = wspace.cur */
    double mult = 1.0, old_e = 0;
    int step;
    for (step = 0; step < steps; ++step) {
        if (update_nlist > 0 && step % update_nlist == 0) build_list(s, xyz, p, &nl);
        eval_forces(s, xyz, &nl, p, f, &e);
        if (trace_e) trace_e[step] = e.epot;
        if (!isfinite(e.epot)) { rc = -1; goto done; }
        if (step != 0) {
            if (e.epot <= old_e) {
                mult *= 1.2;
                memcpy(op, xyz, sizeof(double) * 3 * (size_t)n);
                old_e = e.epot;
            } else {
                mult *= 0.5 / 1.2;
                memcpy(xyz, op, sizeof(double) * 3 * (size_t)n);
                old_e += 1; /* PD adds 1 J/molecule: the next step is always accepted */
                continue;
            }
        } else {
            memcpy(op, xyz, sizeof(double) * 3 * (size_t)n);
            old_e = e.epot;
        }
        for (int k = 0; k < 3 * n; ++k) xyz[k] += f[k] * (step_size * mult);
    }
    rc = step;
    memcpy(xyz, op, sizeof(double) * 3 * (size_t)n); /* cur = old */
    eval_forces(s, xyz, &nl, p, f, &e);
    if (e_final) *e_final = e;
    if (force_final) memcpy(force_final, f, sizeof(double) * 3 * (size_t)n);
done:
    free(nl.row_start);
    free(nl.entries);
    free(f);
    free(op);
    return rc;
}

/* Textbook velocity Verlet; see header. */
int orc_md_vv(const orc_system *s, double *xyz, double *vel, const orc_params *p,
              int steps, double dt, int update_nlist, double *trace_epot, double *trace_ekin) {
    const int n = s->n;
    orc_nlist nl;
    nl.row_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    nl.cap = (int64_t)n * 64 + 1024;
    nl.entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl.cap);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_energies e;
    int rc = steps;
    build_list(s, xyz, p, &nl);
    eval_forces(s, xyz, &nl, p, f, &e);
    for (int step = 0; step < steps; ++step) {
        for (int i = 0; i < n; ++i) {
            const double h = 0.5 * dt / s->mass[i];
            for (int c = 0; c < 3; ++c) {
                vel[3 * i + c] += f[3 * i + c] * h;
                xyz[3 * i + c] += vel[3 * i + c] * dt / kAngstrom;
            }
        }
        if (update_nlist > 0 && (step + 1) % update_nlist == 0) build_list(s, xyz, p, &nl);
        eval_forces(s, xyz, &nl, p, f, &e);
        double ekin = 0;
        for (int i = 0; i < n; ++i) {
            const double h = 0.5 * dt / s->mass[i];
            for (int c = 0; c < 3; ++c) {
                vel[3 * i + c] += f[3 * i + c] * h;
                ekin += 0.5 * s->mass[i] * sqr(vel[3 * i + c]);
            }
        }
        if (trace_epot) trace_epot[step] = e.epot;
        if (trace_ekin) trace_ekin[step] = ekin;
        if (!isfinite(e.epot)) { rc = -1; break; }
    }
    free(nl.row_start);
    free(nl.entries);
    free(f);
    return rc;
}

/* MolecularDynamics::run_core with the VelocityVerlet integrator, as written
 * (protocols/md.cpp:209-233 loop, 788-838 update, protocolbase.cpp:113-122
 * list refresh): each step rebuilds the list when step % update_nlist == 0,
 * evaluates forces, adds f*dt/2m to v (the kinetic energy is taken here),
 * then adds it again, moves p by dt*(v + f*dt/2m) and adds it a third time.
 * No thermostat or barostat (PD's defaults); momentum removal only acts with
 * a thermostat (md.cpp:246-258). */
/* MolecularDynamics::applyBerendsenThermostat (md.cpp:580-597): every
 * velocity times sqrt(1 + dt (Edes / ekin - 1) / tau), Edes = 3 N kB T / 2. */
static void berendsen(int n, double *vel, double ekin, double dt, double tau, double temperature) {
    if (!(tau > 0)) return;
    const double edes = 3.0 * n * kKB * temperature / 2.0;
    const double chi = sqrt(1.0 + dt * (edes / ekin - 1.0) / tau);
    for (int k = 0; k < 3 * n; ++k) vel[k] *= chi;
}

int orc_md_vv_pd(const orc_system *s, double *xyz, double *vel, const orc_params *p,
                 int steps, double dt, int update_nlist, double tau, double temperature,
                 double *trace_epot, double *trace_ekin) {
    const int n = s->n;
    orc_nlist nl;
    nl.row_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    nl.cap = (int64_t)n * 64 + 1024;
    nl.entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl.cap);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_energies e;
    int rc = steps;
    for (int step = 0; step < steps; ++step) {
        if (update_nlist > 0 && step % update_nlist == 0) build_list(s, xyz, p, &nl);
        eval_forces(s, xyz, &nl, p, f, &e);
        double ekin = 0;
        for (int i = 0; i < n; ++i) {
            const double tinvmass = 0.5 * dt / s->mass[i];
            for (int c = 0; c < 3; ++c) vel[3 * i + c] += f[3 * i + c] * tinvmass;
        }
        for (int i = 0; i < n; ++i) /* calcKineticEnergy, md.cpp:504-518: 0.5 m |v|^2 via mag() */
            ekin += 0.5 * s->mass[i] * sqr(sqrt(sqr(vel[3 * i]) + sqr(vel[3 * i + 1]) + sqr(vel[3 * i + 2])));
        berendsen(n, vel, ekin, dt, tau, temperature); /* applyThermostat, md.cpp:816 */
        for (int i = 0; i < n; ++i) {
            const double tinvmass = 0.5 * dt / s->mass[i];
            double dv[3], dr[3];
            for (int c = 0; c < 3; ++c) {
                dv[c] = f[3 * i + c] * tinvmass;
                vel[3 * i + c] += dv[c];
            }
            for (int c = 0; c < 3; ++c) dr[c] = dt * (vel[3 * i + c] + dv[c]);
            for (int c = 0; c < 3; ++c) {
                dr[c] *= 1 / kAngstrom;
                xyz[3 * i + c] += dr[c];
                vel[3 * i + c] += dv[c];
            }
        }
        if (trace_epot) trace_epot[step] = e.epot;
        if (trace_ekin) trace_ekin[step] = ekin;
        if (!isfinite(e.epot)) { rc = -1; break; }
    }
    free(nl.row_start);
    free(nl.entries);
    free(f);
    return rc;
}

/* ---- A11 Langevin integrator: protocols/md.cpp:906-1043 ---- */

/* The coefficients of applyForces_LangevinIntegration for gamma*dt
 * (md.cpp:950-985): exact exponential forms above 0.05, the reference's
 * Taylor series in between. Returns 0 below 1e-5, where the reference runs
 * its VelocityVerlet update instead. */
int orc_langevin_coef(double gamma, double dt, double *c0, double *c1, double *c2, double *varv, double *varr,
                      double *crv) {
    const double gt = gamma * dt;
    if (gt < 0.00001) return 0;
    if (gt > 0.05) {
        *c0 = exp(-gt);
        *c1 = (1 - *c0) / gt;
        *c2 = (1 - *c1) / gt;
        const double eg = exp(-gt);
        *varv = 1.0 - sqr(eg);
        *varr = 2.0 * gt - 3.0 + (4.0 - eg) * eg;
        *crv = sqr(1.0 - eg) / sqrt(*varv * *varr);
        return 1;
    }
    const double g2 = gt * gt, g3 = gt * g2, g4 = g2 * g2, g5 = g2 * g3, g6 = g3 * g3, g7 = g3 * g4, g8 = g4 * g4,
                 g9 = g4 * g5;
    *c0 = 1 - gt + 0.5 * g2 - g3 / 6.0 + g4 / 24.0 - g5 / 120.0 + g6 / 720.0 - g7 / 5040.0 + g8 / 40320.0 -
          g9 / 362880.0;
    *c1 = 1 - (0.5 * gt - g2 / 6.0 + g3 / 24.0 - g4 / 120.0 + g5 / 720.0 - g6 / 5040.0 + g7 / 40320.0 -
               g8 / 362880.0);
    *c2 = 0.5 - gt / 6.0 + g2 / 24.0 - g3 / 120.0 + g4 / 720.0 - g5 / 5040.0 + g6 / 40320.0 - g7 / 362880.0;
    *varr = 2.0 * g3 / 3.0 - g4 / 2.0 + 7.0 * g5 / 30.0 - g6 / 12.0 + 31.0 * g7 / 1260.0 - g8 / 160.0 +
            127.0 * g9 / 90720.0;
    *varv = 2.0 * gt - 2.0 * g2 + 4.0 * g3 / 3.0 - 2.0 * g4 / 3.0 + 4.0 * g5 / 15.0 - 4.0 * g6 / 45.0 +
            8.0 * g7 / 315.0 - 2.0 * g8 / 315.0 + 4.0 * g9 / 2835.0;
    *crv = sqrt(3.0) * (0.5 - 3.0 * gt / 16.0 - 17.0 * g2 / 1280.0 + 17.0 * g3 / 6144.0 + 40967.0 * g4 / 34406400.0 -
                        57203.0 * g5 / 275251200.0 - 1429487.0 * g6 / 13212057600.0);
    return 2;
}

/* The normal deviates the reference draws in a Langevin run: per step
 * (0..steps-2) and atom, nrand(n1x, n2y), nrand(n2x, n1z), nrand(n1y, n2z)
 * (md.cpp:1017-1019), each nrand being two frand() = (1 + rand()) /
 * (RAND_MAX + 1) through Box-Muller at -2 pi (maths/maths.h:138-152,
 * maths/maths.cpp:256-261), after srand(seed). Written per step and atom as
 * (n1x, n1y, n1z, n2x, n2y, n2z); the last step's slots are zero, and so
 * are those of atoms with skip[i] set (hydrogens when LangevinOnHydrogens
 * is off: applyForces_LangevinIntegration_NonHydrogens draws nothing for
 * them, md.cpp:1160-1190). */
void orc_langevin_noise(unsigned seed, int natoms, int steps, const int8_t *skip, double *out) {
    srand(seed);
    const double two_pi = 2.0 * kPI;
    for (int st = 0; st < steps; ++st) {
        for (int i = 0; i < natoms; ++i) {
            double *o = out + ((size_t)st * natoms + i) * 6;
            if (st == steps - 1 || (skip && skip[i])) {
                for (int k = 0; k < 6; ++k) o[k] = 0.0;
                continue;
            }
            double pr[6];
            for (int q = 0; q < 3; ++q) {
                const double z1 = (1.0 + (double)rand()) / ((double)RAND_MAX + 1.0);
                const double z2 = (1.0 + (double)rand()) / ((double)RAND_MAX + 1.0);
                const double l = sqrt(-2.0 * log(z1));
                pr[2 * q] = l * cos(-two_pi * z2);
                pr[2 * q + 1] = l * sin(-two_pi * z2);
            }
            /* call order: (n1x, n2y), (n2x, n1z), (n1y, n2z) */
            o[0] = pr[0]; o[4] = pr[1];
            o[3] = pr[2]; o[2] = pr[3];
            o[1] = pr[4]; o[5] = pr[5];
        }
    }
}

/* MolecularDynamics::run_core with Integrator = Langevin: with hydrogen ==
 * NULL LangevinOnHydrogens (PD's default), otherwise atoms with hydrogen[i]
 * set move by PD's plain half kicks (applyForces_LangevinIntegration_
 * NonHydrogens, md.cpp:1049-1195). Per step, list refresh, forces, then
 * (step > 0) v += (1 - c0/c1) f dt/m / (gamma dt); ekin; (step < steps-1)
 * dr = dt (c1 v + c2 f dt/m) + sigma_r n1, v = c0 v + c0 c2 f dt/m / c1 +
 * sigma_v (crv n1 + sqrt(1 - crv^2) n2) with sigma_v^2 = kB T/m varv and
 * sigma_r^2 = kB T/m varr / gamma^2 (md.cpp:1003-1041). noise as written
 * by orc_langevin_noise. Below gamma dt = 1e-5 the VelocityVerlet update,
 * with the Berendsen thermostat when tau > 0. */
int orc_md_langevin_pd(const orc_system *s, double *xyz, double *vel, const orc_params *p, int steps, double dt,
                       int update_nlist, double temperature, double gamma, double tau, const double *noise,
                       const int8_t *hydrogen, double *trace_epot, double *trace_ekin) {
    double c0, c1, c2, varv, varr, crv;
    /* PD's fallback is applyForces_VelocityVerletIntegration, which applies
     * the thermostat (md.cpp:816); the Langevin update itself does not */
    if (!orc_langevin_coef(gamma, dt, &c0, &c1, &c2, &varv, &varr, &crv))
        return orc_md_vv_pd(s, xyz, vel, p, steps, dt, update_nlist, tau, temperature, trace_epot, trace_ekin);
    const double gt = gamma * dt;
    const int n = s->n;
    orc_nlist nl;
    nl.row_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    nl.cap = (int64_t)n * 64 + 1024;
    nl.entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl.cap);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_energies e;
    int rc = steps;
    for (int step = 0; step < steps; ++step) {
        if (update_nlist > 0 && step % update_nlist == 0) build_list(s, xyz, p, &nl);
        eval_forces(s, xyz, &nl, p, f, &e);
        if (step > 0)
            for (int i = 0; i < n; ++i) {
                if (hydrogen && hydrogen[i]) { /* plain VV half kick (md.cpp:1125-1136) */
                    const double th = 0.5 * dt / s->mass[i];
                    for (int c = 0; c < 3; ++c) vel[3 * i + c] += f[3 * i + c] * th;
                    continue;
                }
                const double tinvmass = dt / s->mass[i];
                for (int c = 0; c < 3; ++c) vel[3 * i + c] += (1.0 - c0 / c1) * f[3 * i + c] * tinvmass / gt;
            }
        double ekin = 0;
        for (int i = 0; i < n; ++i)
            ekin += 0.5 * s->mass[i] * sqr(sqrt(sqr(vel[3 * i]) + sqr(vel[3 * i + 1]) + sqr(vel[3 * i + 2])));
        if (step < steps - 1)
            for (int i = 0; i < n; ++i) {
                if (hydrogen && hydrogen[i]) { /* md.cpp:1163-1180 */
                    const double th = 0.5 * dt / s->mass[i];
                    for (int c = 0; c < 3; ++c) {
                        const double dv = f[3 * i + c] * th;
                        double dr = dt * (vel[3 * i + c] + dv);
                        dr *= 1 / kAngstrom;
                        xyz[3 * i + c] += dr;
                        vel[3 * i + c] += dv;
                    }
                    continue;
                }
                const double tinvmass = dt / s->mass[i];
                const double kTdivm = kKB * temperature / s->mass[i];
                const double sv = sqrt(kTdivm * varv), sr = sqrt(kTdivm * varr) / gamma;
                const double *nz = noise + ((size_t)step * n + i) * 6;
                for (int c = 0; c < 3; ++c) {
                    const double n1 = nz[c], n2 = nz[3 + c];
                    const double dlv = sv * (crv * n1 + sqrt(1.0 - sqr(crv)) * n2);
                    const double dlr = sr * n1;
                    double dr = dt * (c1 * vel[3 * i + c] + c2 * f[3 * i + c] * tinvmass) + dlr;
                    dr *= 1 / kAngstrom;
                    const double dv = (c0 * c2) * f[3 * i + c] * tinvmass / c1 + dlv;
                    xyz[3 * i + c] += dr;
                    vel[3 * i + c] *= c0;
                    vel[3 * i + c] += dv;
                }
            }
        if (trace_epot) trace_epot[step] = e.epot;
        if (trace_ekin) trace_ekin[step] = ekin;
        if (!isfinite(e.epot)) { rc = -1; break; }
    }
    free(nl.row_start);
    free(nl.entries);
    free(f);
    return rc;
}

/* ---- A11 Beeman integrator: protocols/md.cpp:840-904 ---- */

/* MolecularDynamics::run_core with Integrator = Beeman (PD's default
 * integrator), no thermostat: per step, list refresh, forces, then
 * (step > 0) v += 2 f dt/6m; ekin; (step < steps-1) p += dt (v + (4 f -
 * f_old) dt/6m) / A, v += (5 f - f_old) dt/6m, f_old = f. f_old (N,
 * natoms*3) is the workspace's old forces and is updated in place. */
int orc_md_beeman_pd(const orc_system *s, double *xyz, double *vel, double *f_old, const orc_params *p, int steps,
                     double dt, int update_nlist, double tau, double temperature, double *trace_epot,
                     double *trace_ekin) {
    const int n = s->n;
    orc_nlist nl;
    nl.row_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    nl.cap = (int64_t)n * 64 + 1024;
    nl.entries = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nl.cap);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    orc_energies e;
    int rc = steps;
    for (int step = 0; step < steps; ++step) {
        if (update_nlist > 0 && step % update_nlist == 0) build_list(s, xyz, p, &nl);
        eval_forces(s, xyz, &nl, p, f, &e);
        if (step > 0)
            for (int i = 0; i < n; ++i) {
                const double tinvmass = dt / (6.0 * s->mass[i]);
                for (int c = 0; c < 3; ++c) vel[3 * i + c] += (2.0 * f[3 * i + c]) * tinvmass;
            }
        double ekin = 0;
        for (int i = 0; i < n; ++i)
            ekin += 0.5 * s->mass[i] * sqr(sqrt(sqr(vel[3 * i]) + sqr(vel[3 * i + 1]) + sqr(vel[3 * i + 2])));
        berendsen(n, vel, ekin, dt, tau, temperature); /* applyThermostat, md.cpp:874 */
        if (step < steps - 1)
            for (int i = 0; i < n; ++i) {
                const double tinvmass = dt / (6.0 * s->mass[i]);
                for (int c = 0; c < 3; ++c) {
                    const double fc = f[3 * i + c], fo = f_old[3 * i + c];
                    double dr = dt * (vel[3 * i + c] + (4.0 * fc - fo) * tinvmass);
                    dr *= 1 / kAngstrom;
                    const double dv = (5.0 * fc - fo) * tinvmass;
                    xyz[3 * i + c] += dr;
                    vel[3 * i + c] += dv;
                    f_old[3 * i + c] = fc;
                }
            }
        if (trace_epot) trace_epot[step] = e.epot;
        if (trace_ekin) trace_ekin[step] = ekin;
        if (!isfinite(e.epot)) { rc = -1; break; }
    }
    free(nl.row_start);
    free(nl.entries);
    free(f);
    return rc;
}
