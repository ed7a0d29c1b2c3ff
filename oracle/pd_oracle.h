/*
 * pd_oracle.h -- CPU FP64 restatement of PD's per-step forcefield hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libpdgpu.so) never links, loads or calls it.
 *
 * Units follow PD exactly: positions in Angstrom, charges in e, LJ radius is
 * R_min/2 in Angstrom, epsilon is sqrt(J/molecule) (PD stores
 * sqrt(eps_kcal * kcal2J / Na), forcefields/ffparam.cpp:1125), energies in
 * J/molecule, forces in Newton, masses in kg.
 *
 * Each function cites the reference lines it restates. Parity status of
 * every function is listed in the header comment of pd_oracle.c.
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Forcefield knobs of FF_GeneralizedBorn_Still / FF_NonBonded
 * (forcefields/gbff.h:127-133, forcefields/nonbonded.h:55-126). */
typedef struct orc_params {
    double cutoff, inner_cutoff;          /* Cutoff, InnerCutoff            */
    double vdw_cutoff, vdw_inner_cutoff;  /* VdwCutoff, VdwInnerCutoff      */
    double vdw14_scaling, elec14_scaling; /* Vdw14Scaling, Elec14Scaling    */
    double dielectric_solute;             /* DielectricSolute (=Dielectric) */
    double dielectric_solvent;            /* DielectricSolvent              */
    double still_cutoff;                  /* GbsaStillCutoff                */
    double exp_approx_threshold;          /* ExpApproxThreshold             */
    double dielectric_offset;             /* DielectricOffset               */
    double born_radius_offset;            /* BornRadiusOffset               */
    double P1, P2, P3, P4, P5;            /* Still parameters               */
    double padding;                       /* NeighbourList Padding          */
    int elec_mode;  /* energy path: 0 none, 1 normal, 3 energy switch, 4 force switch */
    int vdw_mode;   /* energy path: 0 none, 2 energy switch                            */
} orc_params;

/* Components of PD's Hamiltonian touched by the path (workspace/hamiltonian.h). */
typedef struct orc_energies {
    double epot, epot_vdw, epot_elec, epot_pol, epot_pol_self, epot_pol_cross;
    double virial; /* InternalVirial */
    double epot_bond, epot_angle;
    double epot_torsion;  /* torsions + impropers, as PD's Hamiltonian holds them */
    double epot_improper; /* FF_Bonded::getEImproper                              */
} orc_energies;

/* FF_Bonded's assembled term lists (forcefields/ffbonded.h:22-72), atom
 * indices local to the system. Bond: E = k(d-l)^2, k in J/molecule/A^2.
 * Angle: E = k(theta-theta0)^2, k in J/molecule/rad^2. Dihedral i-a-b-j,
 * up to 4 terms: E = Vn(1+cos(n phi+gamma)) for n > 0, Vn (phi-gamma)^2
 * (difference wrapped to [-pi,pi]) for n == 0; phi is the IUPAC dihedral. */
typedef struct orc_bonded {
    int nbonds, nangles, ntorsions, nimpropers;
    const int32_t *bond_ij;      /* 2 per bond        */
    const double *bond_lk;       /* 2 per bond: l, k  */
    const int32_t *angle_iaj;    /* 3 per angle, apex a in the middle */
    const double *angle_tk;      /* 2 per angle: theta0, k */
    const int32_t *tors_iabj;    /* 4 per torsion     */
    const int32_t *tors_terms;   /* 1..4              */
    const double *tors_vn, *tors_n, *tors_gamma; /* 4 per torsion */
    const int32_t *impr_iabj;
    const int32_t *impr_terms;
    const double *impr_vn, *impr_n, *impr_gamma;
} orc_bonded;

/* Per-instance topology and parameters. */
typedef struct orc_system {
    int n;
    const double *charge;    /* e                                      */
    const double *radius;    /* R_min/2, Angstrom                      */
    const double *epsilon;   /* sqrt(J/molecule)                       */
    const double *gb_radius; /* Still solvation type radius, Angstrom  */
    const double *vj;        /* Still volume V_j                       */
    const double *terms123;  /* Still fixed terms                      */
    const double *mass;      /* kg (integrators only)                  */
    int mid;                 /* BondOrder max index delta              */
    const int8_t *band;      /* mid*(n-1) band of bond orders          */
    int nsparse;             /* off-diagonal bond orders               */
    const int32_t *sp_i, *sp_j, *sp_o;
    const orc_bonded *bonded; /* FF_Bonded evaluated beside GB by the step
                                 drivers (orc_minimise_sd, orc_md_vv); NULL: GB only */
} orc_system;

/* Packed neighbour list: row_start[n+1], entries j | (bondorder&7)<<24. */
typedef struct orc_nlist {
    int32_t *row_start;
    uint32_t *entries;
    int64_t cap;
    int64_t count;
} orc_nlist;

/* Physical constants of PD (forcefields/forcefield.cpp:18-38). */
double orc_const(int which);
enum { ORC_KCAL2J = 0, ORC_NA, ORC_ANGSTROM, ORC_ECONV_J, ORC_HALFECONV_J,
       ORC_ECONV, ORC_HALFECONV, ORC_KB, ORC_AMU };

void orc_default_params(orc_params *p);

/* A2 BondOrder::reinit. Pairs are undirected (a,b); resid gives residue of
 * each atom. Call with band == NULL to obtain *mid only. Returns nsparse. */
int orc_bondorder(int n, const int32_t *resid,
                  int n12, const int32_t *a12, const int32_t *b12,
                  int n13, const int32_t *a13, const int32_t *b13,
                  int n14, const int32_t *a14, const int32_t *b14,
                  int32_t *mid, int8_t *band,
                  int32_t *sp_i, int32_t *sp_j, int32_t *sp_o);

/* Bond order PD's neighbour list would record for (i,j). */
int orc_pair_bondorder(const orc_system *s, int i, int j);

/* A1 NeighbourList::calcNewList for InfiniteSpace. list_cutoff is the
 * list's Cutoff (max requested); padding added inside. Returns 0, or the
 * number of entries needed (>cap) when cap is too small. */
int64_t orc_nlist_build(const orc_system *s, const double *xyz, double list_cutoff,
                        double padding, int shadow, orc_nlist *nl);

/* A3 calcFixedBornRadiiTerms. gb_radius = per atom raw type radius. */
void orc_fixed_terms(int n, const double *gb_radius,
                     int n12, const int32_t *a12, const int32_t *b12, const double *d12,
                     int n13, const int32_t *a13, const int32_t *b13, const double *d13,
                     const orc_params *p, double *vj, double *terms123);

/* A4 calcBornRadii_PairwiseApprox. */
void orc_born_radii(const orc_system *s, const double *xyz, const orc_nlist *nl,
                    const orc_params *p, double *born);

/* A7 force path: A4 + A5 + A6. Forces (N) are ADDED to force (may be NULL),
 * energies are ADDED into e the way the reference adds into wspace.ene.
 * born/deda outputs may be NULL. */
void orc_gb_calc_forces(const orc_system *s, const double *xyz, const orc_nlist *nl,
                        const orc_params *p, double *force, orc_energies *e,
                        double *born, double *deda);

/* A8 FF_NonBonded_CalcForces_T_fast2 (energy/force, FPU sqrt, non-verbose). */
/* FF_SoftVDW::calcForces_Soft (forcefields/ffsoftvdw.cpp:71-145): soft-core
 * repulsion over list pairs of bond order above 1-4, hardness in kcal/mol,
 * olap and vdw_cutoff in A. Adds to force (N) and e->epot; sets
 * e->epot_vdw to the soft energy, as PD does. */
void orc_softvdw_forces(const orc_system *s, const double *xyz, const orc_nlist *nl, double hardness,
                        double olap, double vdw_cutoff, double *force, orc_energies *e);

void orc_nb_calc_forces(const orc_system *s, const double *xyz, const orc_nlist *nl,
                        const orc_params *p, double *force, orc_energies *e);

/* A7 energy path: A4, then A8 with the params' modes, then calcBornEnergy. */
void orc_gb_calc_energies(const orc_system *s, const double *xyz, const orc_nlist *nl,
                          const orc_params *p, double *force, orc_energies *e,
                          double *born);

/* FF_Bonded::calcForces (forcefields/ffbonded.cpp:809-824): bonds, angles,
 * torsions, impropers. Forces (N) are ADDED to force (may be NULL); the
 * four component energies (J/molecule) are ADDED to e->epot_bond, ...,
 * e->epot_torsion (torsions and impropers), e->epot_improper, their sum to
 * e->epot, and the bond virial to e->virial. */
void orc_bonded_forces(const orc_bonded *b, const double *xyz, double *force, orc_energies *e);

/* A11 Minimisation::runcore with SteepestDescent, UpdateNList=update_nlist.
 * xyz is updated in place; returns steps done or -1 on NaN. energies/forces
 * of the final structure are written to e_final / force_final (nullable).
 * trace_e (nullable, length steps) receives ene.epot at each step. */
int orc_minimise_sd(const orc_system *s, double *xyz, const orc_params *p,
                    int steps, double step_size, int update_nlist,
                    double *trace_e, orc_energies *e_final, double *force_final);

/* Velocity Verlet (textbook form): half kick, drift, recompute forces, half
 * kick. vel in m/s, dt in s; nlist rebuilt every update_nlist steps. */
int orc_md_vv(const orc_system *s, double *xyz, double *vel, const orc_params *p,
              int steps, double dt, int update_nlist, double *trace_epot, double *trace_ekin);

/* A11 MolecularDynamics::run_core with PD's VelocityVerlet update as written
 * (protocols/md.cpp:209-233, 788-838): per step, list refresh when
 * step % update_nlist == 0, forces, then v += f*dt/2m three times with the
 * drift dt*(v + f*dt/2m) before the third. trace_ekin receives ene.ekin of
 * each step (taken after the first addition). tau > 0: the Berendsen
 * thermostat (applyBerendsenThermostat, md.cpp:580-597) at temperature K
 * after the kinetic energy, as PD's applyThermostat (md.cpp:816). */
int orc_md_vv_pd(const orc_system *s, double *xyz, double *vel, const orc_params *p,
                 int steps, double dt, int update_nlist, double tau, double temperature,
                 double *trace_epot, double *trace_ekin);

/* A11 Langevin (protocols/md.cpp:906-1043): coefficients for gamma [1/s]
 * and dt [s]; returns 0 (VelocityVerlet), 1 (exponential form) or 2
 * (Taylor form). */
int orc_langevin_coef(double gamma, double dt, double *c0, double *c1, double *c2, double *varv, double *varr,
                      double *crv);

/* The reference's normal deviates after srand(seed): [steps][natoms][6] as
 * (n1x, n1y, n1z, n2x, n2y, n2z), last step and skipped atoms zero. */
void orc_langevin_noise(unsigned seed, int natoms, int steps, const int8_t *skip, double *out);

/* MolecularDynamics::run_core, Integrator = Langevin, constant target
 * temperature [K], friction gamma [1/s], noise as from orc_langevin_noise;
 * hydrogen NULL: LangevinOnHydrogens, else those atoms on plain kicks.
 * tau (Berendsen, s) only acts in the VelocityVerlet fallback below
 * gamma dt = 1e-5, as in PD. */
int orc_md_langevin_pd(const orc_system *s, double *xyz, double *vel, const orc_params *p, int steps, double dt,
                       int update_nlist, double temperature, double gamma, double tau, const double *noise,
                       const int8_t *hydrogen, double *trace_epot, double *trace_ekin);

/* MolecularDynamics::run_core, Integrator = Beeman; f_old (N, natoms*3)
 * in/out; tau > 0: Berendsen thermostat as for orc_md_vv_pd. */
int orc_md_beeman_pd(const orc_system *s, double *xyz, double *vel, double *f_old, const orc_params *p, int steps,
                     double dt, int update_nlist, double tau, double temperature, double *trace_epot,
                     double *trace_ekin);

#ifdef __cplusplus
}
#endif
#endif
