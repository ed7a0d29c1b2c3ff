/*
 * pdgpu.h -- C ABI of the MI355X (gfx950) implementation of PD's per-step
 * forcefield evaluation: the Still Generalized Born Born-radius pass, the
 * fused LJ + Coulomb + GB pair pass with its chain rule, PD's energy-only
 * path, and the neighbour-list rebuild that feeds them.
 *
 * The interface replaced is Physics::ForcefieldBase as implemented by
 * FF_GeneralizedBorn_Still (reference: src/mmlib/forcefields/forcefield.h:93-222,
 * forcefields/gbff.h:120-200) together with the NeighbourList it reads
 * (workspace/neighbourlist.h:187-198). Each entry point below names the
 * reference member it stands in for.
 *
 * Conventions
 *  - Plain C types only; no exceptions cross this boundary. Every function
 *    returns PDGPU_OK (0) or a negative PDGPU_ERR_* code; the message of the
 *    last failure is available from pdgpu_last_error().
 *  - Units at the boundary are PD's (forcefields/forcefield.cpp:18-38):
 *    positions in Angstrom, charges in e, LJ radius R_min/2 in Angstrom,
 *    epsilon as sqrt(J/molecule) (how PD's Particle stores it,
 *    forcefields/ffparam.cpp:1125), energies in J/molecule, forces in
 *    Newton, masses in kg. Arrays of vectors are AoS: x0,y0,z0,x1,...
 *  - Input arrays are copied; the caller keeps ownership.
 *  - A context owns one GPU stream; calls are stream-ordered and return
 *    after the results they produce are on the host. A context belongs to
 *    one host thread.
 *  - Instances (independent molecules) live resident on the device and are
 *    evaluated together when PDGPU_ALL is passed: that is the batched path.
 */
#ifndef PDGPU_H
#define PDGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDGPU_OK 0
#define PDGPU_ERR_INVALID (-1)     /* bad argument or state                      */
#define PDGPU_ERR_NOMEM (-2)       /* device memory budget or allocation failure */
#define PDGPU_ERR_NONFINITE (-3)   /* non-finite or out-of-range result          */
#define PDGPU_ERR_HIP (-4)         /* HIP runtime error                          */
#define PDGPU_ERR_UNSUPPORTED (-5) /* option the implementation does not offer  */
#define PDGPU_ERR_NODEVICE (-6)    /* no usable gfx950 device                    */

#define PDGPU_ALL (-1)

enum pdgpu_precision {
    PDGPU_FP64 = 0,  /* FP64 throughout: the plain method                         */
    PDGPU_MIXED = 1  /* FP32 pair maths, 64-bit fixed-point force accumulation,   */
                     /* FP64 energy reductions                                    */
};

/* Forcefield knobs, mirrored one to one from the public fields of
 * FF_GeneralizedBorn_Still / FF_NonBonded (forcefields/gbff.h:127-133,
 * forcefields/nonbonded.h:55-126), plus implementation settings. */
typedef struct pdgpu_options {
    int32_t precision;            /* pdgpu_precision                               */
    int32_t device;               /* HIP device ordinal                            */
    int32_t fast_mode;            /* FastMode; only 1 (the reference default)      */
    int32_t elec_mode;            /* energy path electrostatics, PD T_ElecMode:    */
                                  /* 0 none, 1 plain, 3 energy switch, 4 force sw. */
    int32_t vdw_mode;             /* energy path vdW: 0 none, 2 energy switch      */
    int32_t tiles_per_wave;       /* 0: choose automatically                       */
    double cutoff;                /* Cutoff                                        */
    double inner_cutoff;          /* InnerCutoff                                   */
    double vdw_cutoff;            /* VdwCutoff                                     */
    double vdw_inner_cutoff;      /* VdwInnerCutoff                                */
    double vdw14_scaling;         /* Vdw14Scaling                                  */
    double elec14_scaling;        /* Elec14Scaling                                 */
    double dielectric_solute;     /* DielectricSolute                              */
    double dielectric_solvent;    /* DielectricSolvent                             */
    double gbsa_still_cutoff;     /* GbsaStillCutoff                               */
    double exp_approx_threshold;  /* ExpApproxThreshold                            */
    double dielectric_offset;     /* DielectricOffset                              */
    double born_radius_offset;    /* BornRadiusOffset                              */
    double p1, p2, p3, p4, p5;    /* Still P1..P5                                  */
    double nlist_padding;         /* NeighbourList Padding (skin)                  */
    int64_t device_memory_budget; /* bytes the context may allocate; 0: 80% free  */
    int32_t max_instances;        /* instance slots; 0: 8192                       */
    int32_t components;           /* PDGPU_FF_* bits evaluated; default all        */
    int32_t use_graphs;           /* 1: step loops replay captured HIP graphs;     */
                                  /* default 0 (measured no faster)                */
    int32_t dense_tile_pairs;     /* off-diagonal tiles with at least this many    */
                                  /* candidate pairs are evaluated densely (lane   */
                                  /* rotation, no lists); 0: 3584 x (list radius / */
                                  /* cutoff)^3, i.e. ~7/8 of the slots in cutoff   */
    int32_t langevin_on_hydrogens; /* LangevinOnHydrogens (protocols/md.h);      */
                                  /* default 1; 0: hydrogens flagged with        */
                                  /* pdgpu_set_hydrogens move by plain kicks     */
    /* FF_SoftVDW (forcefields/ffsoftvdw.h:39-47), used with PDGPU_FF_SOFTVDW  */
    double softvdw_hardness;      /* Hardness, kcal/mol (default 12)               */
    double softvdw_olap_offset;   /* OlapOffset, A (default 0.6)                   */
    double softvdw_cutoff;        /* VdwCutoff of FF_SoftVDW, A (default 15)       */
    int32_t tile_split;           /* wavefronts per block-pair tile: 0 or 1 (one), */
                                  /* 2, 4, 8 or 16. Each takes its own share of    */
                                  /* the tile's candidates / rotation steps and    */
                                  /* flushes its own sums, so results are          */
                                  /* deterministic for a given value but differ in */
                                  /* rounding between values. For single systems   */
                                  /* of a few thousand atoms, whose few tiles      */
                                  /* leave the GPU idle behind the longest one;    */
                                  /* batches run best at 1. Not PD's (no knob).    */
} pdgpu_options;

/* Forcefield components of a context: FF_GeneralizedBorn_Still and the four
 * term kinds of FF_Bonded (its DoBonds / DoAngles / DoTorsions / DoImpropers,
 * forcefields/ffbonded.h:147-150). Bonded terms contribute only for
 * instances given term lists with pdgpu_bonded_set. */
#define PDGPU_FF_GB 1
#define PDGPU_FF_BONDS 2
#define PDGPU_FF_ANGLES 4
#define PDGPU_FF_TORSIONS 8
#define PDGPU_FF_IMPROPERS 16
#define PDGPU_FF_BONDED (PDGPU_FF_BONDS | PDGPU_FF_ANGLES | PDGPU_FF_TORSIONS | PDGPU_FF_IMPROPERS)
#define PDGPU_FF_ALL (PDGPU_FF_GB | PDGPU_FF_BONDED)
/* FF_SoftVDW's soft-core repulsion (calcForces_Soft, forcefields/
 * ffsoftvdw.cpp:71-145), PD's clash-relief term; not part of PDGPU_FF_ALL.
 * Like PD's, its energy is reported as epot_vdw. It cannot be combined with
 * PDGPU_FF_GB in one context (PD's FF_SoftVDW resets ene.epot_vdw, so the
 * pair depends on insertion order): ctx_create returns
 * PDGPU_ERR_UNSUPPORTED. The soft pairs (r < R_i + R_j - OlapOffset) must
 * lie within Cutoff, which holds for any Cutoff above ~5 A. */
#define PDGPU_FF_SOFTVDW 32
/* FF_NonBonded on its own (forcefields/nonbonded.cpp:1137-1229, the
 * FF_NonBonded_CalcForces_T_fast2 pair loop :542-893): vacuum LJ +
 * Coulomb over the cutoff list with the context's elec_mode / vdw_mode
 * (PD's FF_NonBonded default is the force switch, elec_mode 4, see
 * pdgpu_options_default_nonbonded), energies in epot_vdw / epot_elec, its
 * InternalVirial in virial. calcForces and calcEnergies are the same full
 * evaluation (PD's UsePartialRecalc is off by default). Not part of
 * PDGPU_FF_ALL; combines with the FF_Bonded terms, not with PDGPU_FF_GB
 * (which holds its own vacuum terms) or PDGPU_FF_SOFTVDW
 * (PDGPU_ERR_UNSUPPORTED). */
#define PDGPU_FF_NONBONDED 64

/* One molecule. Per-atom arrays have natoms entries. The bond-order band and
 * the sparse list are PD's BondOrder (workspace/bondorder.h:43-112): band has
 * max_index_delta*(natoms-1) entries, band[i*mid + (j-i-1)] = order of (i,j)
 * for j>i; sparse holds pairs (i<j) beyond the band. */
typedef struct pdgpu_topology {
    int32_t natoms;
    int32_t max_index_delta;
    const double *charge;    /* e                                        */
    const double *radius;    /* R_min/2, Angstrom                        */
    const double *epsilon;   /* sqrt(J/molecule)                         */
    const double *gb_radius; /* Still solvation type radius, Angstrom    */
    const double *vj;        /* Still V_j (pdgpu_fixed_terms)            */
    const double *terms123;  /* Still fixed terms (pdgpu_fixed_terms)    */
    const double *mass;      /* kg; may be NULL when no integrator used  */
    const int8_t *band;
    int32_t nsparse;
    int32_t reserved;
    const int32_t *sparse_i, *sparse_j, *sparse_order;
} pdgpu_topology;

/* FF_Bonded's assembled term lists for one instance (forcefields/ffbonded.h:
 * 22-72, as FF_Bonded::setup builds them, forcefields/ffbonded.cpp:13-486),
 * atom indices local to the instance. Bond: E = k (d-l)^2 with l in
 * Angstrom, k in J/molecule/A^2. Angle i-a-j (apex a): E = k (theta-theta0)^2,
 * theta0 in rad, k in J/molecule/rad^2. Torsion / improper i-a-b-j with
 * 1..4 terms: E = Vn (1 + cos(n phi + gamma)) for n > 0, Vn (phi-gamma)^2
 * (difference wrapped to [-pi, pi]) for n == 0; phi is the IUPAC dihedral,
 * Vn in J/molecule, gamma in rad; 4 entries per dihedral in vn/n/gamma. */
typedef struct pdgpu_bonded {
    int32_t nbonds, nangles, ntorsions, nimpropers;
    const int32_t *bond_ij;        /* 2 per bond                    */
    const double *bond_lk;         /* 2 per bond: l, k              */
    const int32_t *angle_iaj;      /* 3 per angle                   */
    const double *angle_tk;        /* 2 per angle: theta0, k        */
    const int32_t *torsion_iabj;   /* 4 per torsion                 */
    const int32_t *torsion_terms;  /* 1 per torsion                 */
    const double *torsion_vn, *torsion_n, *torsion_gamma;
    const int32_t *improper_iabj;
    const int32_t *improper_terms;
    const double *improper_vn, *improper_n, *improper_gamma;
} pdgpu_bonded;

/* Energy components (workspace/hamiltonian.h:23-90), J/molecule. */
typedef struct pdgpu_energies {
    double epot, epot_vdw, epot_elec, epot_pol, epot_pol_self, epot_pol_cross;
    double virial;          /* InternalVirial: GB energy path + bond stretching  */
    double epot_bond, epot_angle;
    double epot_torsion;    /* torsions + impropers, as the Hamiltonian holds it */
    double epot_improper;   /* FF_Bonded::getEImproper                           */
    int64_t pairs;          /* unordered pairs within Cutoff (pair pass)         */
    int64_t born_pairs;     /* ordered pairs in the Born/chain-rule sums         */
    int32_t status;         /* PDGPU_OK or PDGPU_ERR_NONFINITE for this instance */
    int32_t reserved;
} pdgpu_energies;

typedef struct pdgpu_ctx pdgpu_ctx;

/* Fills *o with the defaults of a new FF_GeneralizedBorn_Still
 * (forcefields/gbff.cpp:17-37, forcefields/nonbonded.cpp:968-1004). */
void pdgpu_options_default(pdgpu_options *o);

/* Fills *o with the defaults of a new FF_NonBonded (FF_NonBonded::
 * settodefault, forcefields/nonbonded.cpp:968-1004: Cutoff 15, InnerCutoff
 * 12, VdwCutoff 15, VdwInnerCutoff 12, force switching, Dielectric 1) and
 * components = PDGPU_FF_NONBONDED | PDGPU_FF_BONDED. The 1-4 scalings are
 * 1.0 as constructed; PD's AutoScaling14 copies the parameter file's values
 * (0.5, 0.83333 for amber99) at setup, which the caller sets here. */
void pdgpu_options_default_nonbonded(pdgpu_options *o);

/* Creation of the component and its device state; stands in for the
 * constructor plus ForcefieldBase::ensuresetup (forcefields/forcefield.h:143). */
int pdgpu_ctx_create(const pdgpu_options *o, pdgpu_ctx **out);
void pdgpu_ctx_destroy(pdgpu_ctx *ctx);
const char *pdgpu_last_error(const pdgpu_ctx *ctx);

/* Setup of one molecule: FF_GeneralizedBorn_Still::setup's parameter copy
 * (forcefields/gbff.cpp:40-87) and WorkSpace::bondorder() upload. */
int pdgpu_instance_add(pdgpu_ctx *ctx, const pdgpu_topology *t, int32_t *id);
int pdgpu_instance_remove(pdgpu_ctx *ctx, int32_t id);

/* FF_Bonded::setup for one instance: uploads its term lists (replacing any
 * earlier ones). Energy units J/molecule as described at pdgpu_bonded. */
int pdgpu_bonded_set(pdgpu_ctx *ctx, int32_t id, const pdgpu_bonded *b);

/* Device bytes the records of pdgpu_bonded_set take (no reference
 * counterpart; for admission by memory budget beside pdgpu_instance_bytes). */
int64_t pdgpu_bonded_bytes(int32_t nbonds, int32_t nangles, int32_t ntorsions, int32_t nimpropers);
int pdgpu_instance_natoms(const pdgpu_ctx *ctx, int32_t id);

/* WorkSpace::cur.atom[].p in / out. */
int pdgpu_set_positions(pdgpu_ctx *ctx, int32_t id, const double *xyz);
int pdgpu_get_positions(pdgpu_ctx *ctx, int32_t id, double *xyz);

/* NeighbourList::calcNewList (workspace/neighbourlist.cpp:205-227): rebuilds
 * the device pair-tile list of one instance or all (PDGPU_ALL). */
int pdgpu_nlist_build(pdgpu_ctx *ctx, int32_t id);

/* NeighbourList::getData() (workspace/neighbourlist.h:46): writes PD's packed
 * list, entries j | bondorder<<24, rows in ascending j, row_start[natoms+1].
 * shadow=1 gives the full list, 0 only j<i. When capacity is too small,
 * nothing is written past it, *count is the size needed and PDGPU_ERR_NOMEM
 * is returned. */
int pdgpu_nlist_export(pdgpu_ctx *ctx, int32_t id, int32_t shadow, int32_t *row_start,
                       uint32_t *entries, int64_t capacity, int64_t *count);

/* The same export with no device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. The
 * per-atom counts, row offsets and entries are staged in one static 64 MiB
 * device array per GPU instead of hipMalloc'd buffers; instances whose
 * staging (about 20 B per atom + 4 B per entry) exceeds it get
 * PDGPU_ERR_NOMEM with a message, as does a too small capacity (*count is
 * then the size needed). The instance's list must already exist
 * (pdgpu_nlist_build or an evaluation), else PDGPU_ERR_INVALID. All calls
 * on one GPU share the array: calls that overlap on one GPU race. Output
 * is bit-identical to pdgpu_nlist_export. */
int pdgpu_nlist_export_noalloc(pdgpu_ctx *ctx, int32_t id, int32_t shadow, int32_t *row_start,
                               uint32_t *entries, int64_t capacity, int64_t *count);

/* Forcefield::calcForces over the context's components
 * (forcefields/forcefield.cpp:220-227): FF_Bonded::calcForces
 * (forcefields/ffbonded.cpp:809-824) and FF_GeneralizedBorn_Still::calcForces
 * (forcefields/gbff.cpp:168-180: Born radii, fused pair pass, chain rule).
 * energies: one record for id, or an array indexed by instance id for
 * PDGPU_ALL (may be NULL). forces: natoms*3 Newton for a single id, or NULL
 * to keep them on the device. */
int pdgpu_calc_forces(pdgpu_ctx *ctx, int32_t id, pdgpu_energies *energies, double *forces);

/* Forcefield::calcEnergies (forcefields/forcefield.cpp:212-218):
 * FF_Bonded::calcEnergies (= its calcForces, forcefields/ffbonded.cpp:804)
 * and FF_GeneralizedBorn_Still::calcEnergies (forcefields/gbff.cpp:161-166:
 * Born radii, FF_NonBonded energy path with its forces, as the reference
 * writes them, exact pairwise GB energy). */
int pdgpu_calc_energies(pdgpu_ctx *ctx, int32_t id, pdgpu_energies *energies, double *forces);

/* Device results of the last evaluation: Born radii (Angstrom), forces (N). */
int pdgpu_get_born_radii(pdgpu_ctx *ctx, int32_t id, double *born);
int pdgpu_get_forces(pdgpu_ctx *ctx, int32_t id, double *forces);

/* Minimisation::runcore with SteepestDescent (protocols/minimise.cpp:99-235),
 * run on the device for one instance or all. PD semantics: step p += f *
 * step_size * mult, mult x1.2 on downhill, x0.5/1.2 and restore on uphill.
 * update_nlist: rebuild every that many steps (PD default 1); a negative
 * value rebuilds an instance's list only when one of its atoms has moved
 * more than half the list padding since its last rebuild (the Verlet-skin
 * criterion, an extension PD does not have). grms_tol > 0
 * stops an instance once its force RMS falls below it (kcal/mol/Angstrom),
 * an extension PD does not have. steps_done (per id, may be NULL) receives
 * the steps each instance ran, or -1 if its energy became non-finite. */
int pdgpu_minimise_sd(pdgpu_ctx *ctx, int32_t id, int32_t steps, double step_size,
                      int32_t update_nlist, double grms_tol, int32_t *steps_done);

/* The same minimiser, driven step by step for continuous batching: a
 * server starts a minimisation on an instance as soon as it is added
 * (pdgpu_sd_start), advances every started minimisation together
 * (pdgpu_sd_advance), and collects each one as it finishes
 * (pdgpu_sd_finish: cur = old, final energies and forces; the instance can
 * then be read and removed). Each instance keeps its own step counter,
 * multiplier and list-refresh schedule, so its result does not depend on
 * which other instances share the launches. done_by_id / steps_by_id (may
 * be NULL) are indexed by instance id; *running counts unfinished ones. */
int pdgpu_sd_start(pdgpu_ctx *ctx, int32_t id, int32_t steps, double step_size, int32_t update_nlist,
                   double grms_tol);
int pdgpu_sd_advance(pdgpu_ctx *ctx, int32_t nsteps, int32_t *done_by_id, int32_t *steps_by_id,
                     int32_t *running);
int pdgpu_sd_finish(pdgpu_ctx *ctx, int32_t id, pdgpu_energies *energies, int32_t *steps_done);

/* Velocity Verlet, textbook form (half kick, drift, force, half kick), on the
 * device. vel in/out in m/s (natoms*3, single id; NULL keeps device state),
 * dt in seconds. The neighbour list is rebuilt every update_nlist steps. */
int pdgpu_set_velocities(pdgpu_ctx *ctx, int32_t id, const double *vel);
int pdgpu_get_velocities(pdgpu_ctx *ctx, int32_t id, double *vel);
int pdgpu_md_vv(pdgpu_ctx *ctx, int32_t id, int32_t steps, double dt, int32_t update_nlist,
                pdgpu_energies *energies, double *ekin);

/* MolecularDynamics::runcore with Integrator = VelocityVerlet, as PD writes
 * it (protocols/md.cpp:209-233 step loop, :788-838 update; list refresh
 * protocols/protocolbase.cpp:113-122): each step refreshes the list when
 * step % update_nlist == 0, evaluates forces and applies v += f*dt/2m
 * (ekin taken), v += f*dt/2m, p += dt*(v + f*dt/2m), v += f*dt/2m. No
 * barostat; with pdgpu_set_thermostat(tau > 0) the Berendsen thermostat
 * scales the velocities after each step's kinetic energy, and such runs are
 * limited to 200 steps. energies are those of the last force evaluation and
 * ekin the last step's, as in PD's workspace after run(). */
int pdgpu_md_vv_pd(pdgpu_ctx *ctx, int32_t id, int32_t steps, double dt, int32_t update_nlist,
                   pdgpu_energies *energies, double *ekin);

/* MolecularDynamics::Thermostat = Berendsen for the PD MD calls
 * (pdgpu_md_vv_pd, pdgpu_md_beeman_pd; applyBerendsenThermostat,
 * protocols/md.cpp:580-597, applied after each step's kinetic energy as
 * applyThermostat is, md.cpp:816 and :874): every velocity of an instance
 * times sqrt(1 + dt (3 N kB T / 2 / ekin - 1) / tau). tau in s (PD's
 * BerendsenTau, default 1e-13), temperature in K; tau = 0 turns it off.
 * PD also removes the total momenta every 200 steps while a thermostat is
 * on, for every integrator (md.cpp:246-258); that is not built, so any MD
 * call (pdgpu_md_langevin_pd included) runs at most 200 steps while tau > 0
 * (PDGPU_ERR_UNSUPPORTED beyond). */
int pdgpu_set_thermostat(pdgpu_ctx *ctx, double tau, double temperature);

/* Particle::isHydrogen flags of an instance's atoms (natoms bytes, 0/1;
 * system/fundamentals.h:223), used by the Langevin update when
 * langevin_on_hydrogens is 0 (applyForces_LangevinIntegration_NonHydrogens,
 * protocols/md.cpp:1049-1195: hydrogens take PD's plain half kicks and draw
 * no deviates). Zero when an instance is added. */
int pdgpu_set_hydrogens(pdgpu_ctx *ctx, int32_t id, const uint8_t *is_hydrogen);

/* MolecularDynamics::runcore with Integrator = Langevin and
 * LangevinOnHydrogens (PD's default; protocols/md.cpp:906-1043 update,
 * :209-233 step loop): each step refreshes the list when due, evaluates
 * forces, then (after the first step) v += (1 - c0/c1) f dt/m / (gamma dt),
 * takes ekin, and (before the last step) moves p by dt (c1 v + c2 f dt/m)
 * + sigma_r n1 and sets v = c0 v + c0 c2 f dt/m / c1 + sigma_v (crv n1 +
 * sqrt(1 - crv^2) n2), with PD's coefficients for gamma dt (exponential
 * form above 0.05, its Taylor series below; below 1e-5 the VelocityVerlet
 * update, as PD does). temperature in K (constant), friction gamma in 1/s
 * (PD's FricCoeff, default 1e13). Normal deviates: noise == NULL draws
 * them on the device from Philox4x32-10 keyed by seed over (atom,
 * instance, step counter of the context); a single id may instead pass
 * noise = [steps][natoms][6] doubles (n1x, n1y, n1z, n2x, n2y, n2z), e.g.
 * the reference's libc rand() stream, for a step-for-step reproduction of
 * a PD run. */
int pdgpu_md_langevin_pd(pdgpu_ctx *ctx, int32_t id, int32_t steps, double dt, int32_t update_nlist,
                         double temperature, double friction, uint64_t seed, const double *noise,
                         pdgpu_energies *energies, double *ekin);

/* MolecularDynamics::runcore with Integrator = Beeman (PD's default
 * integrator; protocols/md.cpp:840-904; Berendsen as for pdgpu_md_vv_pd
 * when pdgpu_set_thermostat set tau > 0, at most 200 steps): after the first
 * step v += 2 f dt/6m, ekin taken, and before the last step p += dt (v +
 * (4 f - f_old) dt/6m), v += (5 f - f_old) dt/6m, f_old = f. The old
 * forces (WorkSpace::old in PD) are kept per atom between calls and are
 * zero when an instance is added. */
int pdgpu_md_beeman_pd(pdgpu_ctx *ctx, int32_t id, int32_t steps, double dt, int32_t update_nlist,
                       pdgpu_energies *energies, double *ekin);

/* Result hand-off formats. A function returning int64_t gives the byte
 * count of its output (or a negative pdgpu_status) and writes only when cap
 * is large enough (text outputs also get a terminating NUL, so they need
 * cap > length).
 *
 * SnapShot::writeMIME (workspace/snapshot.cpp:316-348) of arrays: "!PSP
 * name", PD's 6-bit coding (tools/io.cpp:79-163) of a 128-byte header
 * (natoms, epot, four zero status words; PD leaves the remaining 100 bytes
 * uninitialised, here they are zero) followed by p (A), v (m/s), f (N) per
 * atom, then "!ENDPSP". */
int64_t pdgpu_mime_encode(const double *xyz, const double *vel, const double *force, int32_t natoms,
                          double epot, const char *name, char *out, int64_t cap);

/* WorkSpace::save().writeMIME(name) (workspace/workspace.cpp:165-172) of
 * an instance: its device positions, velocities, the forces and total
 * energy of its last evaluation, gathered in one copy. */
int64_t pdgpu_snapshot_mime(pdgpu_ctx *ctx, int32_t id, const char *name, char *out, int64_t cap);

/* OutTra_NAMD's DCD trajectory (fileio/psfdcd.cpp:104-247): the 276-byte
 * header write_DCD_header writes (its title blocks zero-padded where PD
 * writes uninitialised stack bytes), and one append_DCD_step frame of an
 * instance (x, y, z as float records), converted on the device and copied
 * once. pdgpu_dcd_create / pdgpu_dcd_append do the same on a file. */
int64_t pdgpu_dcd_header(int32_t natoms, uint8_t *out, int64_t cap);
int64_t pdgpu_dcd_frame(pdgpu_ctx *ctx, int32_t id, uint8_t *out, int64_t cap);
int pdgpu_dcd_create(const char *path, int32_t natoms);
int pdgpu_dcd_append(pdgpu_ctx *ctx, int32_t id, const char *path);

/* Bristol trajectory format (OutTra_BTF, fileio/tra.cpp:28-180, records of
 * fileio/tratypes.cpp:12-226), the reference's own trajectory file.
 * Includes are PD's BFT_StandardIncludes flags; positions are always
 * written. */
#define PDGPU_BTF_PHIPSI 2  /* per residue (phi, psi), radians, as float    */
#define PDGPU_BTF_ENERGY 8  /* the 64-float energy record (DefaultIncludes) */
#define PDGPU_BTF_FORCE  16 /* per atom p + f * 5e8 (f in N), as float      */

/* One atom of the system definition block (BTF_SystemDefinitionEntry::
 * setFrom, fileio/tratypes.cpp:130-172): the Particle's names (pdbname,
 * type_name, rawname; NUL-terminated, at most the first 8 bytes are kept,
 * zero-padded as strncpy does), its residue's name and index (parentname,
 * ir), the reference position posRef (A) and isKnownStructure, the
 * covalent 1-2 partners in cov12atom order (at most 6), and charge
 * (e), radius (A), epsilon and mass in PD's units. */
typedef struct pdgpu_btf_atom {
    const char *pdbname, *type_name, *rawname, *parentname;
    int32_t ir;
    int32_t known_structure;
    double ref[3];
    int32_t n_cov12;
    int32_t cov12[6];
    double charge, radius, epsilon, mass;
} pdgpu_btf_atom;

/* What a frame's energy record takes from the workspace besides the last
 * evaluation's energies: WorkSpace::Step (written as Step and as time),
 * ene.cRMS, ene.dRMS, ene.ekin and ene.etot (J/molecule) and ene.data1..10
 * (BTF_Energy::save, fileio/tratypes.cpp:194-231). */
typedef struct pdgpu_btf_state {
    int64_t step;
    double crms, drms, ekin, etot;
    double data[10];
} pdgpu_btf_state;

/* The file's start (BTF_Header::save then OutTra_BTF::create,
 * tratypes.cpp:60-84, tra.cpp:28-89): the 18,084-byte header (version 2,
 * Type = 1 | includes, residue and atom counts, the frame size, the offset
 * of the first frame, the date, ID strings "null", PD's property and
 * energy names; text areas zero), "SYSTDEFI", one 220-byte entry per atom,
 * "TRASTART". date_dmy (day, month, year) may be NULL for today's local
 * date. Where PD leaves bytes uninitialised (the ID strings after "null",
 * parentname after its terminator, unused cov12atom slots) this writes
 * zeros. */
int64_t pdgpu_btf_header(const pdgpu_btf_atom *atoms, int32_t natoms, int32_t nresidues, int32_t includes,
                         const int32_t *date_dmy, uint8_t *out, int64_t cap);

/* One OutTra_BTF::append record of an instance (tra.cpp:91-180): "TRAE",
 * its device positions as float x, y, z per atom; with PDGPU_BTF_PHIPSI,
 * (phi, psi) of each residue from its backbone atoms (backbone: N, CA, C
 * instance-local indices per residue, -1 where absent; MoleculeBase::
 * calcResiduePhiPsi, system/molecule.cpp:613-651, 0 where undefined); with
 * PDGPU_BTF_ENERGY, the energy record from the instance's last evaluation
 * (kcal/mol) and *state; with PDGPU_BTF_FORCE, p + f * 5e8 of the last
 * forces. pdgpu_btf_append does the same on a file. */
int64_t pdgpu_btf_frame(pdgpu_ctx *ctx, int32_t id, int32_t includes, const int32_t *backbone, int32_t nresidues,
                        const pdgpu_btf_state *state, uint8_t *out, int64_t cap);
int pdgpu_btf_append(pdgpu_ctx *ctx, int32_t id, int32_t includes, const int32_t *backbone, int32_t nresidues,
                     const pdgpu_btf_state *state, const char *path);

/* MolecularDynamics::UpdateTra with an OutTra_NAMD in the workspace's
 * trajectory container (protocols/md.cpp:247, ProtocolBase::every): with
 * path non-NULL, the PD MD calls (pdgpu_md_vv_pd, _langevin_pd,
 * _beeman_pd) append a DCD frame of instance id after every step whose
 * count (from 0 in each call) is a positive multiple of `every`; the file
 * gets its header with the first frame. Frames are packed on the device
 * and written in batches, at the latest when the MD call returns. path ==
 * NULL detaches; removing the instance detaches too. */
int pdgpu_traj_dcd(pdgpu_ctx *ctx, int32_t id, const char *path, int32_t every);

/* Grid-shape autotuner (no reference counterpart): times the tile kernels
 * of a force evaluation of id (or PDGPU_ALL) reps times at each of several
 * wavefronts-per-CU settings and keeps the fastest per kernel for later
 * evaluations. chosen (may be NULL) receives 4 values: enumeration, Born
 * sum, pair pass, chain rule. Results do not depend on the setting. */
int pdgpu_autotune(pdgpu_ctx *ctx, int32_t id, int32_t reps, int32_t *chosen);

/* Optional timing of the kernel groups below with HIP events recorded on
 * the context's stream around each launch (no reference counterpart; used
 * by the benchmark). pdgpu_profile(ctx, 1) clears and starts recording,
 * (ctx, 0) clears and stops; pdgpu_profile_read sums milliseconds and
 * launch counts per group since the last clear. */
enum pdgpu_kernel_group {
    PDGPU_KERNEL_NLIST = 0,     /* bounding boxes, tile rows, row scan      */
    PDGPU_KERNEL_ATOM = 1,      /* per-atom prepare / Born finalise / forces */
    PDGPU_KERNEL_BORN = 2,      /* Born-radius pair sum                      */
    PDGPU_KERNEL_PAIR = 3,      /* fused LJ + Coulomb + GB pair pass         */
    PDGPU_KERNEL_CHAIN = 4,     /* chain rule through the Born radii         */
    PDGPU_KERNEL_INTEGRATE = 5, /* integrator / minimiser updates            */
    PDGPU_KERNEL_ENUM = 6,      /* per-evaluation pair enumeration           */
    PDGPU_KERNEL_BONDED = 7,    /* FF_Bonded terms                           */
    PDGPU_KERNEL_NGROUPS = 8
};
int pdgpu_profile(pdgpu_ctx *ctx, int32_t enable);
int pdgpu_profile_read(pdgpu_ctx *ctx, double *ms, int64_t *launches, int32_t n);

/* Host-side setup helpers (no device work).
 * BondOrder::reinit (workspace/bondorder.cpp:170-195) from undirected 1-2,
 * 1-3, 1-4 pair lists and residue numbers. Call with band == NULL to obtain
 * *max_index_delta (and size band as mid*(natoms-1)); sparse arrays need
 * n12+n13+n14 entries. */
int pdgpu_bondorder_build(int32_t natoms, const int32_t *resid,
                          int32_t n12, const int32_t *a12, const int32_t *b12,
                          int32_t n13, const int32_t *a13, const int32_t *b13,
                          int32_t n14, const int32_t *a14, const int32_t *b14,
                          int32_t *max_index_delta, int8_t *band, int32_t *nsparse,
                          int32_t *sparse_i, int32_t *sparse_j, int32_t *sparse_order);

/* calcFixedBornRadiiTerms (forcefields/gbff.cpp:307-381): V_j and terms123
 * from type radii and the equilibrium 1-2 / 1-3 distances. */
int pdgpu_fixed_terms(int32_t natoms, const double *gb_radius,
                      int32_t n12, const int32_t *a12, const int32_t *b12, const double *d12,
                      int32_t n13, const int32_t *a13, const int32_t *b13, const double *d13,
                      const pdgpu_options *o, double *vj, double *terms123);

/* Input preparation (host only, no device work): what PD does before a
 * forcefield sees a molecule.
 *
 * pdgpu_ff_read reads a PD parameter file (FFParamSet::readLib,
 * forcefields/ffparam.cpp:1719-2200: TYPE, BOND, ANGLE, TORSION, IMPROPER,
 * MOLECULE blocks with LOAD / DELATOM / ATOM_APPEND_MODE / CONNECT /
 * CHARGE / IMPROPER, ALIAS, CLASS / CLASS_PREFIX, the 1-4 scalings,
 * INCLUDE, and the SOLVATION_GENERALISED_BORN section's GENERALPARAMETERS
 * and SOLVATIONTYPE radii, forcefields/gbff.cpp:185-290). Units become
 * PD's: sqrt(J/molecule) epsilon, kg, J/molecule force constants, radians.
 * pdgpu_prep_error() describes the last failure of these calls. */
typedef struct pdgpu_ff pdgpu_ff;
typedef struct pdgpu_molecule pdgpu_molecule;
int pdgpu_ff_read(const char *path, pdgpu_ff **out);
void pdgpu_ff_free(pdgpu_ff *ff);
const char *pdgpu_prep_error(void);
/* the file's VDW14SCALING, ELEC14SCALING and GB P1..P5 (gb_p: 5 values) */
int pdgpu_ff_info(const pdgpu_ff *ff, double *vdw14, double *elec14, double *gb_p);

/* NewProtein(ffps, sequence) (system/genpolymer.cpp:12-164,256-270): the
 * sequence grammar of Sequence::BioSequence::append (sequence/
 * sequence.cpp:165-452: '-' separated names, '(...)' one-letter runs,
 * 'X[n]' and '<...>[n]' repeats, '*' polymer start / end with the class's
 * terminal codes, aliases), residues placed by their CONNECT link points,
 * the covalent 1-2 / 1-3 / 1-4 structure (system/molecule.cpp:123-322),
 * an alpha helix and then the extended chain (phi = psi = omega = pi,
 * system/molecule.cpp:677-721,1099-1180), and the FF_Bonded term lists
 * with PD's wildcard parameter lookup and the equilibrium 1-2 / 1-3
 * distances (forcefields/ffbonded.cpp:13-460). */
int pdgpu_peptide_build(const pdgpu_ff *ff, const char *sequence, pdgpu_molecule **out);
void pdgpu_molecule_free(pdgpu_molecule *m);
/* counts: natoms, nresidues, n12, n13, n14 (pairs i < j), nbonds, nangles,
 * ntorsions, nimpropers */
int pdgpu_molecule_counts(const pdgpu_molecule *m, int32_t *counts);
/* per-atom arrays (any may be NULL): positions (A), charge (e), radius
 * (A), epsilon (sqrt(J/molecule)), mass (kg), the Still radius of the atom
 * type, residue index, and Particle::isHydrogen (the type name starts with
 * 'H', forcefields/ffparam.cpp:739-797) */
int pdgpu_molecule_atoms(const pdgpu_molecule *m, double *xyz, double *charge, double *radius, double *epsilon,
                         double *mass, double *gb_radius, int32_t *resid, uint8_t *is_hydrogen);
/* pointers into the molecule: PDB name, type name, raw name, residue name */
int pdgpu_molecule_atom_names(const pdgpu_molecule *m, int32_t i, const char **pdbname, const char **type_name,
                              const char **rawname, const char **parentname);
/* the 1-2 (order 2), 1-3 or 1-4 pairs i < j in PD's order (atom i
 * ascending, then its cov12atom / cov13atom / cov14atom order), with the
 * equilibrium distances of 1-2 and 1-3 pairs in d (A) */
int pdgpu_molecule_pairs(const pdgpu_molecule *m, int32_t order, int32_t *a, int32_t *b, double *d);
/* each atom's cov12atom list (6 slots, -1 past n_cov12) */
int pdgpu_molecule_cov12(const pdgpu_molecule *m, int32_t *cov12, int32_t *n_cov12);
/* per residue N, CA, C atom indices (-1 where absent) */
int pdgpu_molecule_backbone(const pdgpu_molecule *m, int32_t *n_ca_c);
/* MoleculeBase::setResiduePhiPsi, then setResidueOmega unless omega is
 * NaN or ir == 0 (radians); and calcResiduePhiPsi / calcResidueOmega */
int pdgpu_molecule_set_phipsi(pdgpu_molecule *m, int32_t ir, double phi, double psi, double omega);
int pdgpu_molecule_get_phipsi(const pdgpu_molecule *m, int32_t ir, double *phi, double *psi, double *omega);
/* the FF_Bonded term lists into caller arrays sized by the counts, in the
 * layout pdgpu_bonded_set takes */
typedef struct pdgpu_bonded_out {
    int32_t nbonds, nangles, ntorsions, nimpropers;
    int32_t *bond_ij;
    double *bond_lk;
    int32_t *angle_iaj;
    double *angle_tk;
    int32_t *torsion_iabj;
    int32_t *torsion_terms;
    double *torsion_vn, *torsion_n, *torsion_gamma;
    int32_t *improper_iabj;
    int32_t *improper_terms;
    double *improper_vn, *improper_n, *improper_gamma;
} pdgpu_bonded_out;
int pdgpu_molecule_bonded(const pdgpu_molecule *m, pdgpu_bonded_out *out);

/* Device bytes one instance of natoms atoms with the given bond-order sizes
 * would take in this context (admission control). */
int64_t pdgpu_instance_bytes(const pdgpu_ctx *ctx, int32_t natoms, int32_t max_index_delta,
                             int32_t nsparse);
int64_t pdgpu_bytes_free(const pdgpu_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
