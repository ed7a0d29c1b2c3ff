// pdgpu_integrate.h -- device side of the step drivers: PD's steepest
// descent minimiser (protocols/minimise.cpp:99-235) and a velocity Verlet
// integrator, resident on the GPU for many instances at once. Each instance
// carries its own step counter, so minimisations can start and finish at
// different times while sharing launches (continuous batching).
#pragma once
#include "pdgpu_device.h"

namespace pdgpu {

// Per-instance integrator state.
struct InstMD {
    double mult;      // SD step multiplier (m_StepMultiplier)
    double old_e;     // SD reference energy, J/molecule (m_OldEnergy)
    double epot;      // last total energy, J/molecule
    double f2;        // sum of |f|^2 over atoms, (kcal/mol/A)^2
    double ekin;      // kinetic energy, J
    double step_size; // StepSize (A per N)
    double grms_tol;  // force-RMS stop, kcal/mol/A (0: off)
    int action;       // see below
    int running;      // minimisation started and not finished
    int done;         // finished (converged, out of steps, or failed)
    int steps_done;   // SD iterations run, -1 on failure
    int step;         // next iteration number
    int max_steps;    // Steps
    int update_nlist; // UpdateNList
    int pad;
};

enum { kActNone = 0, kActMove = 1, kActRestore = 2, kActSave = 3 };

// Total energy of an instance in kcal/mol from the device record.
template <typename R>
__device__ __forceinline__ double inst_epot_kcal(const DevState &S, int s) {
    double e = 0;
    for (int k = kEVdw; k <= kEImproper; ++k) e += acc_read<R>(S.inst_energy, (size_t)s * kENum + k, kEnergyScale);
    return e;
}

__global__ void k_sd_start(InstMD *md, int s, int max_steps, double step_size, double grms_tol, int update_nlist) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    InstMD &m = md[s];
    m.mult = 1.0;
    m.old_e = 0;
    m.epot = 0;
    m.f2 = 0;
    m.ekin = 0;
    m.step_size = step_size;
    m.grms_tol = grms_tol;
    m.action = kActNone;
    m.running = 1;
    m.done = max_steps <= 0 ? 1 : 0;
    m.steps_done = 0;
    m.step = 0;
    m.max_steps = max_steps;
    m.update_nlist = update_nlist;
}

// Rebuild flags of running minimisations: PD refreshes the list when
// step % UpdateNList == 0 (protocols/protocolbase.cpp:113-122).
// With update_nlist < 0 an instance rebuilds when k_moved has flagged it
// (the Verlet-skin criterion, an extension PD does not have); the flags are
// cleared here for the next step.
// One workgroup: the flags per instance, and *any = whether any is set
// (written, not or-ed, so no reset is needed between steps).
__global__ void __launch_bounds__(1024) k_sd_need(const InstMD *md, int *need, int *moved, int *any, int s0,
                                                  int s1) {
    int found = 0;
    for (int s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const InstMD &m = md[s];
        const bool sched = m.update_nlist > 0 && (m.step % m.update_nlist) == 0;
        const bool skin = m.update_nlist < 0 && moved[s] != 0;
        const bool n = m.running && !m.done && (sched || skin);
        need[s] = n ? 1 : 0;
        moved[s] = 0;
        found |= n ? 1 : 0;
    }
    found = __syncthreads_or(found);
    if (threadIdx.x == 0) *any = found;
}

// The per-instance |f|^2 sums (InstMD::f2) are added by k_finalize: an SD
// step passes it &md[0].f2 with this stride.
static_assert(sizeof(InstMD) % sizeof(double) == 0, "InstMD is read with a stride in doubles");
constexpr int kInstMDStride = (int)(sizeof(InstMD) / sizeof(double));

// PD's accept / reject rule (protocols/minimise.cpp:181-235), one thread per
// instance, on the energy just evaluated.
template <typename R>
__global__ void k_sd_decide(DevState S, InstMD *md, const int *natoms, int s0, int s1, double jpk) {
    const int s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    InstMD &m = md[s];
    m.action = kActNone;
    if (natoms[s] <= 0 || !m.running || m.done) { m.f2 = 0; return; }
    const double E = inst_epot_kcal<R>(S, s) * jpk;
    const double grms = sqrt(m.f2 / (3.0 * natoms[s]));
    m.f2 = 0;
    m.epot = E;
    if (S.inst_status[s] != 0 || !isfinite(E)) {
        m.done = 1;
        m.steps_done = -1; // Minimisation::runcore returns -1 (minimise.cpp:123-127)
        return;
    }
    const int step = m.step++;
    m.steps_done = m.step;
    if (step != 0 && !(E <= m.old_e)) {
        // uphill: restore the saved structure, shrink the step
        m.mult *= 0.5 / 1.2;
        m.old_e += 1; // PD adds 1 J/molecule: the next step is always accepted
        m.action = kActRestore;
    } else {
        if (step != 0) m.mult *= 1.2;
        m.old_e = E;
        if (m.grms_tol > 0 && grms < m.grms_tol) {
            m.action = kActSave; // converged: keep this structure
            m.done = 1;
            return;
        }
        m.action = kActMove;
    }
    if (m.step >= m.max_steps) m.done = 1;
}

// Per-atom SD update: save and move, restore, or save only (converged).
__global__ void k_sd_update(DevState S, const InstMD *md, double *opx, double *opy, double *opz, double *px,
                            double *py, double *pz, int atom_begin, int atom_end, double f_to_n) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= atom_end) return;
    const int b = a / kBlk;
    const int inst = S.blk_inst[b];
    if (inst < 0 || (a % kBlk) >= S.blk_n[b]) return;
    const int act = md[inst].action;
    if (act == kActMove || act == kActSave) {
        opx[a] = px[a]; opy[a] = py[a]; opz[a] = pz[a];
        if (act == kActMove) {
            const double s = f_to_n * (md[inst].step_size * md[inst].mult);
            px[a] += S.fx[a] * s;
            py[a] += S.fy[a] * s;
            pz[a] += S.fz[a] * s;
        }
    } else if (act == kActRestore) {
        px[a] = opx[a]; py[a] = opy[a]; pz[a] = opz[a];
    }
}

// cur = old at the end of a minimisation (protocols/minimise.cpp:173);
// failed instances keep their positions, as the reference returns early.
__global__ void k_copy_back(DevState S, const InstMD *md, const double *opx, const double *opy, const double *opz,
                            double *px, double *py, double *pz, int atom_begin, int atom_end) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= atom_end) return;
    const int b = a / kBlk;
    const int inst = S.blk_inst[b];
    if (inst < 0 || (a % kBlk) >= S.blk_n[b]) return;
    if (md[inst].steps_done < 0) return;
    px[a] = opx[a]; py[a] = opy[a]; pz[a] = opz[a];
}

// Velocity Verlet half kick (and drift when drift != 0). Forces are
// kcal/mol/A; f_to_n converts to N, minv is 1/kg, positions in A, v in m/s.
__global__ void k_vv(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz, double *px,
                     double *py, double *pz, int atom_begin, int atom_end, double dt, double f_to_n, int drift,
                     int want_ekin) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double h = 0.5 * dt * minv[a] * f_to_n;
        const double ux = vx[a] + S.fx[a] * h, uy = vy[a] + S.fy[a] * h, uz = vz[a] + S.fz[a] * h;
        vx[a] = ux; vy[a] = uy; vz[a] = uz;
        if (drift) {
            const double s = dt / 1e-10;
            px[a] += ux * s; py[a] += uy * s; pz[a] += uz * s;
        }
        if (minv[a] > 0) ek = 0.5 / minv[a] * (ux * ux + uy * uy + uz * uz);
    }
    if (want_ekin) {
        ek = wave_sum(ek);
        if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
    }
}

// PD's VelocityVerlet update as written (protocols/md.cpp:788-838), with
// dv = f*dt/2m from the forces of this step: v += dv (the kinetic energy is
// taken here, md.cpp:815), v += dv, p += dt*(v + dv) / Angstrom, v += dv.
__global__ void k_vv_pd(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz,
                        double *px, double *py, double *pz, int atom_begin, int atom_end, double dt,
                        double f_to_n, int want_ekin) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double tinvmass = 0.5 * dt * minv[a];
        const double dx = S.fx[a] * f_to_n * tinvmass, dy = S.fy[a] * f_to_n * tinvmass,
                     dz = S.fz[a] * f_to_n * tinvmass;
        double ux = vx[a] + dx, uy = vy[a] + dy, uz = vz[a] + dz;
        if (minv[a] > 0) {
            const double m = sqrt(ux * ux + uy * uy + uz * uz);
            ek = 0.5 / minv[a] * (m * m);
        }
        ux += dx; uy += dy; uz += dz;
        const double s = 1.0 / 1e-10;
        px[a] += (dt * (ux + dx)) * s;
        py[a] += (dt * (uy + dy)) * s;
        pz[a] += (dt * (uz + dz)) * s;
        vx[a] = ux + dx; vy[a] = uy + dy; vz[a] = uz + dz;
    }
    if (want_ekin) {
        ek = wave_sum(ek);
        if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
    }
}

// k_finalize and the closing half kick of a velocity-Verlet step in one
// pass; with next != 0 also the following step's opening half kick and
// drift (the same arithmetic as the kernels in sequence), and the following
// evaluation's set-up (k_prepare's work: FP32 positions of the drifted
// atoms, Born sums and block records cleared, instance records cleared), so
// that evaluation starts without a set-up launch. Only the last step's
// energies are read, so intermediate steps clear their block records
// instead of folding them into the instance records.
template <typename R>
__global__ void k_finalize_kick(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz,
                                double *px, double *py, double *pz, int atom_begin, int atom_end, double dt,
                                double f_to_n, int want_ekin, int next) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    if (!next) blk_fold<R>(S, b, inst);
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double fx = -acc_read<R>(S.acc_fx, a, kForceScale);
        const double fy = -acc_read<R>(S.acc_fy, a, kForceScale);
        const double fz = -acc_read<R>(S.acc_fz, a, kForceScale);
        S.fx[a] = fx; S.fy[a] = fy; S.fz[a] = fz;
        if (!(isfinite(fx) && isfinite(fy) && isfinite(fz))) atomicOr(S.inst_status + inst, kStatNonFinite);
        const double h = 0.5 * dt * minv[a] * f_to_n;
        double ux = vx[a] + fx * h, uy = vy[a] + fy * h, uz = vz[a] + fz * h;
        if (minv[a] > 0) ek = 0.5 / minv[a] * (ux * ux + uy * uy + uz * uz);
        if (next) {
            ux += fx * h; uy += fy * h; uz += fz * h;
            const double s = dt / 1e-10;
            px[a] += ux * s; py[a] += uy * s; pz[a] += uz * s;
        }
        vx[a] = ux; vy[a] = uy; vz[a] = uz;
    }
    if (next && in) {
        blk_clear(S, a);
        acc_store<R>(S.acc_g, a, 0.0, 1.0);
        if constexpr (Prec<R>::fixed) store_xyz32(S, a);
        if (lane_id() == 0 && inst >= 0 && S.blk_first[b] == b) inst_clear(S, inst);
    }
    if (want_ekin) {
        ek = wave_sum(ek);
        if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
    }
}

// ------------------------------------------ Langevin (protocols/md.cpp:906-1043)
// Per-call constants of PD's Langevin update (host-computed with the
// reference's formulas, md.cpp:950-985).
struct LangevinCoef {
    double c0, c1, c2; // velocity decay and force weights
    double varv, varr; // variances of the velocity and position increments, in units of kB T / m
    double crv, scrv;  // their correlation and sqrt(1 - crv^2)
    double gamma, gt;  // friction [1/s] and gamma * dt
    double kT;         // kB * T [J]
};

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ten rounds of
// two 32x32->64 multiplies with the published multipliers and Weyl key
// increments.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        const unsigned lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const unsigned lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// 53-bit uniform in (0, 1], as PD's frand() excludes 0 (maths/maths.cpp:256)
__device__ __forceinline__ double uniform53(unsigned hi, unsigned lo) {
    return ((double)((((unsigned long long)hi << 32) | lo) >> 11) + 1.0) * 0x1p-53;
}

// The six standard normal deviates of one atom and step, (n1x, n1y, n1z,
// n2x, n2y, n2z), by Box-Muller as PD's nrand (maths/maths.h:138-152) over
// the counter (atom, instance | pair << 24, step) and key seed.
__device__ __forceinline__ void langevin_normals(uint2 key, unsigned atom, unsigned slot, unsigned long long step,
                                                 double n[6]) {
    for (int q = 0; q < 3; ++q) {
        const uint4 r = philox4x32_10(make_uint4(atom, slot | ((unsigned)q << 24), (unsigned)step,
                                                 (unsigned)(step >> 32)),
                                      key);
        const double u1 = uniform53(r.x, r.y), u2 = uniform53(r.z, r.w);
        const double l = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        n[q] = l * cs;
        n[3 + q] = l * sn;
    }
}

// One Langevin update (LangevinOnHydrogens), after the step's forces: with
// first == 0 the velocity is finished with the new forces, v += (1 -
// c0/c1) f dt/m / (gamma dt); the kinetic energy is taken; with last == 0
// positions and velocities advance: dr = dt (c1 v + c2 f dt/m) + sigma_r
// n1, v = c0 v + c0 c2 f dt/m / c1 + sigma_v (crv n1 + sqrt(1 - crv^2) n2)
// (md.cpp:998-1041). noise, when given, holds the step's deviates per
// atom of the instance as (n1x, n1y, n1z, n2x, n2y, n2z); otherwise they
// come from langevin_normals at the step counter *ctr.
__global__ void k_langevin_pd(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz,
                              double *px, double *py, double *pz, int atom_begin, int atom_end, double dt,
                              double f_to_n, LangevinCoef L, int first, int last, int want_ekin,
                              const double *noise, const unsigned long long *ctr, uint2 key,
                              const uint8_t *is_h, int on_h) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double im = minv[a], tm = dt * im;
        const double f[3] = {S.fx[a] * f_to_n, S.fy[a] * f_to_n, S.fz[a] * f_to_n};
        double u[3] = {vx[a], vy[a], vz[a]};
        // LangevinOnHydrogens off: hydrogens take PD's plain half kicks
        // and draw no deviates (md.cpp:1125-1180)
        const bool plain = !on_h && is_h[a];
        if (!first)
            for (int c = 0; c < 3; ++c)
                u[c] += plain ? f[c] * (0.5 * dt * im) : (1.0 - L.c0 / L.c1) * f[c] * tm / L.gt;
        if (im > 0) {
            const double m = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            ek = 0.5 / im * (m * m);
        }
        if (!last && plain) {
            double p[3] = {px[a], py[a], pz[a]};
            for (int c = 0; c < 3; ++c) {
                const double dv = f[c] * (0.5 * dt * im);
                double dr = dt * (u[c] + dv);
                dr *= 1.0 / 1e-10;
                p[c] += dr;
                u[c] += dv;
            }
            px[a] = p[0]; py[a] = p[1]; pz[a] = p[2];
        } else if (!last) {
            const int local = a - S.blk_first[b] * kBlk;
            double n[6];
            if (noise) {
                for (int k = 0; k < 6; ++k) n[k] = noise[(size_t)local * 6 + k];
            } else {
                langevin_normals(key, (unsigned)local, (unsigned)inst, *ctr, n);
            }
            const double kTdivm = L.kT * im;
            const double sv = sqrt(kTdivm * L.varv), sr = sqrt(kTdivm * L.varr) / L.gamma;
            double p[3] = {px[a], py[a], pz[a]};
            for (int c = 0; c < 3; ++c) {
                const double n1 = n[c], n2 = n[3 + c];
                const double dlv = sv * (L.crv * n1 + L.scrv * n2), dlr = sr * n1;
                double dr = dt * (L.c1 * u[c] + L.c2 * f[c] * tm) + dlr;
                dr *= 1.0 / 1e-10;
                const double dv = (L.c0 * L.c2) * f[c] * tm / L.c1 + dlv;
                p[c] += dr;
                u[c] *= L.c0;
                u[c] += dv;
            }
            px[a] = p[0]; py[a] = p[1]; pz[a] = p[2];
        }
        vx[a] = u[0]; vy[a] = u[1]; vz[a] = u[2];
    }
    if (want_ekin) {
        ek = wave_sum(ek);
        if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
    }
}

__global__ void k_counter_step(unsigned long long *ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr += 1;
}

// ---------------------------------------------- Beeman (protocols/md.cpp:840-904)
// One Beeman update after the step's forces: with first == 0 v += 2 f
// dt/6m; the kinetic energy is taken; with last == 0 p += dt (v + (4 f -
// f_old) dt/6m) / A, v += (5 f - f_old) dt/6m and f_old = f. f_old (N) is
// the workspace's old forces, kept per atom between calls.
__global__ void k_beeman_pd(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz,
                            double *px, double *py, double *pz, double *ox, double *oy, double *oz, int atom_begin,
                            int atom_end, double dt, double f_to_n, int first, int last, int want_ekin) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double im = minv[a], tm = dt * im / 6.0;
        const double f[3] = {S.fx[a] * f_to_n, S.fy[a] * f_to_n, S.fz[a] * f_to_n};
        double u[3] = {vx[a], vy[a], vz[a]};
        if (!first)
            for (int c = 0; c < 3; ++c) u[c] += (2.0 * f[c]) * tm;
        if (im > 0) {
            const double m = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            ek = 0.5 / im * (m * m);
        }
        if (!last) {
            const double fo[3] = {ox[a], oy[a], oz[a]};
            double p[3] = {px[a], py[a], pz[a]};
            for (int c = 0; c < 3; ++c) {
                double dr = dt * (u[c] + (4.0 * f[c] - fo[c]) * tm);
                dr *= 1.0 / 1e-10;
                p[c] += dr;
                u[c] += (5.0 * f[c] - fo[c]) * tm;
            }
            px[a] = p[0]; py[a] = p[1]; pz[a] = p[2];
            ox[a] = f[0]; oy[a] = f[1]; oz[a] = f[2];
        }
        vx[a] = u[0]; vy[a] = u[1]; vz[a] = u[2];
    }
    if (want_ekin) {
        ek = wave_sum(ek);
        if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
    }
}

// ------------------------------ Berendsen thermostat (protocols/md.cpp:580-597)
// PD's applyThermostat runs after the kinetic energy of a step, so a step
// with the thermostat splits in two kernels: the velocity update up to the
// kinetic energy (summed per instance), then every velocity times
// chi = sqrt(1 + dt (Edes / ekin - 1) / tau), Edes = 3 N kB T / 2, and the
// rest of the update.
__device__ __forceinline__ double berendsen_chi(double ekin, int natoms, double dt, double tau, double kT) {
    const double edes = 3.0 * natoms * kT / 2.0;
    return sqrt(1.0 + dt * (edes / ekin - 1.0) / tau);
}

// kind 0: VelocityVerlet as PD writes it (first half kick); kind 1: Beeman
// (finishing kick 2 f dt/6m after the first step). Adds the kinetic energy.
__global__ void k_md_thermo_a(DevState S, InstMD *md, const double *minv, double *vx, double *vy, double *vz,
                              int atom_begin, int atom_end, double dt, double f_to_n, int kind, int first) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = in && inst >= 0 && (a % kBlk) < S.blk_n[b];
    double ek = 0;
    if (v) {
        const double im = minv[a];
        const double f[3] = {S.fx[a] * f_to_n, S.fy[a] * f_to_n, S.fz[a] * f_to_n};
        double u[3] = {vx[a], vy[a], vz[a]};
        for (int c = 0; c < 3; ++c) {
            if (kind == 0) u[c] += f[c] * (0.5 * dt * im);
            else if (!first) u[c] += (2.0 * f[c]) * (dt * im / 6.0);
        }
        if (im > 0) {
            const double m = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            ek = 0.5 / im * (m * m);
        }
        vx[a] = u[0]; vy[a] = u[1]; vz[a] = u[2];
    }
    ek = wave_sum(ek);
    if (lane_id() == 0 && in && inst >= 0 && ek != 0.0) unsafeAtomicAdd(&md[inst].ekin, ek);
}

// chi, then the rest of the step: kind 0 PD's VelocityVerlet (v += f dt/2m,
// p += dt (v + f dt/2m), v += f dt/2m); kind 1 Beeman's update unless last.
__global__ void k_md_thermo_b(DevState S, const InstMD *md, const int *natoms, const double *minv, double *vx,
                              double *vy, double *vz, double *px, double *py, double *pz, double *ox, double *oy,
                              double *oz, int atom_begin, int atom_end, double dt, double f_to_n, double tau,
                              double kT, int kind, int last) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= atom_end) return;
    const int b = a / kBlk;
    const int inst = S.blk_inst[b];
    if (inst < 0 || (a % kBlk) >= S.blk_n[b]) return;
    const double chi = berendsen_chi(md[inst].ekin, natoms[inst], dt, tau, kT);
    const double im = minv[a];
    const double f[3] = {S.fx[a] * f_to_n, S.fy[a] * f_to_n, S.fz[a] * f_to_n};
    double u[3] = {vx[a] * chi, vy[a] * chi, vz[a] * chi};
    double p[3] = {px[a], py[a], pz[a]};
    if (kind == 0) {
        for (int c = 0; c < 3; ++c) {
            const double dv = f[c] * (0.5 * dt * im);
            u[c] += dv;
            double dr = dt * (u[c] + dv);
            dr *= 1.0 / 1e-10;
            p[c] += dr;
            u[c] += dv;
        }
    } else if (!last) {
        const double tm = dt * im / 6.0;
        const double fo[3] = {ox[a], oy[a], oz[a]};
        for (int c = 0; c < 3; ++c) {
            double dr = dt * (u[c] + (4.0 * f[c] - fo[c]) * tm);
            dr *= 1.0 / 1e-10;
            p[c] += dr;
            u[c] += (5.0 * f[c] - fo[c]) * tm;
        }
        ox[a] = f[0]; oy[a] = f[1]; oz[a] = f[2];
    }
    vx[a] = u[0]; vy[a] = u[1]; vz[a] = u[2];
    px[a] = p[0]; py[a] = p[1]; pz[a] = p[2];
}

__global__ void k_ekin_clear(InstMD *md, int s0, int s1) {
    const int s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s < s1) md[s].ekin = 0;
}

} // namespace pdgpu
