// pdgpu_bonded.h -- FF_Bonded on the device (forcefields/ffbonded.cpp:809-1590):
// harmonic bonds, harmonic angles, Fourier / harmonic dihedrals for
// torsions and impropers. One work item per term; every term reads its
// 2-4 atoms' FP64 positions, and the gradients of a workgroup's terms are
// summed in LDS before they reach the same per-atom accumulators as the pair
// kernels (64-bit fixed point in the mixed path, FP64 in the FP64 path), so
// bonded and GB forces come out of one k_finalize. The arithmetic is FP64 in
// both precisions. (One global atomic per term and component instead of the
// LDS window measured several times slower once the terms are off their
// minima; at a freshly built geometry most bond and angle gradients round to
// zero and skip their atomics, which hides the cost.)
#pragma once

#include "pdgpu_device.h"

namespace pdgpu {

// component bits (pdgpu_options::components)
constexpr int kFFGB = 1, kFFBonds = 2, kFFAngles = 4, kFFTorsions = 8, kFFImpropers = 16;

// Term records. Atom indices are global (padded layout); the owning
// instance is stored as slot + 1 so that zeroed memory is a free record.
struct BondedState {
    const int4 *bd_idx;      // (i, j, slot+1, 0)
    const double2 *bd_par;   // (l [A], k [kcal/mol/A^2])
    const int4 *an_idx;      // (i, apex a, j, slot+1)
    const double2 *an_par;   // (theta0 [rad], k [kcal/mol/rad^2])
    const int4 *dh_idx;      // (i, a, b, j)
    const int *dh_meta;      // (slot+1) | terms << 24 | improper << 28
    const double4 *dh_vn;    // Vn [kcal/mol], 4 terms
    const double4 *dh_n;     // periodicity n (0: harmonic)
    const double4 *dh_gamma; // phase gamma [rad]
    int b0, b1, a0, a1, d0, d1; // record ranges evaluated
    int comps;
};

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3_sub(const D3 &a, const D3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 d3_cross(const D3 &a, const D3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double d3_dot(const D3 &a, const D3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 d3_lin(double s, const D3 &a, double t, const D3 &b) {
    return {s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z};
}
__device__ __forceinline__ D3 d3_at(const DevState &S, int a) { return {S.px[a], S.py[a], S.pz[a]}; }

// Gradient sink of one workgroup. Terms are stored instance by instance in
// atom order, so the 256 terms of a workgroup touch a narrow range of atoms:
// their gradients are summed in an LDS window of kWin atoms starting at the
// smallest atom index of the workgroup, and the window is added to the
// global accumulators once at the end. Atoms beyond the window go straight
// to the global accumulators. In the mixed path every contribution is
// rounded to 64-bit fixed point before it is added, so the sums are exact
// integers whatever the grouping.
constexpr int kWin = 1024;

template <typename R> struct GradSink {
    unsigned long long (*win)[kWin]; // [3][kWin]: int64 fixed point (mixed) or double bits (FP64)
    int amin;
    const DevState *S;

    __device__ void add(int a, const D3 &g, int *st) const {
        const double v[3] = {g.x, g.y, g.z};
        const int k = a - amin;
        if (k >= 0 && k < kWin) {
            for (int c = 0; c < 3; ++c) {
                if constexpr (Prec<R>::fixed) {
                    const double sc = v[c] * kForceScale;
                    if (!(fabs(sc) < kFixedLimit)) { atomicOr(st, (sc == sc) ? kStatOverflow : kStatNonFinite); continue; }
                    const long long iv = to_fixed(sc);
                    if (iv != 0) atomicAdd(&win[c][k], (unsigned long long)iv);
                } else {
                    if (!(fabs(v[c]) < 1e300)) atomicOr(st, kStatNonFinite);
                    if (v[c] != 0.0) atomicAdd(reinterpret_cast<double *>(&win[c][k]), v[c]);
                }
            }
        } else {
            acc_add<R>(S->acc_fx, a, g.x, kForceScale, st);
            acc_add<R>(S->acc_fy, a, g.y, kForceScale, st);
            acc_add<R>(S->acc_fz, a, g.z, kForceScale, st);
        }
    }
};

// Energy slot of the block record += e for every valid lane (blk: the block
// of the term's first atom). Called by the whole wavefront; lanes are grouped
// by record field, one atomic per distinct field (a wavefront's terms mostly
// share one or two blocks). In the mixed path each term's energy is rounded
// to fixed point on its own and the integers are summed, so an instance's
// total does not depend on which other terms share its wavefront.
template <typename R>
__device__ void add_term_energy(const DevState &S, bool valid, int inst, int blk, int slot, double e) {
    unsigned long long rem = __ballot(valid);
    if (!rem) return;
    const int key = valid ? blk * kBAcc + slot : -1;
    if constexpr (Prec<R>::fixed) {
        long long iv = 0;
        if (valid) {
            const double s = e * kEnergyScale;
            if (fabs(s) < kFixedLimit) iv = to_fixed(s);
            else atomicOr(S.inst_status + inst, (s == s) ? kStatOverflow : kStatNonFinite);
        }
        while (rem) {
            const int lead = __ffsll((unsigned long long)rem) - 1;
            const int key0 = __shfl(key, lead, 64);
            const bool mine = key == key0;
            rem &= ~__ballot(mine);
            const long long tot = wave_sum(mine ? iv : 0ll);
            if (lane_id() == lead && tot != 0) atomicAdd(S.blk_acc + key0, (unsigned long long)tot);
        }
    } else {
        if (valid && !(fabs(e) < 1e300)) atomicOr(S.inst_status + inst, kStatNonFinite);
        while (rem) {
            const int lead = __ffsll((unsigned long long)rem) - 1;
            const int key0 = __shfl(key, lead, 64);
            const bool mine = key == key0;
            rem &= ~__ballot(mine);
            const double tot = wave_sum(mine ? e : 0.0);
            if (lane_id() == lead && tot != 0.0) unsafeAtomicAdd(reinterpret_cast<double *>(S.blk_acc) + key0, tot);
        }
    }
}

// Dihedral i-a-b-j. phi is the IUPAC angle, which is what the reference's
// phi = -atan2(sin, cos) over its normals evaluates to
// (forcefields/ffbonded.cpp:1269-1296). The gradient uses the closed form of
// dphi/dx (Blondel & Karplus, J. Comput. Chem. 17:1132, 1996) instead of the
// reference's two-branch cos/sin derivative; both are the exact gradient.
template <typename R>
__device__ double dihedral_term(const DevState &S, const GradSink<R> &G, const int4 &q, int terms, const double4 &vn,
                                const double4 &nn, const double4 &gm, int *st) {
    const D3 ri = d3_at(S, q.x), ra = d3_at(S, q.y), rb = d3_at(S, q.z), rj = d3_at(S, q.w);
    const D3 b1 = d3_sub(ra, ri), b2 = d3_sub(rb, ra), b3 = d3_sub(rj, rb);
    const D3 m = d3_cross(b1, b2), n = d3_cross(b2, b3);
    const double m2 = d3_dot(m, m), n2 = d3_dot(n, n), b22 = d3_dot(b2, b2);
    // reciprocals and square roots as in the FP64 pair terms (hardware
    // estimate + two Newton steps)
    const double ib2 = Prec<double>::rsqrt(b22), lb2 = b22 * ib2;
    const double phi = atan2(lb2 * d3_dot(b1, n), d3_dot(m, n));
    const double V[4] = {vn.x, vn.y, vn.z, vn.w}, N[4] = {nn.x, nn.y, nn.z, nn.w}, Ph[4] = {gm.x, gm.y, gm.z, gm.w};
    double e = 0, dedphi = 0;
    for (int t = 0; t < 4; ++t) {
        if (t >= terms) break;
        if (N[t] > 0) {
            double s, c;
            sincos(N[t] * phi + Ph[t], &s, &c);
            e += V[t] * (1.0 + c);
            dedphi -= N[t] * V[t] * s;
        } else {
            double d = phi - Ph[t];
            if (d < -M_PI) d += 2.0 * M_PI;
            else if (d > M_PI) d -= 2.0 * M_PI;
            e += V[t] * d * d;
            dedphi += 2.0 * V[t] * d;
        }
    }
    const D3 gi = d3_lin(-dedphi * lb2 * Prec<double>::rcp(m2), m, 0.0, m);
    const D3 gj = d3_lin(dedphi * lb2 * Prec<double>::rcp(n2), n, 0.0, n);
    const double ib22 = ib2 * ib2;
    const double p1 = d3_dot(b1, b2) * ib22, p3 = d3_dot(b3, b2) * ib22;
    G.add(q.x, gi, st);
    G.add(q.y, d3_lin(-(1.0 + p1), gi, p3, gj), st);
    G.add(q.z, d3_lin(-(1.0 + p3), gj, p1, gi), st);
    G.add(q.w, gj, st);
    return e;
}

// LDS of one workgroup of the bonded terms.
struct BondedLds {
    unsigned long long win[3][kWin];
    int amin;
};

// Workgroup `group` (256 threads) of the record ranges [b0,b1) bonds,
// [a0,a1) angles, [d0,d1) dihedrals, one thread per record. Called by whole
// workgroups (it synchronises the workgroup).
template <typename R> __device__ __forceinline__ void bonded_group(const DevState &S, const BondedState &B, int group,
                                                                   BondedLds &L) {
    unsigned long long(*win)[kWin] = L.win;
    int &s_amin = L.amin;
    const int t = group * 256 + threadIdx.x;
    const int nb = B.b1 - B.b0, na = B.a1 - B.a0, nd = B.d1 - B.d0;
    // decode this thread's record
    int kind = -1, r = 0, inst = -1;
    int4 q = make_int4(0, 0, 0, 0);
    if (t < nb) {
        r = B.b0 + t;
        const int4 id = B.bd_idx[r];
        if (id.z > 0 && (B.comps & kFFBonds)) { kind = 0; inst = id.z - 1; q = make_int4(id.x, id.y, id.x, id.x); }
    } else if (t < nb + na) {
        r = B.a0 + (t - nb);
        const int4 id = B.an_idx[r];
        if (id.w > 0 && (B.comps & kFFAngles) && B.an_par[r].y != 0.0) {
            kind = 1; inst = id.w - 1; q = make_int4(id.x, id.y, id.z, id.x);
        }
    } else if (t < nb + na + nd) {
        r = B.d0 + (t - nb - na);
        const int meta = B.dh_meta[r];
        const int impr = (meta >> 28) & 1;
        if ((meta & 0xFFFFFF) > 0 && (B.comps & (impr ? kFFImpropers : kFFTorsions))) {
            kind = 2 + impr; inst = (meta & 0xFFFFFF) - 1; q = B.dh_idx[r];
        }
    }
    // LDS window at the workgroup's smallest atom index
    if (threadIdx.x == 0) s_amin = 0x7FFFFFFF;
    for (int k = threadIdx.x; k < 3 * kWin; k += blockDim.x) (&win[0][0])[k] = 0ull;
    __syncthreads();
    if (kind >= 0) atomicMin(&s_amin, min(min(q.x, q.y), min(q.z, q.w)));
    __syncthreads();
    const GradSink<R> G{win, s_amin, &S};
    int slot = kEBond;
    double e = 0, vir = 0;
    if (kind == 0) {
        // calcBondForces (ffbonded.cpp:885-940): E = k(d-l)^2
        int *st = S.inst_status + inst;
        const double2 p = B.bd_par[r];
        const D3 v = d3_sub(d3_at(S, q.x), d3_at(S, q.y));
        const double d2 = d3_dot(v, v), id = Prec<double>::rsqrt(d2), d = d2 * id;
        const double dd = d - p.x;
        e = p.y * dd * dd;
        const double dedd = 2.0 * p.y * dd;
        vir = dedd * d; // InternalVirial += d * force (ffbonded.cpp:924)
        const D3 g = d3_lin(dedd * id, v, 0.0, v);
        G.add(q.x, g, st);
        G.add(q.y, d3_lin(-1.0, g, 0.0, g), st);
    } else if (kind == 1) {
        // calcAngleForces (ffbonded.cpp:994-1120): E = k(theta-theta0)^2
        slot = kEAngle;
        int *st = S.inst_status + inst;
        const double2 p = B.an_par[r];
        const D3 ra = d3_at(S, q.y);
        const D3 vi = d3_sub(d3_at(S, q.x), ra), vj = d3_sub(d3_at(S, q.z), ra);
        const double li = Prec<double>::rsqrt(d3_dot(vi, vi)), lj = Prec<double>::rsqrt(d3_dot(vj, vj));
        const D3 ui = d3_lin(li, vi, 0.0, vi), uj = d3_lin(lj, vj, 0.0, vj);
        const double c = d3_dot(ui, uj);
        const double th = acos(c), sn = sin(th);
        const double dth = th - p.x;
        e = p.y * dth * dth;
        const double dedth = 2.0 * p.y * dth;
        // dtheta/dx_i = (u_i cos - u_j) / (|v_i| sin)
        const double isn = Prec<double>::rcp(sn);
        const D3 gi = d3_lin(dedth * li * isn * c, ui, -dedth * li * isn, uj);
        const D3 gj = d3_lin(dedth * lj * isn * c, uj, -dedth * lj * isn, ui);
        G.add(q.x, gi, st);
        G.add(q.z, gj, st);
        G.add(q.y, d3_lin(-1.0, gi, -1.0, gj), st);
    } else if (kind >= 2) {
        // calcDihedralForces (ffbonded.cpp:1199-1440)
        slot = kind == 3 ? kEImproper : kETorsion;
        const int meta = B.dh_meta[r];
        e = dihedral_term<R>(S, G, q, (meta >> 24) & 7, B.dh_vn[r], B.dh_n[r], B.dh_gamma[r], S.inst_status + inst);
    }
    add_term_energy<R>(S, kind >= 0, inst, q.x / kBlk, slot, e);
    add_term_energy<R>(S, kind == 0, inst, q.x / kBlk, kEVirial, vir);
    __syncthreads();
    // flush the window (only atoms that received something)
    for (int k = threadIdx.x; k < kWin; k += blockDim.x) {
        for (int c = 0; c < 3; ++c) {
            const unsigned long long v = win[c][k];
            if (v == 0ull) continue;
            void *base = c == 0 ? S.acc_fx : c == 1 ? S.acc_fy : S.acc_fz;
            if constexpr (Prec<R>::fixed) {
                atomicAdd(reinterpret_cast<unsigned long long *>(base) + (s_amin + k), v);
            } else {
                unsafeAtomicAdd(reinterpret_cast<double *>(base) + (s_amin + k), __longlong_as_double((long long)v));
            }
        }
    }
}

// The bonded terms on their own launch (contexts without the GB pass, and
// the energy path). With GB forces they run as extra workgroups of k_chain.
template <typename R> __global__ void __launch_bounds__(256) k_bonded(DevState S, BondedState B) {
    __shared__ BondedLds L;
    bonded_group<R>(S, B, blockIdx.x, L);
}

// The chain-rule pass (workgroups [0, ngroups)) and the bonded terms
// (workgroups [ngroups, gridDim.x)) in one launch: both add to the force
// sums the Born-radius kernel cleared, and the bonded terms' latency hides
// behind the tiles instead of taking a launch of its own.
template <typename R>
__global__ void __launch_bounds__(256) k_chain_bonded(DevState S, PairParams P, int blk_begin, int blk_end, int tpw,
                                                      int ngroups, BondedState B) {
    __shared__ union {
        ChainLds<R> chain;
        BondedLds bonded;
    } U;
    if ((int)blockIdx.x < ngroups) chain_work<R>(S, P, blk_begin, blk_end, tpw, ngroups, U.chain);
    else bonded_group<R>(S, B, (int)blockIdx.x - ngroups, U.bonded);
}

// Clears the force accumulators where the GB kernels (which clear them in
// k_born_final) do not run.
template <typename R>
__global__ void k_clear_grad(DevState S, int atom_begin, int atom_end, int slot_begin, int slot_end) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot_begin + t < slot_end) inst_clear(S, slot_begin + t); // the per-instance records too
    const int a = atom_begin + t;
    if (a >= atom_end) return;
    blk_clear(S, a);
    acc_store<R>(S.acc_fx, a, 0.0, 1.0);
    acc_store<R>(S.acc_fy, a, 0.0, 1.0);
    acc_store<R>(S.acc_fz, a, 0.0, 1.0);
    acc_store<R>(S.acc_deda, a, 0.0, 1.0);
}

} // namespace pdgpu
