// pdgpu_pair32.h -- the mixed-precision force pass (FastMode pair pass of
// FF_GeneralizedBorn_Still::calcForcesIncludingVacuo, forcefields/gbff.cpp:
// 1529-1729) written for the FP32 VALU of gfx950.
//
// Same tiles, lists and dense-tile bit rows as k_pair<float, 0>, with the
// instruction count per pair cut where the generic kernel spends it:
//  - per-atom records (x, y, z, q) and (R, sqrt(eps), a, c/a) as two float4,
//    written once per evaluation by k_born_final, so a lane gets an atom
//    with two 16-byte LDS reads; the exponent of the GB term is one product
//    r^2 (c/a_i)(c/a_j) with c^2 = log2(e)/4, no reciprocal;
//  - the pair term has no branches: far pairs (no GB term) and pairs without
//    Coulomb / LJ get zero factors, so a wavefront never runs a code path
//    twice; lanes without a pair are masked off around the whole term;
//  - the pair's class (bond order <= 2, 1-4, other) selects its Coulomb and
//    LJ scalings from a 12-entry table held in lanes 0-11 (a lane permute,
//    no LDS allocation) instead of compares and selects;
//  - dense tiles read the rotating j atom from an LDS copy of the j block
//    stored twice in a row, so the address of step k is one add; the j-side
//    sums still rotate with DPP wave_ror:1, fused into their update
//    (v_sub_f32_dpp); the pair bits of a lane are pre-rotated once per tile
//    so that step k reads bit k - 1;
//  - one switching function for Coulomb, GB and LJ when their radii agree
//    (template SAME);
//  - i-side force sums and per-lane energy sums of a tile in FP32 (at most
//    64 terms per lane before the flush to FP64 and 64-bit fixed point).
//    Summing the lane energies in FP64 instead was measured to leave the
//    mixed energy error unchanged: it comes from the FP32 pair terms
//    themselves (coordinates and products rounded to 24 bits), a random
//    error of ~2e-7 of each pair energy.
#pragma once
#include "pdgpu_device.h"

namespace pdgpu {

constexpr float kLn2f = 0.69314718055994531f;

// FP32 constants of the pass, derived once per launch from PairParams.
struct MixK {
    float cut2, inner, invSw, inner_s, m4invSw; // inner_s = inner*invSw, m4invSw = -4*invSw
    float vcut, vinner, vinvSw, vinner_s, vm4invSw;
    float s14v, s14e, ke_in, ke_out, pm, thr2;
};

__device__ __forceinline__ MixK mix_consts(const PairParams &P) {
    MixK K;
    K.cut2 = (float)P.cut2;
    K.inner = (float)P.inner;
    K.invSw = (float)P.invSw;
    K.inner_s = K.inner * K.invSw;
    K.m4invSw = -4.0f * K.invSw;
    K.vcut = (float)P.vcut;
    K.vinner = (float)P.vinner;
    K.vinvSw = (float)P.vinvSw;
    K.vinner_s = K.vinner * K.vinvSw;
    K.vm4invSw = -4.0f * K.vinvSw;
    K.s14v = (float)P.s14v;
    K.s14e = (float)P.s14e;
    K.ke_in = P.econv_in32;
    K.ke_out = P.econv_out32;
    K.pm = (float)P.premul;
    K.thr2 = (float)(P.exp_thresh * 1.4426950408889634);
    return K;
}

// Per-lane energy sums of a tile (pol holds the switched GB pair energy
// once; the reference counts it twice, gbff.cpp:1627).
struct MixE {
    float vdw = 0, elec = 0, pol = 0;
    unsigned int pairs = 0;
};

// PD's switching polynomial S = (1 - d^2)^2, d = (max(r, inner) - inner) /
// (cut - inner), and dS/dr (gbff.cpp:1605-1614); below the inner radius
// S = 1, dS = -0.
__device__ __forceinline__ void mix_switch(float r, float inner, float invSw, float inner_s, float m4invSw,
                                           float &Sw, float &dS) {
    const float dn = fmaf(fmaxf(r, inner), invSw, -inner_s);
    const float s = fmaf(-dn, dn, 1.0f);
    Sw = s * s;
    dS = (dn * s) * m4invSw;
}

// One pair inside the cutoff, branch-free (the caller masks lanes without a
// pair). sc = (Coulomb, LJ) scaling of the pair's class: (0, 0) for bond
// order <= 2, the 1-4 scalings, or (1, 1). kin / kout: Coulomb prefactors
// for near pairs (solute dielectric) and far pairs (solvent), held in
// vector registers by the caller. Returns (dE/dr)/r, the factor of
// (x_i - x_j) in dE/dx_i; adds dE/da_i to dai and sets dE/da_j in daj.
// ai, bj = (R, sqrt(eps), a, c/a) of atoms i and j.
template <bool SAME>
__device__ __forceinline__ float mix_pair(const MixK &K, float kin, float kout, float qi, const float4 &ai, float qj,
                                          const float4 &bj, float r2, float2 sc, MixE &E, float &dai, float &daj) {
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float qq = qi * qj;
    const float u2 = -(r2 * (ai.w * bj.w)); // log2(e) * u, u = -r^2 / (4 a_i a_j)
    const bool near = u2 > K.thr2;          // u above ExpApproxThreshold: GB term
    float Sw, dS;
    mix_switch(r, K.inner, K.invSw, K.inner_s, K.m4invSw, Sw, dS);
    // Coulomb: solute dielectric for near pairs, solvent for far ones,
    // scaled by the pair's class (none for bond order <= 2)
    const float ke = (near ? kin : kout) * sc.x;
    const float ep = (qq * ke) * rinv;
    // LJ in the R_min form, scaled by the class; rr is clamped (only ever
    // reached for atoms within 0.002 A of each other) so that a class
    // scaling of 0 never meets an infinite power
    float rr = fminf((ai.x + bj.x) * rinv, 1.0e3f);
    if constexpr (!SAME) rr = r < K.vcut ? rr : 0.0f;
    const float rr2 = rr * rr;
    const float B6 = rr2 * rr2 * rr2;
    const float A12 = B6 * B6;
    const float eps = (ai.y * bj.y) * sc.y;
    const float vp = eps * fmaf(-2.0f, B6, A12);   // 2 eps (A/2 - B)
    const float vfr = (eps * -12.0f) * (A12 - B6); // r dE_vdw/dr
    float vS = Sw, vd = dS;
    if constexpr (!SAME) mix_switch(r, K.vinner, K.vinvSw, K.vinner_s, K.vm4invSw, vS, vd);
    // GB: u above the threshold only (gbff.cpp:1575)
    const float aa = ai.z * bj.z;
    const float tg = __builtin_amdgcn_exp2f(u2);
    const float fgb = __builtin_amdgcn_rsqf(fmaf(tg, aa, r2));
    const float epol = near ? (qq * K.pm) * fgb : 0.0f;
    const float eL = epol * (fgb * fgb);
    const float dedr = -eL * (r * fmaf(-0.25f, tg, 1.0f));
    const float dedal = ((-eL * tg) * fmaf(-kLn2f, u2, 1.0f)) * Sw; // (1 - u), switched
    // dE/dr: switched LJ + Coulomb + 2 x switched GB
    float fr = fmaf(vS * rinv, vfr, vd * vp) + fmaf(-Sw * rinv, ep, dS * ep);
    fr = fmaf(2.0f, fmaf(Sw, dedr, dS * epol), fr);
    E.vdw = fmaf(vS, vp, E.vdw);
    E.elec = fmaf(Sw, ep, E.elec);
    E.pol = fmaf(Sw, epol, E.pol);
    dai += dedal * bj.z;
    daj = dedal * ai.z;
    return fr * rinv;
}

// Table of pair-class scalings (Coulomb, LJ), entry k held by lane k:
// entries 0-7 by list bond order (0 for orders <= 2, the 1-4 scalings for 3,
// 1 above); entries 8-11 by dense-tile bits, 8 + excluded + 2 * (1-4).
struct ClassTable {
    float e, v;
    __device__ __forceinline__ explicit ClassTable(const MixK &K) {
        const int k = lane_id();
        const bool one = (k >= 4 && k < 8) || k == 8, s14 = k == 3 || k == 10;
        e = one ? 1.0f : s14 ? K.s14e : 0.0f;
        v = one ? 1.0f : s14 ? K.s14v : 0.0f;
    }
    // entry k (k < 64, the same or another per lane); a lane permute reads
    // lanes 0-11, so every lane of the wavefront must be active
    __device__ __forceinline__ float2 at(int k) const {
        return make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(k << 2, __float_as_int(e))),
                           __int_as_float(__builtin_amdgcn_ds_bpermute(k << 2, __float_as_int(v))));
    }
};

// DPP wave_ror:1 (every lane has a source, so bound_ctrl never applies);
// this form lets the compiler fuse the move into the VALU op that uses it.
__device__ __forceinline__ float rotd(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x13C, 0xF, 0xF, true));
}

// One tile at a time per wavefront, each wavefront with its own 8 KB of
// LDS --
//   dense tile: the j block's (x,y,z,q) at [0, 128) and (R, sqrt eps, a,
//               c/a) at [128, 256), each block stored twice in a row;
//   list tile:  (x,y,z,q) of I at [0,64), J at [64,128), params at
//               [128,192), [192,256); then 4 + 4 double rows of i- and
//               j-side sums.
// (A variant with the four wavefronts of a workgroup sharing each tile was
// measured: the 8k-atom chain's pass fell from 22 to 14 us, the batch's rose
// from 0.83 to 1.34 ms; routing only the big list tiles to it was slower on
// both.)
struct MixLdsWave {
    float4 v[256];
    double f[8][kBlk];
};

// SAME: the vdW and elec/GB switching ranges coincide. ALLIN: every list
// entry lies within the cutoff (GbsaStillCutoff <= Cutoff, so the born-side
// entries do too), and only the tail of the last chunk needs masking.
template <bool SAME, bool ALLIN>
__global__ void __launch_bounds__(256) k_pair_mixed(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ MixLdsWave lds[kWavesPerGroup];
    const int t_begin = S.row_off[blk_begin], t_end = min(S.row_off[blk_end], S.plist_cap);
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    MixLdsWave &M = lds[w];
    const MixK K = mix_consts(P);
    const ClassTable sc_tab(K);
    // the two Coulomb prefactors in vector registers (as the operands of
    // one select per pair; from scalar registers each would need a move)
    float kin = K.ke_in, kout = K.ke_out;
    asm volatile("" : "+v"(kin), "+v"(kout));
    InstSink<float> sink;
    for (int c = wave_global_id() * tpw; c < u_end; c += nw * tpw) {
        const int cend = min(u_end, c + tpw);
        for (int u = c; u < cend; ++u) {
            const TileItem it = tile_item(u, t_begin, nt, lg);
            const int t = it.t;
            const int2 tw = S.tiles[t];
            const int I = tw.x, J = tw.y & 0xFFFFFF;
            const int inst = S.blk_inst[I];
            if (inst < 0) continue;
            const bool diag = (tw.y >> 24) & kTileDiag;
            const int2 cnt = S.tile_cnt[it.rec];
            const bool dense = diag || cnt.x < 0;
            const int ntot = dense ? 0 : cnt.x + cnt.y;
            if (!dense && ntot == 0) continue;
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const float4 qa = S.xyzq32[a], pa = S.pp32[a];
            float4 qb = qa, pb = pa;
            if (!diag) {
                const double4 sh = tile_shift<float>(S, I, J);
                qb = S.xyzq32[b];
                pb = S.pp32[b];
                qb.x += (float)sh.x; qb.y += (float)sh.y; qb.z += (float)sh.z;
            }
            MixE E;
            int *st = S.inst_status + inst;
            if (dense) {
                const DiagBits D = diag_bits(S, S.tile_slot[t]);
                M.v[l] = qb; M.v[l + kBlk] = qb;
                M.v[128 + l] = pb; M.v[128 + kBlk + l] = pb;
                const int ldir = rot_dir(); // 1 or 63, the same in every lane
                unsigned long long T = step_bits(D.mc, l, ldir);
                const unsigned long long X = step_bits(D.mx, l, ldir), F = step_bits(D.m4, l, ldir);
                if (diag && l >= 32) T &= ~(1ull << 31); // step 32 meets each pair twice: the lower lane takes it
                wave_lds_sync();
                float fix = 0, fiy = 0, fiz = 0, dai_acc = 0;
                float fjx = 0, fjy = 0, fjz = 0, daj_acc = 0;
                // this segment's steps k0+1 .. k1; the j atom of step k is lane
                // (l + k*ldir) & 63, read from the doubled copy
                const int steps = diag ? 32 : kBlk;
                const int k0 = seg_step(it.s, steps, lg), k1 = seg_step(it.s + 1, steps, lg);
                const int dj = ldir == 1 ? 1 : -1;
                int jx = (ldir == 1 ? l : l + kBlk) + k0 * dj;
                for (int h = k0 >> 5; h < (k1 + 31) >> 5; ++h) {
                    const unsigned int Tw = (unsigned int)(T >> (32 * h)), Xw = (unsigned int)(X >> (32 * h)),
                                       Fw = (unsigned int)(F >> (32 * h));
                    const int kk1 = min(32, k1 - 32 * h);
                    for (int kk = max(0, k0 - 32 * h); kk < kk1; ++kk) {
                        jx += dj;
                        const float4 qj = M.v[jx], pj = M.v[128 + jx];
                        // both j records are requested together, ahead of the
                        // pair mask (left to itself the compiler would sink the
                        // second read into the masked block, a second LDS round
                        // trip per step)
                        asm volatile("" ::"v"(pj.x), "v"(pj.y), "v"(pj.z), "v"(pj.w), "v"(qj.w));
                        const float dx = qa.x - qj.x, dy = qa.y - qj.y, dz = qa.z - qj.z;
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        // the j-side sums travel with their atoms on every lane
                        fjx = rotd(fjx); fjy = rotd(fjy); fjz = rotd(fjz); daj_acc = rotd(daj_acc);
                        const float2 sc = sc_tab.at(8 + (int)((Xw >> kk) & 1u) + 2 * (int)((Fw >> kk) & 1u));
                        if (((Tw >> kk) & 1u) && r2 <= K.cut2) {
                            float daj;
                            const float g = mix_pair<SAME>(K, kin, kout, qa.w, pa, qj.w, pj, r2, sc, E, dai_acc, daj);
                            const float gx = g * dx, gy = g * dy, gz = g * dz;
                            fix += gx; fiy += gy; fiz += gz;
                            fjx -= gx; fjy -= gy; fjz -= gz;
                            daj_acc += daj;
                            ++E.pairs;
                        }
                    }
                }
                // after step k1 the j-side sum of atom a sits in lane (a - k1*ldir) & 63
                fjx = rot_home(fjx, k1, ldir); fjy = rot_home(fjy, k1, ldir); fjz = rot_home(fjz, k1, ldir);
                daj_acc = rot_home(daj_acc, k1, ldir);
                if (diag) {
                    const double gx = (double)fix + (double)fjx;
                    const double gy = (double)fiy + (double)fjy;
                    const double gz = (double)fiz + (double)fjz;
                    const double gd = (double)dai_acc + (double)daj_acc;
                    if (l < S.blk_n[I]) {
                        acc_add<float>(S.acc_fx, a, gx, kForceScale, st);
                        acc_add<float>(S.acc_fy, a, gy, kForceScale, st);
                        acc_add<float>(S.acc_fz, a, gz, kForceScale, st);
                        acc_add<float>(S.acc_deda, a, gd, kForceScale, st);
                    }
                } else {
                    if (l < S.blk_n[I]) {
                        acc_add<float>(S.acc_fx, a, fix, kForceScale, st);
                        acc_add<float>(S.acc_fy, a, fiy, kForceScale, st);
                        acc_add<float>(S.acc_fz, a, fiz, kForceScale, st);
                        acc_add<float>(S.acc_deda, a, dai_acc, kForceScale, st);
                    }
                    if (l < S.blk_n[J]) {
                        acc_add<float>(S.acc_fx, b, fjx, kForceScale, st);
                        acc_add<float>(S.acc_fy, b, fjy, kForceScale, st);
                        acc_add<float>(S.acc_fz, b, fjz, kForceScale, st);
                        acc_add<float>(S.acc_deda, b, daj_acc, kForceScale, st);
                    }
                }
            } else {
                const int capm1 = (kTileCap >> lg) - 1;
                const uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
                // list entries are fetched two chunks ahead of their use
                int e_n1 = l < ntot ? (int)L[entry_pos(l, cnt.x, capm1)] : 0;
                int e_n2 = l + kBlk < ntot ? (int)L[entry_pos(l + kBlk, cnt.x, capm1)] : 0;
                M.v[l] = qa; M.v[kBlk + l] = qb;
                M.v[128 + l] = pa; M.v[192 + l] = pb;
                for (int k = 0; k < 8; ++k) M.f[k][l] = 0;
                wave_lds_sync();
                for (int q0 = 0; q0 < ntot; q0 += kBlk) {
                    const int q = q0 + l;
                    const int e = e_n1;
                    e_n1 = e_n2;
                    e_n2 = q + 2 * kBlk < ntot ? (int)L[entry_pos(q + 2 * kBlk, cnt.x, capm1)] : 0;
                    const int ii = e & 63, jj = (e >> 6) & 63, bo = (e >> 12) & 7;
                    const float4 qi = M.v[ii], qj = M.v[kBlk + jj];
                    const float4 pi = M.v[128 + ii], pj = M.v[192 + jj];
                    const float dx = qi.x - qj.x, dy = qi.y - qj.y, dz = qi.z - qj.z;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float2 sc = sc_tab.at(bo);
                    // born-side entries reach past Cutoff only when GbsaStillCutoff is larger
                    if (q < ntot && (ALLIN || r2 <= K.cut2)) {
                        float dai = 0, daj;
                        const float g = mix_pair<SAME>(K, kin, kout, qi.w, pi, qj.w, pj, r2, sc, E, dai, daj);
                        if (!ALLIN) ++E.pairs;
                        const float gx = g * dx, gy = g * dy, gz = g * dz;
                        atomicAdd(&M.f[0][ii], (double)gx);
                        atomicAdd(&M.f[1][ii], (double)gy);
                        atomicAdd(&M.f[2][ii], (double)gz);
                        atomicAdd(&M.f[3][ii], (double)dai);
                        atomicAdd(&M.f[4][jj], (double)-gx);
                        atomicAdd(&M.f[5][jj], (double)-gy);
                        atomicAdd(&M.f[6][jj], (double)-gz);
                        atomicAdd(&M.f[7][jj], (double)daj);
                    }
                }
                if (ALLIN && l == 0) E.pairs = (unsigned int)ntot;
                wave_lds_sync();
                if (l < S.blk_n[I]) {
                    acc_add<float>(S.acc_fx, a, M.f[0][l], kForceScale, st);
                    acc_add<float>(S.acc_fy, a, M.f[1][l], kForceScale, st);
                    acc_add<float>(S.acc_fz, a, M.f[2][l], kForceScale, st);
                    acc_add<float>(S.acc_deda, a, M.f[3][l], kForceScale, st);
                }
                if (l < S.blk_n[J]) {
                    acc_add<float>(S.acc_fx, b, M.f[4][l], kForceScale, st);
                    acc_add<float>(S.acc_fy, b, M.f[5][l], kForceScale, st);
                    acc_add<float>(S.acc_fz, b, M.f[6][l], kForceScale, st);
                    acc_add<float>(S.acc_deda, b, M.f[7][l], kForceScale, st);
                }
            }
            // energies: one fixed-point rounding per tile and field
            sink.to(S, I, inst);
            sink.energy(0, wave_sum((double)E.vdw));
            sink.energy(1, wave_sum((double)E.elec));
            sink.energy(2, 2.0 * wave_sum((double)E.pol));
            sink.n[kCPairs] += wave_sum(E.pairs);
            wave_lds_sync();
        }
    }
    sink.flush(S);
}

} // namespace pdgpu
