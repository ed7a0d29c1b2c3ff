// pdgpu_device.h -- gfx950 kernels of the GB / nonbonded hot path.
//
// Work decomposition. Atoms of every instance are laid out in 64-atom
// blocks (one block = one wavefront's worth of atoms; an instance starts on
// a block boundary, its last block is padded). The neighbour list is a list
// of block-pair tiles (I, J<=I) whose bounding boxes come within the list
// radius. A wavefront processes a tile with one i-atom per lane and the 64
// j-atoms held in registers, rotated one lane per step with DPP wave_ror:1,
// so every lane meets every j once and the j-side accumulators travel with
// their atom. No LDS is needed for the pair loop. Unordered pairs are
// computed once (Newton's third law); the ordered Born-radius and
// chain-rule sums get both directions from the one symmetric pair term.
//
// Precision policies. R=double is the plain FP64 path (double atomics).
// R=float is the mixed path: FP32 pair maths on block-relative coordinates,
// forces / dE/da / Born sums accumulated in 64-bit fixed point (integer adds
// commute, so results are bitwise reproducible whatever the tile order),
// energies summed per lane in FP64 and accumulated in 64-bit fixed point.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdgpu {

constexpr int kBlk = 64;             // atoms per block = wavefront width
constexpr int kWavesPerGroup = 4;    // wavefronts per workgroup of pair kernels

// tile flag bits (stored in the high byte of the J word)
constexpr int kTileDiag = 1;   // I == J: only j < i pairs
constexpr int kTileNear = 2;   // some pair has index gap < max_index_delta: band lookups
constexpr int kTileSparse = 4; // some off-diagonal bond order falls in the tile

// per-instance energy record slots; the potential energy is the sum of
// slots kEVdw .. kEImproper
enum { kEVdw = 0, kEElec, kEPolCross, kEPolSelf, kEBond, kEAngle, kETorsion, kEImproper, kEVirial, kENum = 10 };
enum { kCPairs = 0, kCBornPairs, kCNum = 2 };
// per-block record of an evaluation: energy slots kEVdw .. then the counters
constexpr int kBAcc = kENum + kCNum;
// per-instance status bits
constexpr int kStatNonFinite = 1;
constexpr int kStatOverflow = 2;

// fixed-point scales of the mixed path
constexpr double kForceScale = 4294967296.0;      // 2^32 per kcal/mol/A
constexpr double kGpolScale = 1099511627776.0;    // 2^40 per Born-sum unit
constexpr double kEnergyScale = 4294967296.0;     // 2^32 per kcal/mol
constexpr double kFixedLimit = 4.0e18;            // |value*scale| beyond this is flagged

// Uniform parameters of one context (kcal/mol, Angstrom, e).
struct PairParams {
    double cut2;           // Cutoff^2
    double inner;          // InnerCutoff
    double invSw;          // 1/(Cutoff-InnerCutoff), 0 for a zero width
    double vcut, vinner;   // VdwCutoff, VdwInnerCutoff
    double vinvSw;         // 1/(VdwCutoff-VdwInnerCutoff), 0 for a zero width
    double s14v, s14e;     // 1-4 scalings
    double econv_in;       // econv / DielectricSolute
    double econv_out;      // econv / DielectricSolvent
    float econv_in32, econv_out32; // the same rounded once, for the FP32 pair maths
    int same_switch;       // the vdW switch equals the elec/GB switch (same radii)
    double soft_h;         // FF_SoftVDW Hardness, kcal/mol
    double soft_olap;      // FF_SoftVDW OlapOffset, A
    double soft_cut2;      // FF_SoftVDW VdwCutoff^2
    double premul;         // -halfeconv*(1/eps_in - 1/eps_out)
    double still2;         // min(GbsaStillCutoff, list radius)^2
    double invP5, P4, P5;
    double exp_thresh;     // ExpApproxThreshold
    double halfeconv;      // Born radius numerator
    double born_off;       // BornRadiusOffset
    double list_r;         // list radius: max(Cutoff,VdwCutoff)+Padding
    double list_lim2;      // (max(Cutoff,VdwCutoff)+Padding)^2 as the reference forms it
    // force-switch constants of the energy path
    double fs_sA, fs_sB, fs_sC, fs_sD, fs_sE, fs_cutV, fs_eshift;
    double cut, inner2;
    int elec_mode, vdw_mode;
    int dense_pairs; // off-diagonal tiles with at least this many candidates are stored dense
    int lg_split;    // log2 of the segments each tile is split into (tile_items)
};

// Device pointers of one context.
struct DevState {
    // per atom (padded layout, cap_atoms)
    const double *px, *py, *pz;   // master positions, Angstrom
    const double *q;              // charge
    const double2 *lj64;          // (R_min/2, sqrt(eps) in sqrt(kcal/mol))
    const double2 *gb64;          // (Still type radius, V_j)
    const double *t123;           // terms123
    float4 *xyzq32;               // mixed: block-relative position + charge
    const float2 *lj32, *gb32;
    float4 *pp32;                 // mixed: (R_min/2, sqrt eps, Born radius a, c/a), c^2 = log2(e)/4
    void *born;                   // R[]: Born radii
    void *acc_g;                  // Born sums (int64 fixed / double)
    void *acc_deda;               // dE/da (int64 fixed / double)
    void *acc_fx, *acc_fy, *acc_fz; // dE/dx (int64 fixed / double)
    double *fx, *fy, *fz;         // finalized forces, kcal/mol/A (= -dE/dx)
    // per block
    const int *blk_inst;          // instance slot or -1
    const int *blk_n;             // valid atoms in block
    const int *blk_first;         // first block of the instance
    double4 *blk_ctr;             // reference centre (x,y,z) of the block
    float4 *blk_half;             // bounding-box half extents (x,y,z)
    double4 *blk_box;             // bounding-box centre (x,y,z)
    int *row_count;               // tiles per I row
    int *row_off;                 // exclusive scan of row_count (+ total at [nblocks])
    int2 *tiles;                  // (I, J | flags<<24)
    int *tile_slot;               // candidate slot of each tile
    // per instance
    const int *inst_mid;
    const long long *inst_band_off;
    const int *inst_sp_off, *inst_sp_n;
    const int8_t *band;
    const int4 *sparse;           // (i_local, j_local, order, 0), i<j
    const int *inst_need;         // optional per-instance rebuild flags (null: all)
    const int *any_need;          // optional: 0 means no instance is flagged (skip the rebuild)
    double *rx, *ry, *rz;         // positions at each atom's last list rebuild
    int *inst_moved;              // set when an atom moved more than half the padding since
    uint16_t *plist;              // per-tile compacted pair lists, kTileCap entries each
    int2 *tile_cnt;               // (born-side count, other count) per tile
    int plist_cap;                // tiles the list buffer holds
    int *overflow;                // set when a tile index exceeds plist_cap
    uint16_t *cand;               // candidate pairs of the last rebuild, kTileCap per tile slot
    int *cand_n;                  // candidates per tile slot
    const long long *inst_cslot;  // first tile slot of each instance's candidate region
    void *inst_energy;            // [slot*kENum]
    unsigned long long *inst_count; // [slot*kCNum]
    int *inst_status;
    // per block [block*kBAcc]: the evaluation's partial energies (fixed point
    // or double bits, per the precision) and pair counts of the tiles of its
    // I row and of its bonded terms, folded into the instance record by the
    // evaluation's last kernel
    unsigned long long *blk_acc;
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int rot_i(int v) {
    // wave_ror:1; every lane has a source, so the 'old' operand is never
    // selected and passing v lets the result reuse v's register
    return __builtin_amdgcn_update_dpp(v, v, 0x13C, 0xF, 0xF, false);
}
__device__ __forceinline__ float rot(float v) { return __int_as_float(rot_i(__float_as_int(v))); }
// Lane offset of one rotation, 1 or 63: after k rotations lane l holds the
// value of lane (l + k*rot_dir()) & 63. The same in every lane; read from
// lane 0 so that the compiler treats it (and what derives from it) as
// wavefront-uniform.
__device__ __forceinline__ int rot_dir() {
    return __builtin_amdgcn_readfirstlane((rot_i(lane_id()) - lane_id()) & 63);
}
// Wavefront index within the workgroup, as a wavefront-uniform value (the
// compiler's divergence analysis does not know that threadIdx.x / 64 is
// uniform, so tile indices, records and loop bounds derived from it would
// otherwise live in vector registers and branch through exec masks).
__device__ __forceinline__ int wave_in_group() { return __builtin_amdgcn_readfirstlane(threadIdx.x / 64); }
__device__ __forceinline__ double rot(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = rot_i((int)(b & 0xffffffffLL));
    const int hi = rot_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ int rot(int v) { return rot_i(v); }

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// precision traits
template <typename R> struct Prec;
template <> struct Prec<double> {
    static constexpr bool fixed = false;
    // 1/sqrt(x) and 1/x from the hardware estimates (relative error about
    // 2^-23) refined by two Newton steps in FMA form: within an ulp or two of
    // the correctly rounded values for the positive normal arguments of the
    // pair terms (squared distances, radius products). The IEEE square root
    // and division sequences cost ~25 and ~10 instructions.
    __device__ static double rsqrt(double x) {
        double y = __builtin_amdgcn_rsq(x);
        const double h = 0.5 * x;
        double e = fma(-(h * y), y, 0.5);
        y = fma(y, e, y);
        e = fma(-(h * y), y, 0.5);
        return fma(y, e, y);
    }
    __device__ static double rcp(double x) {
        double y = __builtin_amdgcn_rcp(x);
        double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
        e = fma(-x, y, 1.0);
        return fma(y, e, y);
    }
    __device__ static double expo(double x) { return exp(x); }
    // cos/sin of th = pi * f (f in [0,1])
    __device__ static double cospi_(double f) { return cospi(f); }
    __device__ static double sinpi_(double f) { return sinpi(f); }
};
template <> struct Prec<float> {
    static constexpr bool fixed = true;
    __device__ static float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static float expo(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
    // v_cos_f32 / v_sin_f32 take revolutions: cos(pi f) = cos(2 pi (f/2))
    __device__ static float cospi_(float f) { return __builtin_amdgcn_cosf(0.5f * f); }
    __device__ static float sinpi_(float f) { return __builtin_amdgcn_sinf(0.5f * f); }
};

// Accumulate v into an 8-byte slot: fixed point (mixed) or double (FP64).
// Round-to-nearest-even double -> int64 (what __double2ll_rn gives). Below
// 2^51 in magnitude, adding 1.5 * 2^52 leaves the rounded integer in the
// low mantissa bits, two instructions instead of the general sequence.
__device__ __forceinline__ long long to_fixed(double s) {
    if (fabs(s) < 0x1p51) {
        const double m = 0x1.8p52;
        return __double_as_longlong(s + m) - __double_as_longlong(m);
    }
    return __double2ll_rn(s);
}

template <typename R>
__device__ __forceinline__ void acc_add(void *base, size_t idx, double v, double scale, int *status) {
    if constexpr (Prec<R>::fixed) {
        const double s = v * scale;
        if (!(fabs(s) < kFixedLimit)) {
            atomicOr(status, (s == s) ? kStatOverflow : kStatNonFinite);
            return;
        }
        const long long iv = to_fixed(s);
        if (iv != 0) atomicAdd(reinterpret_cast<unsigned long long *>(base) + idx, (unsigned long long)iv);
    } else {
        if (!(fabs(v) < 1e300)) atomicOr(status, kStatNonFinite);
        if (v != 0.0) unsafeAtomicAdd(reinterpret_cast<double *>(base) + idx, v);
    }
}

template <typename R>
__device__ __forceinline__ double acc_read(const void *base, size_t idx, double scale) {
    if constexpr (Prec<R>::fixed) {
        return (double)reinterpret_cast<const long long *>(base)[idx] / scale;
    } else {
        return reinterpret_cast<const double *>(base)[idx];
    }
}

template <typename R>
__device__ __forceinline__ void acc_store(void *base, size_t idx, double v, double scale) {
    if constexpr (Prec<R>::fixed) {
        const double s = v * scale;
        reinterpret_cast<long long *>(base)[idx] = (fabs(s) < kFixedLimit) ? to_fixed(s) : 0;
    } else {
        reinterpret_cast<double *>(base)[idx] = v;
    }
}

// Bond order of (il, jl), il > jl, local indices of instance inst
// (workspace/neighbourlist.cpp:281-289 band rule, :162-203 off-diagonal patch).
__device__ __forceinline__ int pair_bondorder(const DevState &S, int inst, int il, int jl, int flags) {
    const int d = il - jl;
    const int mid = S.inst_mid[inst];
    if (d < mid) return S.band[S.inst_band_off[inst] + (long long)jl * mid + (d - 1)] & 7;
    int bo = 7;
    if (flags & kTileSparse) {
        const int off = S.inst_sp_off[inst], n = S.inst_sp_n[inst];
        for (int k = 0; k < n; ++k) {
            const int4 e = S.sparse[off + k];
            if (e.x == jl && e.y == il) bo = e.z & 7;
        }
    }
    return bo;
}

// The same lookup with the band rows of the tile's J block staged in LDS:
// sb holds rows jl0 .. jl0+63 (mid entries each) of the instance's band.
__device__ __forceinline__ int pair_bondorder_lds(const DevState &S, int inst, int il, int jl, int flags, int mid,
                                                  const int8_t *sb, int jl0) {
    const int d = il - jl;
    if (d < mid) return sb[(jl - jl0) * mid + (d - 1)] & 7;
    return pair_bondorder(S, inst, il, jl, flags & kTileSparse);
}

// Work items of the tile kernels. With PairParams::lg_split = g every tile
// is split into 2^g segments, each processed by one wavefront on its own:
//  - candidates: segment s owns candidates [s*c, (s+1)*c) of the tile's
//    candidate slot, c being the count over 2^g rounded up to a multiple of
//    4 (so that 8-byte candidate loads stay aligned);
//  - lists: the region [s*kTileCap/2^g, (s+1)*kTileCap/2^g) of the tile's
//    list, into which k_filter compacts that segment's candidates, with its
//    own (born-side, other) counts in tile_cnt[t*2^g + s];
//  - dense tiles: rotation steps [s*K/2^g, (s+1)*K/2^g) of the tile's K.
// Items are numbered segment-major: item u of a launch over nt tiles from
// t_begin is segment u / nt of tile t_begin + u % nt, so the chunk of
// consecutive items a wavefront takes holds different tiles and a tile's
// segments go to different wavefronts. g = 0 is one wavefront per tile. A
// larger g shortens the longest chain of dependent work of a launch (a full
// list tile is 64 steps of 64 pairs) for systems with too few tiles to fill
// the GPU; per-segment sums are flushed on their own, so results depend on g
// (deterministically) and not on the grid. Per-segment list counts are kept
// at tile_cnt[(t << g) + s].
struct TileItem {
    int t, s;
    size_t rec; // index of the segment's tile_cnt record
};
__device__ __forceinline__ TileItem tile_item(int u, int t_begin, int nt, int lg) {
    if (lg == 0) return {t_begin + u, 0, (size_t)(t_begin + u)};
    const int s = u / nt, t = t_begin + (u - s * nt);
    return {t, s, ((size_t)t << lg) + s};
}

// First and one-past-last rotation step of segment s of a dense tile with
// `steps` steps.
__device__ __forceinline__ int seg_step(int s, int steps, int lg) { return (s * steps) >> lg; }

// v as lane (l + k*ldir) & 63 holds it, i.e. after k rotations
template <typename T> __device__ __forceinline__ T rot_to(T v, int k, int ldir) {
    return (k * ldir) & 63 ? __shfl(v, (lane_id() + k * ldir) & 63, 64) : v;
}
// j-side sums travelling with their atoms for k rotations, brought back to
// the atom's own lane
template <typename T> __device__ __forceinline__ T rot_home(T v, int k, int ldir) {
    return (k * ldir) & 63 ? __shfl(v, (lane_id() - k * ldir) & 63, 64) : v;
}

// Persistent-grid chunk assignment. Consecutive chunks go to workgroups on
// the same XCD (workgroups are dealt round-robin over 8 XCDs), so the j-atom
// data of neighbouring tiles is shared through one L2.
// ngroups: the workgroups taking part (the first ngroups of the grid).
__device__ __forceinline__ int wave_global_id(int ngroups) {
    const int nx = 8;
    int b = blockIdx.x;
    if ((ngroups % nx) == 0) b = (blockIdx.x % nx) * (ngroups / nx) + blockIdx.x / nx;
    return b * kWavesPerGroup + wave_in_group();
}
__device__ __forceinline__ int wave_global_id() { return wave_global_id(gridDim.x); }

// ------------------------------------------------------ per-instance setup
// Lane a % 64 < kBAcc of a block-aligned per-atom launch clears one field of
// its block's record.
__device__ __forceinline__ void blk_clear(const DevState &S, int a) {
    const int f = a & (kBlk - 1);
    if (f < kBAcc) S.blk_acc[(size_t)(a / kBlk) * kBAcc + f] = 0ull;
}

// Adds block b's record into its instance's (one wavefront per block, in the
// evaluation's last kernel): at most kBAcc atomics per block instead of one
// per wavefront of every pair kernel on the same few addresses.
template <typename R> __device__ __forceinline__ void blk_fold(const DevState &S, int b, int inst) {
    const int f = lane_id();
    if (inst < 0 || f >= kBAcc) return;
    const unsigned long long v = S.blk_acc[(size_t)b * kBAcc + f];
    if (v == 0ull) return;
    if (f >= kENum) {
        atomicAdd(S.inst_count + (size_t)inst * kCNum + (f - kENum), v);
    } else if constexpr (Prec<R>::fixed) {
        atomicAdd(reinterpret_cast<unsigned long long *>(S.inst_energy) + (size_t)inst * kENum + f, v);
    } else {
        unsafeAtomicAdd(reinterpret_cast<double *>(S.inst_energy) + (size_t)inst * kENum + f,
                        __longlong_as_double((long long)v));
    }
}

__device__ __forceinline__ void inst_clear(const DevState &S, int s) {
    unsigned long long *e = reinterpret_cast<unsigned long long *>(S.inst_energy) + (size_t)s * kENum;
    for (int k = 0; k < kENum; ++k) e[k] = 0;
    for (int k = 0; k < kCNum; ++k) S.inst_count[(size_t)s * kCNum + k] = 0;
    S.inst_status[s] = 0;
}


// --------------------------------------------------- neighbour-list build
// Bounding box and reference centre of each block (one wave per block).
// A flagged rebuild with no flagged instance does nothing.
__device__ __forceinline__ bool build_idle(const DevState &S) { return S.any_need && *S.any_need == 0; }

__global__ void k_block_bounds(DevState S, int blk_begin, int blk_end) {
    if (build_idle(S)) return;
    const int b = blk_begin + blockIdx.x * (blockDim.x / 64) + wave_in_group();
    if (b >= blk_end) return;
    if (S.blk_inst[b] < 0) return;
    if (S.inst_need && !S.inst_need[S.blk_inst[b]]) return;
    const int l = lane_id();
    const bool v = l < S.blk_n[b];
    const size_t a = (size_t)b * kBlk + l;
    double x = v ? S.px[a] : 0, y = v ? S.py[a] : 0, z = v ? S.pz[a] : 0;
    S.rx[a] = x; S.ry[a] = y; S.rz[a] = z; // reference for the displacement test
    double mnx = v ? x : 1e300, mny = v ? y : 1e300, mnz = v ? z : 1e300;
    double mxx = v ? x : -1e300, mxy = v ? y : -1e300, mxz = v ? z : -1e300;
    for (int o = 32; o > 0; o >>= 1) {
        mnx = fmin(mnx, __shfl_xor(mnx, o, 64)); mxx = fmax(mxx, __shfl_xor(mxx, o, 64));
        mny = fmin(mny, __shfl_xor(mny, o, 64)); mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
        mnz = fmin(mnz, __shfl_xor(mnz, o, 64)); mxz = fmax(mxz, __shfl_xor(mxz, o, 64));
    }
    if (l == 0) {
        const double cx = 0.5 * (mnx + mxx), cy = 0.5 * (mny + mxy), cz = 0.5 * (mnz + mxz);
        S.blk_box[b] = make_double4(cx, cy, cz, 0.0);
        S.blk_ctr[b] = make_double4(cx, cy, cz, 0.0);
        // half extents rounded up, plus a margin that covers float rounding
        const float m = 1e-3f;
        S.blk_half[b] = make_float4((float)(0.5 * (mxx - mnx)) + m, (float)(0.5 * (mxy - mny)) + m,
                                    (float)(0.5 * (mxz - mnz)) + m, 0.f);
    }
}

// Verlet-skin test: flags an instance when one of its atoms has moved more
// than half the list padding since the instance's last rebuild (then a pair
// from beyond the list radius may have come inside the cutoff).
__global__ void k_moved(DevState S, double lim2, int blk_begin, int blk_end) {
    const int b = blk_begin + blockIdx.x * (blockDim.x / 64) + wave_in_group();
    if (b >= blk_end) return;
    const int inst = S.blk_inst[b];
    if (inst < 0) return;
    const int l = lane_id();
    const size_t a = (size_t)b * kBlk + l;
    bool far = false;
    if (l < S.blk_n[b]) {
        const double dx = S.px[a] - S.rx[a], dy = S.py[a] - S.ry[a], dz = S.pz[a] - S.rz[a];
        far = !(dx * dx + dy * dy + dz * dz <= lim2); // non-finite positions count as moved
    }
    if (__ballot(far) && l == 0) atomicOr(S.inst_moved + inst, 1);
}

__device__ __forceinline__ bool boxes_within(const DevState &S, int I, int J, double r) {
    const double4 ci = S.blk_box[I], cj = S.blk_box[J];
    const float4 hi = S.blk_half[I], hj = S.blk_half[J];
    const double dx = fmax(0.0, fabs(ci.x - cj.x) - (double)hi.x - (double)hj.x);
    const double dy = fmax(0.0, fabs(ci.y - cj.y) - (double)hi.y - (double)hj.y);
    const double dz = fmax(0.0, fabs(ci.z - cj.z) - (double)hi.z - (double)hj.z);
    return dx * dx + dy * dy + dz * dz <= r * r;
}

__device__ __forceinline__ int tile_flags(const DevState &S, int inst, int I, int J) {
    const int first = S.blk_first[I];
    int f = 0;
    if (I == J) f |= kTileDiag;
    const int gap = (I - J) * kBlk - (kBlk - 1); // smallest il - jl over the tile
    if (gap < S.inst_mid[inst]) f |= kTileNear;
    const int n = S.inst_sp_n[inst];
    if (n > 0) {
        const int off = S.inst_sp_off[inst];
        const int i0 = (I - first) * kBlk, j0 = (J - first) * kBlk;
        for (int k = 0; k < n; ++k) {
            const int4 e = S.sparse[off + k];
            if (e.x >= j0 && e.x < j0 + kBlk && e.y >= i0 && e.y < i0 + kBlk) { f |= kTileSparse; break; }
        }
    }
    return f;
}

// pass 0: count tiles per I row; pass 1: write them at row_off[I].
template <int PASS>
__global__ void k_tile_rows(DevState S, PairParams P, int blk_begin, int blk_end) {
    if (build_idle(S)) return;
    const int I = blk_begin + blockIdx.x * (blockDim.x / 64) + wave_in_group();
    if (I >= blk_end) return;
    const int inst = S.blk_inst[I];
    const int l = lane_id();
    if (inst < 0) {
        if (PASS == 0 && l == 0) S.row_count[I] = 0;
        return;
    }
    if (PASS == 0 && S.inst_need && !S.inst_need[inst]) return; // keep this instance's rows
    const int first = S.blk_first[I];
    const int base = (PASS == 1) ? S.row_off[I] : 0;
    const int reserved = (PASS == 1) ? S.row_count[I] : 0; // never write past the counted row
    // candidate slot of this row's first tile (tile index relative to the
    // instance's first row, offset by the instance's region)
    const long long slot0 = (PASS == 1) ? S.inst_cslot[inst] + (base - S.row_off[first]) : 0;
    int total = 0;
    for (int J0 = first; J0 <= I; J0 += 64) {
        const int J = J0 + l;
        const bool hit = (J <= I) && boxes_within(S, I, J, P.list_r);
        const unsigned long long m = __ballot(hit);
        if (PASS == 1 && hit) {
            const int local = total + (int)__popcll(m & ((1ull << l) - 1ull));
            if (local < reserved) {
                S.tiles[base + local] = make_int2(I, J | (tile_flags(S, inst, I, J) << 24));
                S.tile_slot[base + local] = (int)(slot0 + local);
            }
        }
        total += (int)__popcll(m);
    }
    if (PASS == 0 && l == 0) S.row_count[I] = total;
}

// Exclusive scan of row_count over [0, n) into row_off, total at row_off[n].
// One workgroup of 1024 threads; deterministic.
__global__ void __launch_bounds__(1024) k_scan_rows(DevState S, int n) {
    if (build_idle(S)) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = t * per, b1 = min(n, b0 + per);
    int s = 0;
    for (int i = b0; i < b1; ++i) s += S.row_count[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b0; i < b1; ++i) { S.row_off[i] = run; run += S.row_count[i]; }
    if (t == 1023) S.row_off[n] = part[1023];
}

// ----------------------------------------------------- per-atom preparation
// Mixed path: block-relative FP32 coordinates. Also clears the Born sums.
// Block-relative FP32 copy of the positions (mixed path), against the block
// centres of the instance's last rebuild.
__device__ __forceinline__ void store_xyzq32(const DevState &S, int a) {
    const int b = a / kBlk;
    const double4 c = S.blk_ctr[b];
    S.xyzq32[a] = make_float4((float)(S.px[a] - c.x), (float)(S.py[a] - c.y), (float)(S.pz[a] - c.z),
                              (float)S.q[a]);
}

// The same for the coordinates only, where the charge in .w is already
// there: every list build writes the whole record (k_xyzq32) and charges
// only change with an upload, which is followed by a build.
__device__ __forceinline__ void store_xyz32(const DevState &S, int a) {
    const int b = a / kBlk;
    const double4 c = S.blk_ctr[b];
    float *p = reinterpret_cast<float *>(S.xyzq32 + a);
    p[0] = (float)(S.px[a] - c.x);
    p[1] = (float)(S.py[a] - c.y);
    p[2] = (float)(S.pz[a] - c.z);
}

__global__ void k_xyzq32(DevState S, int atom_begin, int atom_end) {
    if (build_idle(S)) return;
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (a < atom_end) store_xyzq32(S, a);
}

template <typename R>
__global__ void k_prepare(DevState S, int atom_begin, int atom_end, int slot_begin, int slot_end) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot_begin + t < slot_end) inst_clear(S, slot_begin + t); // the per-instance records too
    const int a = atom_begin + t;
    if (a >= atom_end) return;
    blk_clear(S, a);
    acc_store<R>(S.acc_g, a, 0.0, 1.0);
    if constexpr (Prec<R>::fixed) store_xyz32(S, a);
}

// Per-lane view of a tile's atoms.
template <typename R> struct AtomXYZ { R x, y, z, q; };

template <typename R>
__device__ __forceinline__ AtomXYZ<R> load_xyzq(const DevState &S, size_t a, const double4 &shift) {
    AtomXYZ<R> r;
    if constexpr (Prec<R>::fixed) {
        const float4 v = S.xyzq32[a];
        r.x = v.x + (float)shift.x; r.y = v.y + (float)shift.y; r.z = v.z + (float)shift.z; r.q = v.w;
    } else {
        r.x = S.px[a]; r.y = S.py[a]; r.z = S.pz[a]; r.q = S.q[a];
    }
    return r;
}

// FP32 path: j coordinates are re-expressed relative to block I's centre.
template <typename R>
__device__ __forceinline__ double4 tile_shift(const DevState &S, int I, int J) {
    if constexpr (Prec<R>::fixed) {
        const double4 ci = S.blk_ctr[I], cj = S.blk_ctr[J];
        return make_double4(cj.x - ci.x, cj.y - ci.y, cj.z - ci.z, 0.0);
    } else {
        return make_double4(0, 0, 0, 0);
    }
}

template <typename R> __device__ __forceinline__ R ld_born(const DevState &S, size_t a) {
    return reinterpret_cast<const R *>(S.born)[a];
}

// ------------------------------------------------------ pair enumeration
// Two stages, the way PD's list is used: candidates are enumerated when the
// list is rebuilt, and every evaluation selects from them the pairs inside
// its cutoffs.
//
// k_cand (at a rebuild) walks each tile of the rebuilt instances once:
// lane = i-atom, the 64 j-atoms rotate through the lanes (DPP wave_ror:1),
// and pairs with r^2 within the list radius (Cutoff + Padding, as PD's list
// holds them) are compacted with ballot + mbcnt into the tile's candidate
// slot: entry = i_lane | j_lane << 6 | bond_order << 12. The slot of the
// k-th tile of an instance is inst_cslot + k; an instance keeps its tile
// order until its own next rebuild, so its candidates stay valid while other
// instances rebuild.
//
// k_filter (every evaluation) reads the candidates 64 at a time from LDS-
// staged tile coordinates and compacts, in the same order, the born /
// chain-rule pairs (bond order > 2, r <= Still cutoff) into the front of the
// tile's pair list and the remaining pair-pass pairs (r <= Cutoff) into its
// back. The compute kernels then work through these lists 64 pairs at a time
// with every lane busy.
//
// Per-tile sums of the compute kernels are gathered in LDS with atomic adds
// before one flush to the global accumulators. They are double in both
// precisions: on gfx950 the LDS FP32 atomic add (ds_add_f32) measured ~60x
// more LDS wait cycles than ds_add_f64 for the same access pattern (pair
// kernel 2.3 ms -> 0.39 ms on 1024 conformers).
constexpr int kTileCap = 4096;
constexpr int kMaxSplit = 16; // segments per tile at most (256-entry list regions)

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long lanes_below() {
    const int l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}

// Candidate slot of tile t (of an instance whose first block is `first`).
__device__ __forceinline__ long long cand_slot(const DevState &S, int inst, int first, int t) {
    return S.inst_cslot[inst] + (t - S.row_off[first]);
}

// Band rows staged in LDS per near tile: 64 rows of up to kBandStage
// entries. Wider bands (max_index_delta > kBandStage) are read from global
// memory.
constexpr int kBandStage = 64;

template <typename R>
__global__ void __launch_bounds__(256) k_cand(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ int8_t sband[kWavesPerGroup][kBlk * kBandStage];
    if (build_idle(S)) return;
    const int t_begin = S.row_off[blk_begin], t_end = S.row_off[blk_end];
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (t_end - t_begin + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    // the list radius, widened by a relative 1e-5 so that FP32 rounding never
    // drops a pair the cutoff tests of the evaluation would keep
    const R reach2 = (R)(P.list_lim2 * (1.0 + 1e-5));
    const unsigned long long below = lanes_below();
    for (int c = t_begin + wave_global_id() * tpw; c < t_end; c += nw * tpw) {
        const int cend = min(t_end, c + tpw);
        for (int t = c; t < cend; ++t) {
            const int2 tw = S.tiles[t];
            const int I = tw.x, J = tw.y & 0xFFFFFF, fl = tw.y >> 24;
            const int inst = S.blk_inst[I], first = S.blk_first[I];
            if (inst < 0 || (S.inst_need && !S.inst_need[inst])) continue;
            const long long slot = cand_slot(S, inst, first, t);
            const int il = (I - first) * kBlk + l;
            const bool vi = l < S.blk_n[I];
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            R xj = pj.x, yj = pj.y, zj = pj.z;
            const int nJ = S.blk_n[J];
            // lane offset of one rotation (the same for every lane): after k
            // rotations lane l holds the j-atom of lane (l + k*ldir) & 63
            const int ldir = rot_dir();
            const bool diag = fl & kTileDiag;
            const bool lookup = fl & (kTileNear | kTileSparse);
            const int mid = S.inst_mid[inst], jl0 = (J - first) * kBlk;
            // rows jl0.. of the band: the instance's band is allocated for every
            // padded atom, so all 64 rows are in the buffer
            const bool staged = (fl & kTileNear) && mid <= kBandStage;
            if (staged) {
                const int8_t *src = S.band + S.inst_band_off[inst] + (long long)jl0 * mid;
                for (int k = l; k < kBlk * mid; k += kBlk) sband[w][k] = src[k];
                wave_lds_sync();
            }
            uint16_t *L = S.cand + slot * kTileCap;
            if (diag) {
                // dense tile: per-lane bit rows over all 64 partners (both
                // orders) -- candidate, excluded (bond order <= 2), 1-4 pair
                unsigned long long mc = 0, mx = 0, m4 = 0;
                for (int k = 1; k < kBlk; ++k) {
                    xj = rot(xj); yj = rot(yj); zj = rot(zj);
                    const int jlane = (l + k * ldir) & 63;
                    const R dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    if (vi && jlane < nJ && r2 <= reach2) {
                        const int jl = jl0 + jlane, hi = max(il, jl), lo = min(il, jl);
                        const int bo = staged ? pair_bondorder_lds(S, inst, hi, lo, fl, mid, sband[w], jl0)
                                              : pair_bondorder(S, inst, hi, lo, fl);
                        const unsigned long long bit = 1ull << jlane;
                        mc |= bit;
                        if (bo <= 2) mx |= bit;
                        if (bo == 3) m4 |= bit;
                    }
                }
                unsigned long long *M = reinterpret_cast<unsigned long long *>(L);
                M[l] = mc;
                M[kBlk + l] = mx;
                M[2 * kBlk + l] = m4;
                if (l == 0) S.cand_n[slot] = 0;
                wave_lds_sync();
                continue;
            }
            int n = 0;
            unsigned long long mc = 0, mx = 0, m4 = 0;
            for (int k = 0; k < kBlk; ++k) {
                if (k > 0) { xj = rot(xj); yj = rot(yj); zj = rot(zj); }
                const int jlane = (l + k * ldir) & 63;
                const R dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                const R r2 = dx * dx + dy * dy + dz * dz;
                const bool ok = vi && jlane < nJ && r2 <= reach2;
                int bo = 7;
                if (ok && lookup) {
                    const int jl = jl0 + jlane;
                    bo = staged ? pair_bondorder_lds(S, inst, il, jl, fl, mid, sband[w], jl0)
                                : pair_bondorder(S, inst, il, jl, fl);
                }
                const unsigned long long m = __ballot(ok);
                if (ok) {
                    L[n + (int)__popcll(m & below)] = (uint16_t)(l | (jlane << 6) | (bo << 12));
                    const unsigned long long bit = 1ull << jlane;
                    mc |= bit;
                    if (bo <= 2) mx |= bit;
                    if (bo == 3) m4 |= bit;
                }
                n += (int)__popcll(m);
            }
            if (n >= P.dense_pairs) {
                // dense enough that lane rotation beats list + LDS atomics:
                // keep the bit rows instead (cand_n < 0 marks the format)
                wave_lds_sync();
                unsigned long long *M = reinterpret_cast<unsigned long long *>(L);
                M[l] = mc;
                M[kBlk + l] = mx;
                M[2 * kBlk + l] = m4;
                n = -1;
            }
            if (l == 0) S.cand_n[slot] = n;
            wave_lds_sync();
        }
    }
}

template <typename R> struct PosR { R x, y, z; };

// r^2 of a candidate. FP32: the mixed pair pass's own operation order, so
// that the r^2 it computes for a kept entry is exactly the one tested here.
template <typename R> __device__ __forceinline__ R filter_r2(R dx, R dy, R dz) {
    if constexpr (Prec<R>::fixed) return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    else return dx * dx + dy * dy + dz * dz;
}

template <typename R>
__global__ void __launch_bounds__(256) k_filter(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ PosR<R> si[kWavesPerGroup][kBlk], sj[kWavesPerGroup][kBlk];
    const int t_begin = S.row_off[blk_begin], t_end = S.row_off[blk_end];
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    const R cut2 = (R)P.cut2, still2 = (R)P.still2;
    const unsigned long long below = lanes_below();
    for (int c = wave_global_id() * tpw; c < u_end; c += nw * tpw) {
        const int cend = min(u_end, c + tpw);
        for (int u = c; u < cend; ++u) {
            const TileItem it = tile_item(u, t_begin, nt, lg);
            const int t = it.t;
            // the tile record and its candidate slot come first; the slot's
            // count and first 256 candidates (every slot holds kTileCap
            // entries, and tile_slot only ever names allocated slots) are
            // requested together with the block's records. A lane reads four
            // consecutive 2-byte candidates per 8-byte load, so a tile of up
            // to 256 candidates is one round trip.
            const int2 tw = S.tiles[t];
            const long long slot = S.tile_slot[t];
            const unsigned long long *C4 =
                reinterpret_cast<const unsigned long long *>(S.cand + slot * kTileCap);
            unsigned long long v_nx = C4[l];
            const int n = S.cand_n[slot];
            const int I = tw.x, J = tw.y & 0xFFFFFF;
            const int inst = S.blk_inst[I];
            if (inst < 0 || t >= S.plist_cap) {
                if (l == 0 && t < S.plist_cap) S.tile_cnt[it.rec] = make_int2(0, 0);
                if (l == 0 && t >= S.plist_cap) atomicOr(S.overflow, 1);
                continue;
            }
            if ((tw.y >> 24) & kTileDiag) { // dense tiles are evaluated from their bit rows
                if (l == 0) S.tile_cnt[it.rec] = make_int2(0, 0);
                continue;
            }
            if (n < 0) { // dense off-diagonal tile
                if (l == 0) S.tile_cnt[it.rec] = make_int2(-1, -1);
                continue;
            }
            // this segment's candidates [c0, c1) and list region
            const int segc = (((n + (1 << lg) - 1) >> lg) + 3) & ~3;
            const int c0 = it.s * segc, c1 = min(n, c0 + segc);
            if (c0 >= c1) {
                if (l == 0) S.tile_cnt[it.rec] = make_int2(0, 0);
                continue;
            }
            if (c0 > 0 && (c0 >> 2) + l < kTileCap / 4) v_nx = C4[(c0 >> 2) + l];
            const int capm1 = (kTileCap >> lg) - 1;
            uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            si[w][l] = {pi.x, pi.y, pi.z};
            sj[w][l] = {pj.x, pj.y, pj.z};
            wave_lds_sync();
            int nb = 0, no = 0;
            // chunks of 256 candidates: lane l loads 4l .. 4l+3 (loads that
            // would pass the end of the slot are not issued; their lanes only
            // hand out candidates at or beyond c1) and step k hands lane l
            // candidate q0 + 64k + l through a lane permute, so entries are
            // taken in candidate order: within 64 consecutive list entries
            // (one rotation step of the enumeration) no i or j atom repeats,
            // which keeps the later passes' LDS sums free of same-address
            // conflicts
            for (int q0 = c0; q0 < c1; q0 += 4 * kBlk) {
                const unsigned long long v = v_nx;
                if (q0 + 4 * kBlk < c1 && ((q0 + 4 * kBlk) >> 2) + l < kTileCap / 4)
                    v_nx = C4[((q0 + 4 * kBlk) >> 2) + l];
                const unsigned int vlo = (unsigned int)v, vhi = (unsigned int)(v >> 32);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int src = 16 * k + (l >> 2);
                    const unsigned int lo = __shfl(vlo, src, 64), hi = __shfl(vhi, src, 64);
                    const unsigned int word = (l & 2) ? hi : lo;
                    const int e = (int)((word >> (16 * (l & 1))) & 0xFFFFu);
                    const int q = q0 + kBlk * k + l;
                    const bool in = q < c1;
                    const PosR<R> A = si[w][e & 63], B = sj[w][(e >> 6) & 63];
                    const R dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    const R r2 = filter_r2<R>(dx, dy, dz);
                    const bool born = in && ((e >> 12) & 7) > 2 && r2 <= still2;
                    const bool oth = in && !born && r2 <= cut2;
                    const unsigned long long mb = __ballot(born), mo = __ballot(oth);
                    if (born) L[nb + (int)__popcll(mb & below)] = (uint16_t)e;
                    if (oth) L[capm1 - (no + (int)__popcll(mo & below))] = (uint16_t)e;
                    nb += (int)__popcll(mb);
                    no += (int)__popcll(mo);
                }
            }
            if (l == 0) S.tile_cnt[it.rec] = make_int2(nb, no);
            wave_lds_sync();
        }
    }
}

// Position of the q-th entry of a list region of capm1 + 1 entries: born-side
// entries from the front, the others from the back.
__device__ __forceinline__ int entry_pos(int q, int nb, int capm1) { return q < nb ? q : capm1 - (q - nb); }

// Per-lane energy sums over one tile, in the pass's precision (a lane adds
// at most 64 terms before the tile's sums are reduced in FP64).
template <typename R> struct LaneEnergy {
    R vdw = 0, elec = 0, pol = 0, vir = 0;
    unsigned int pairs = 0;
};

// Per-wavefront running sums of the instance-level results of the pair
// passes (energies, pair counts) over the tiles the wavefront processes,
// kept per I block: they are added to the block's record (blk_acc) with one
// atomic per field when the wavefront moves on to a tile of another row and
// at its end, and the records are folded into the instance's by the
// evaluation's last kernel. One atomic per field per wavefront on the
// instance record itself serialises a single large instance (thousands of
// wavefronts on the same few addresses). In fixed point every tile's energy
// is rounded on its own before it is added, exactly as a per-tile atomic
// would round it, so totals do not depend on how tiles are spread over
// wavefronts.
template <typename R> struct InstSink {
    using Acc = typename std::conditional<Prec<R>::fixed, long long, double>::type;
    int blk = -1, inst = -1;
    Acc e[4] = {0, 0, 0, 0}; // vdw, elec, pol cross, virial
    unsigned long long n[kCNum] = {0, 0};
    int bad = 0;

    __device__ __forceinline__ void to(const DevState &S, int next_blk, int next_inst) {
        if (next_blk != blk) {
            flush(S);
            blk = next_blk;
            inst = next_inst;
        }
    }
    __device__ __forceinline__ void energy(int k, double v) {
        if (v == 0.0) return;
        if constexpr (Prec<R>::fixed) {
            const double sv = v * kEnergyScale;
            if (!(fabs(sv) < kFixedLimit)) {
                bad |= (sv == sv) ? kStatOverflow : kStatNonFinite;
                return;
            }
            e[k] += to_fixed(sv);
        } else {
            if (!(fabs(v) < 1e300)) bad |= kStatNonFinite;
            e[k] += v;
        }
    }
    // wave-uniform: call from every lane of the wavefront
    __device__ __forceinline__ void take(const DevState &S, int next_blk, int next_inst, const LaneEnergy<R> &le) {
        to(S, next_blk, next_inst);
        energy(0, wave_sum((double)le.vdw));
        energy(1, wave_sum((double)le.elec));
        energy(2, wave_sum((double)le.pol));
        energy(3, wave_sum((double)le.vir));
        n[kCPairs] += wave_sum(le.pairs);
    }
    __device__ __forceinline__ void flush(const DevState &S) {
        if (blk >= 0 && lane_id() == 0) {
            const int field[4] = {kEVdw, kEElec, kEPolCross, kEVirial};
            unsigned long long *rec = S.blk_acc + (size_t)blk * kBAcc;
            for (int k = 0; k < 4; ++k) {
                if (e[k] == 0) continue;
                if constexpr (Prec<R>::fixed)
                    atomicAdd(rec + field[k], (unsigned long long)e[k]);
                else
                    unsafeAtomicAdd(reinterpret_cast<double *>(rec) + field[k], e[k]);
            }
            for (int k = 0; k < kCNum; ++k)
                if (n[k]) atomicAdd(rec + kENum + k, n[k]);
            if (bad) atomicOr(S.inst_status + inst, bad);
        }
        for (int k = 0; k < 4; ++k) e[k] = 0;
        for (int k = 0; k < kCNum; ++k) n[k] = 0;
        bad = 0;
    }
};

// ------------------------------------------------------------ Born radii A4
// Gpol_i += P4*V_j*CCF/r^4 over pairs with bond order > 2 and r <= Still
// cutoff (forcefields/gbff.cpp:432-465), both directions per unordered pair.
// Every accumulator is flushed at the end of each tile, so the sums an atom
// receives do not depend on how tiles are spread over wavefronts.
// Dense tiles are evaluated without lists or LDS atomics: lane = atom i, the
// j atoms rotate one lane per step (DPP wave_ror:1) and carry their own
// j-side sums with them. Diagonal tiles (a block with itself, ~80% of their
// 2,016 pairs inside the cutoff for peptides) are always dense: steps 1..32
// visit every unordered pair once (at step 32 both lanes of a pair meet it;
// the lower lane takes it) and after step 32 the j-side sum of atom a sits
// in lane (a + 32) & 63. Off-diagonal tiles are dense when their candidate
// count reached PairParams::dense_pairs at the last rebuild: 64 steps, after
// which the j-side sums are home again. The candidate / exclusion / 1-4 bits
// of each pair come from the tile's bit rows written at that rebuild.
struct DiagBits {
    unsigned long long mc, mx, m4;
};

__device__ __forceinline__ DiagBits diag_bits(const DevState &S, long long slot) {
    const unsigned long long *M = reinterpret_cast<const unsigned long long *>(S.cand + slot * kTileCap);
    const int l = lane_id();
    return {M[l], M[kBlk + l], M[2 * kBlk + l]};
}

__device__ __forceinline__ unsigned long long rotr64(unsigned long long v, int s) {
    return s ? ((v >> s) | (v << (64 - s))) : v;
}

// The tile's pair bits of this lane re-indexed by rotation step: bit k-1
// describes the pair met at step k, j lane (l + k*ldir) & 63, so a step
// tests one bit of a 32-bit half at a uniform position.
__device__ __forceinline__ unsigned long long step_bits(unsigned long long m, int l, int ldir) {
    return ldir == 1 ? rotr64(m, (l + 1) & 63) : rotr64(__builtin_bitreverse64(m), (64 - l) & 63);
}

// does lane l take the pair with lane jlane at rotation step k
__device__ __forceinline__ bool diag_take(const DiagBits &D, int jlane, int k, int l, bool diag) {
    return ((D.mc >> jlane) & 1ull) && (!diag || k < 32 || l < 32);
}
__device__ __forceinline__ int diag_bo(const DiagBits &D, int jlane) {
    return ((D.mx >> jlane) & 1ull) ? 2 : (((D.m4 >> jlane) & 1ull) ? 3 : 7);
}

// Still's pairwise Born term P4*CCF/r^4 (forcefields/gbff.cpp:448-460)
template <typename R>
__device__ __forceinline__ R born_term(R r2, R sR, R invP5, R P4, R P5) {
    // one reciprocal serves both quotients: t = 1/(sR^2 r^4)
    const R s2 = sR * sR, r4 = r2 * r2;
    const R t = Prec<R>::rcp(s2 * r4);
    const R x = (r2 * r4) * t; // r^2 / sR^2
    R ccf = (R)1;
    if (x <= invP5) {
        const R h = (R)0.5 - (R)0.5 * Prec<R>::cospi_(x * P5);
        ccf = h * h;
    }
    return P4 * ccf * (s2 * t);
}

template <typename R> struct BornAtom { R x, y, z, rad, vj; };

template <typename R>
__global__ void __launch_bounds__(256) k_born(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    // list tiles: I atoms at [0, 64), J atoms at [64, 128); dense tiles: the
    // J block twice in a row, so that step k reads lane (l + k*ldir) & 63 at
    // one address per lane
    __shared__ BornAtom<R> st[kWavesPerGroup][2 * kBlk];
    __shared__ double gi_acc[kWavesPerGroup][kBlk], gj_acc[kWavesPerGroup][kBlk];
    const int t_begin = S.row_off[blk_begin], t_end = min(S.row_off[blk_end], S.plist_cap);
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    const R invP5 = (R)P.invP5, P4 = (R)P.P4, P5 = (R)P.P5, still2 = (R)P.still2;
    InstSink<R> sink;
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
            const int tslot = S.tile_slot[t];
            if (diag || cnt.x < 0) {
                const DiagBits D = diag_bits(S, tslot);
                const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
                const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
                const AtomXYZ<R> pj = diag ? pi : load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
                R rad, vj, rj, vjj;
                if constexpr (Prec<R>::fixed) {
                    const float2 g = S.gb32[a], h = S.gb32[b];
                    rad = g.x; vj = g.y; rj = h.x; vjj = h.y;
                } else {
                    const double2 g = S.gb64[a], h = S.gb64[b];
                    rad = g.x; vj = g.y; rj = h.x; vjj = h.y;
                }
                st[w][l] = {pj.x, pj.y, pj.z, rj, vjj};
                st[w][l + kBlk] = {pj.x, pj.y, pj.z, rj, vjj};
                const int ldir = rot_dir();
                // candidates that are not excluded, by step
                unsigned long long T = step_bits(D.mc, l, ldir) & ~step_bits(D.mx, l, ldir);
                if (diag && l >= 32) T &= ~(1ull << 31); // step 32 meets each pair twice: the lower lane takes it
                wave_lds_sync();
                R gj = 0;
                double gi = 0;
                int np = 0;
                const int steps = diag ? 32 : kBlk;
                const int k0 = seg_step(it.s, steps, lg), k1 = seg_step(it.s + 1, steps, lg);
                const int dj = ldir == 1 ? 1 : -1;
                int jx = (ldir == 1 ? l : l + kBlk) + k0 * dj;
                for (int h = k0 >> 5; h < (k1 + 31) >> 5; ++h) {
                    const unsigned int Tw = (unsigned int)(T >> (32 * h));
                    const int kk1 = min(32, k1 - 32 * h);
                    for (int kk = max(0, k0 - 32 * h); kk < kk1; ++kk) {
                        jx += dj;
                        const BornAtom<R> B = st[w][jx];
                        gj = rot(gj);
                        const R dx = pi.x - B.x, dy = pi.y - B.y, dz = pi.z - B.z;
                        const R r2 = dx * dx + dy * dy + dz * dz;
                        if (((Tw >> kk) & 1u) && r2 <= still2) {
                            const R term = born_term<R>(r2, rad + B.rad, invP5, P4, P5);
                            gi += (double)(term * B.vj);
                            gj += term * vj;
                            ++np;
                        }
                    }
                }
                wave_lds_sync();
                int *st = S.inst_status + inst;
                gj = rot_home(gj, k1, ldir);
                if (diag) {
                    const double tot = gi + (double)gj;
                    if (l < S.blk_n[I]) acc_add<R>(S.acc_g, a, tot, kGpolScale, st);
                } else {
                    if (l < S.blk_n[I]) acc_add<R>(S.acc_g, a, gi, kGpolScale, st);
                    if (l < S.blk_n[J]) acc_add<R>(S.acc_g, b, (double)gj, kGpolScale, st);
                }
                const unsigned int npw = wave_sum((unsigned int)np);
                sink.to(S, I, inst);
                sink.n[kCBornPairs] += 2ull * npw;
                continue;
            }
            const int nb = cnt.x;
            if (nb == 0) continue;
            // list entries are fetched one chunk ahead of their use
            const uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
            int e_nx = l < nb ? (int)L[l] : 0;
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            if constexpr (Prec<R>::fixed) {
                const float2 gi = S.gb32[a], gj = S.gb32[b];
                st[w][l] = {pi.x, pi.y, pi.z, gi.x, gi.y};
                st[w][kBlk + l] = {pj.x, pj.y, pj.z, gj.x, gj.y};
            } else {
                const double2 gi = S.gb64[a], gj = S.gb64[b];
                st[w][l] = {pi.x, pi.y, pi.z, gi.x, gi.y};
                st[w][kBlk + l] = {pj.x, pj.y, pj.z, gj.x, gj.y};
            }
            gi_acc[w][l] = 0;
            gj_acc[w][l] = 0;
            wave_lds_sync();
            for (int q0 = 0; q0 < nb; q0 += kBlk) {
                const int q = q0 + l;
                const int e = e_nx;
                e_nx = q + kBlk < nb ? (int)L[q + kBlk] : 0;
                if (q < nb) {
                    const int ii = e & 63, jj = (e >> 6) & 63;
                    const BornAtom<R> A = st[w][ii], B = st[w][kBlk + jj];
                    const R dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    const R term = born_term<R>(r2, A.rad + B.rad, invP5, P4, P5);
                    atomicAdd(&gi_acc[w][ii], (double)(term * B.vj));
                    atomicAdd(&gj_acc[w][jj], (double)(term * A.vj));
                }
            }
            wave_lds_sync();
            int *st = S.inst_status + inst;
            if (l < S.blk_n[I]) acc_add<R>(S.acc_g, a, gi_acc[w][l], kGpolScale, st);
            if (l < S.blk_n[J]) acc_add<R>(S.acc_g, b, gj_acc[w][l], kGpolScale, st);
            sink.to(S, I, inst);
            sink.n[kCBornPairs] += (unsigned long long)(2 * nb);
            wave_lds_sync();
        }
    }
    sink.flush(S);
}

// Born radius a_i = -halfeconv/Gpol + offset (gbff.cpp:463); self term
// premul*q^2/a and its dE/da seed (gbff.cpp:1723-1728); clears force sums.
// Launched over whole blocks: each wavefront covers exactly one block.
template <typename R>
__global__ void k_born_final(DevState S, PairParams P, int atom_begin, int atom_end) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const bool in = a < atom_end;
    const int inst = in ? S.blk_inst[b] : -1;
    const bool v = inst >= 0 && (a % kBlk) < S.blk_n[b];
    double eself = 0;
    if (in) {
        acc_store<R>(S.acc_fx, a, 0.0, 1.0);
        acc_store<R>(S.acc_fy, a, 0.0, 1.0);
        acc_store<R>(S.acc_fz, a, 0.0, 1.0);
        acc_store<R>(S.acc_deda, a, 0.0, 1.0);
    }
    if (v) {
        const double G = S.t123[a] + acc_read<R>(S.acc_g, a, kGpolScale);
        const double ai = -P.halfeconv / G + P.born_off;
        reinterpret_cast<R *>(S.born)[a] = (R)ai;
        if constexpr (Prec<R>::fixed) {
            const float2 lj = S.lj32[a];
            S.pp32[a] = make_float4(lj.x, lj.y, (float)ai, (float)(0.60056120439322491 / ai));
        }
        const double qi = S.q[a];
        eself = P.premul * (qi * qi) / ai;
        if (!isfinite(eself)) atomicOr(S.inst_status + inst, kStatNonFinite);
        acc_store<R>(S.acc_deda, a, -eself / ai, kForceScale);
    }
    const double tot = wave_sum(eself);
    if (lane_id() == 0 && inst >= 0 && tot != 0.0)
        acc_add<R>(S.blk_acc, (size_t)b * kBAcc + kEPolSelf, tot, kEnergyScale, S.inst_status + inst);
}

// Switching function S and dS/dr of PD (forcefields/gbff.cpp:1605-1614).
template <typename R>
__device__ __forceinline__ void switch_fn(R r, R inner, R invSw, R &S, R &dS) {
    const R d = r - inner;
    const R s = (R)1 - (invSw * d) * (invSw * d);
    dS = (R)(-4) * (invSw * invSw) * d * s;
    S = s * s;
}

// The same without a branch: below the inner radius d clamps to 0, giving
// S = 1 and dS = -0, so S*x + y*dS == x exactly there.
template <typename R>
__device__ __forceinline__ void switch_fn_clamped(R r, R inner, R invSw, R &S, R &dS) {
    switch_fn<R>(fmax(r, inner), inner, invSw, S, dS);
}

template <typename R> struct PairAtom { R x, y, z, q, rad, eps, a; };

// One pair of the fused pass. Returns (dE/dr)/r (multiplies xi - xj) and
// adds dE/d(a) contributions to dai / daj (MODE 0) and energies to E.
template <typename R, int MODE>
__device__ __forceinline__ R pair_term(const PairParams &P, const PairAtom<R> &A, const PairAtom<R> &B, int bo,
                                       R r2, LaneEnergy<R> &E, R &dai, R &daj) {
    const R inner = (R)P.inner, invSw = (R)P.invSw;
    const R vcut = (R)P.vcut, vinner = (R)P.vinner, vinvSw = (R)P.vinvSw;
    const R premul = (R)P.premul;
    const R rinv = Prec<R>::rsqrt(r2);
    const R r = r2 * rinv;
    const R qiqj = A.q * B.q;
    R g = 0;
    E.pairs++;
    if (MODE == 0) {
        // Near pairs (u above the threshold) take the GB term and the
        // solute-dielectric Coulomb, far pairs the solvent-dielectric Coulomb
        // only; pairs of bond order <= 2 take no Coulomb and no LJ. Only the
        // GB term is a branch: the Coulomb and LJ parts run for every lane
        // with zero factors where they do not apply, so lanes of a wavefront
        // that disagree on near/far or on the switching regions do not run
        // the same code twice.
        const R aiaj = A.a * B.a;
        const R u = (R)(-0.25) * r2 * Prec<R>::rcp(aiaj);
        const bool b3 = bo >= 3, near = u > (R)P.exp_thresh;
        const R s14e = (bo == 3) ? (R)P.s14e : (R)1;
        R Sw, dS;
        switch_fn_clamped<R>(r, inner, invSw, Sw, dS);
        R econv;
        if constexpr (Prec<R>::fixed) econv = near ? P.econv_in32 : P.econv_out32;
        else econv = near ? P.econv_in : P.econv_out;
        R elec_pot = b3 ? econv * s14e * qiqj * rinv : (R)0;
        const R elec_force = Sw * (-elec_pot * rinv) + elec_pot * dS;
        elec_pot *= Sw;
        if (near) {
            const R tg = Prec<R>::expo(u);
            const R fgb = Prec<R>::rsqrt(r2 + tg * aiaj);
            const R L = fgb * fgb;
            R epol = premul * qiqj * fgb;
            R dedr = -epol * (r - (R)0.25 * r * tg) * L;
            R dedal = -epol * tg * ((R)1 - u) * L;
            dedr = Sw * dedr + dS * epol;
            dedal *= Sw;
            epol *= Sw;
            g = dedr * (R)2 * rinv;
            E.pol += (R)2 * epol;
            dai = dedal * B.a;
            daj = dedal * A.a;
        }
        // rr = 0 for pairs without LJ keeps bonded pairs at tiny separations
        // (a collapsing structure) from forming 0 * inf
        const bool lj = b3 && r < vcut;
        const R s14v = (bo == 3) ? (R)P.s14v : (R)1;
        const R rr = lj ? (A.rad + B.rad) * rinv : (R)0;
        const R rr2 = rr * rr;
        const R Bp = rr2 * rr2 * rr2;
        const R Ap = Bp * Bp;
        const R eps = lj ? s14v * A.eps * B.eps : (R)0;
        R vS, vd;
        if (P.same_switch) { // identical inputs give identical values
            vS = Sw;
            vd = dS;
        } else {
            switch_fn_clamped<R>(r, vinner, vinvSw, vS, vd);
        }
        R vdw_pot = ((R)2 * eps) * ((R)0.5 * Ap - Bp);
        const R vdw_force = vS * (-((R)12 * eps * rinv) * (Ap - Bp)) + vdw_pot * vd;
        vdw_pot *= vS;
        E.vdw += vdw_pot;
        E.elec += elec_pot;
        g += (vdw_force + elec_force) * rinv;
    } else {
        // FF_NonBonded energy path, bond-order scale tables {0,0,0,s14,1,1,1,1}
        const R sv = (bo <= 2) ? (R)0 : (bo == 3 ? (R)P.s14v : (R)1);
        const R se = (bo <= 2) ? (R)0 : (bo == 3 ? (R)P.s14e : (R)1);
        R vdw_pot = 0, vdw_force = 0, elec_pot = 0, elec_force = 0;
        if (P.vdw_mode > 0 && r < vcut) {
            const R rr = (A.rad + B.rad) * rinv;
            const R rr2 = rr * rr;
            const R Bp = rr2 * rr2 * rr2;
            const R Ap = Bp * Bp;
            const R eps = sv * A.eps * B.eps;
            vdw_pot = ((R)2 * eps) * ((R)0.5 * Ap - Bp);
            vdw_force = -((R)12 * eps * rinv) * (Ap - Bp);
            if (P.vdw_mode == 2 && r > vinner) {
                R vS, vd;
                switch_fn<R>(r, vinner, vinvSw, vS, vd);
                vdw_force = vS * vdw_force + vdw_pot * vd;
                vdw_pot *= vS;
            }
        }
        const R cut2 = (R)P.cut2;
        if (P.elec_mode == 1 || P.elec_mode == 3) {
            elec_pot = (R)P.econv_in * se * qiqj * rinv;
            elec_force = -elec_pot * rinv;
            if (P.elec_mode == 3 && r > inner) {
                R Sw, dS;
                switch_fn<R>(r, inner, invSw, Sw, dS);
                elec_force = Sw * elec_force + elec_pot * dS;
                elec_pot *= Sw;
            }
        } else if (P.elec_mode == 4) {
            elec_pot = (R)P.econv_in * se * qiqj;
            if (r > inner) {
                const R c2r = cut2 - r2;
                elec_force = -elec_pot * (R)P.fs_sA * (rinv * rinv) * (c2r * c2r) *
                             (cut2 - (R)3 * (R)P.inner2 + (R)2 * r2);
                elec_pot *= -((R)P.fs_sA * r * ((R)P.fs_sB * (rinv * rinv) + (R)P.fs_sC + (R)P.fs_sD * r2 +
                                                (R)P.fs_sE * (r2 * r2)) -
                              (R)P.fs_cutV);
            } else {
                elec_force = -elec_pot * (rinv * rinv);
                elec_pot *= (rinv - (R)P.fs_eshift);
            }
        }
        E.vdw += vdw_pot;
        E.elec += elec_pot;
        const R fm = vdw_force + elec_force;
        E.vir += r * fm;
        g = fm * rinv;
        // exact pairwise GB energy, no threshold branch (MODE 2: FF_NonBonded alone)
        if (MODE == 1 && r <= (R)P.cut) {
            R Sw = (R)1, dS;
            if (r >= inner) switch_fn<R>(r, inner, invSw, Sw, dS);
            const R s2 = r * r;
            const R aiaj = A.a * B.a;
            const R fgb = Prec<R>::rsqrt(s2 + aiaj * Prec<R>::expo(-s2 * (R)0.25 * Prec<R>::rcp(aiaj)));
            E.pol += premul * (R)2 * Sw * qiqj * fgb;
        }
    }
    return g;
}

// --------------------------------------------------- fused pair pass A5
// MODE 0: FastMode force pass (forcefields/gbff.cpp:1529-1729).
// MODE 1: energy path -- FF_NonBonded with the context's switching modes
//         (forcefields/nonbonded.cpp:670-873) plus the exact pairwise GB
//         energy of calcBornEnergy (forcefields/gbff.cpp:520-579).
// MODE 2: FF_NonBonded alone (forcefields/nonbonded.cpp:1205-1229).
template <typename R, int MODE>
__global__ void __launch_bounds__(256) k_pair(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ PairAtom<R> si[kWavesPerGroup][kBlk], sj[kWavesPerGroup][kBlk];
    __shared__ double fi[kWavesPerGroup][4][kBlk], fj[kWavesPerGroup][4][kBlk];
    const int t_begin = S.row_off[blk_begin], t_end = min(S.row_off[blk_end], S.plist_cap);
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    const R cut2 = (R)P.cut2;
    InstSink<R> sink;
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
            const int tslot = S.tile_slot[t];
            if (diag || cnt.x < 0) {
                const DiagBits D = diag_bits(S, tslot);
                const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
                const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
                const AtomXYZ<R> pj = diag ? pi : load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
                PairAtom<R> A, B;
                if constexpr (Prec<R>::fixed) {
                    const float2 u = S.lj32[a], v = S.lj32[b];
                    A = {pi.x, pi.y, pi.z, pi.q, u.x, u.y, ld_born<R>(S, a)};
                    B = {pj.x, pj.y, pj.z, pj.q, v.x, v.y, ld_born<R>(S, b)};
                } else {
                    const double2 u = S.lj64[a], v = S.lj64[b];
                    A = {pi.x, pi.y, pi.z, pi.q, u.x, u.y, ld_born<R>(S, a)};
                    B = {pj.x, pj.y, pj.z, pj.q, v.x, v.y, ld_born<R>(S, b)};
                }
                R fjx = 0, fjy = 0, fjz = 0, daj_acc = 0;
                double fix = 0, fiy = 0, fiz = 0, dai_acc = 0;
                LaneEnergy<R> E;
                const int ldir = rot_dir();
                const int steps = diag ? 32 : kBlk;
                const int k0 = seg_step(it.s, steps, lg), k1 = seg_step(it.s + 1, steps, lg);
                B.x = rot_to(B.x, k0, ldir); B.y = rot_to(B.y, k0, ldir); B.z = rot_to(B.z, k0, ldir);
                B.q = rot_to(B.q, k0, ldir); B.rad = rot_to(B.rad, k0, ldir); B.eps = rot_to(B.eps, k0, ldir);
                B.a = rot_to(B.a, k0, ldir);
                for (int k = k0 + 1; k <= k1; ++k) {
                    B.x = rot(B.x); B.y = rot(B.y); B.z = rot(B.z); B.q = rot(B.q);
                    B.rad = rot(B.rad); B.eps = rot(B.eps); B.a = rot(B.a);
                    fjx = rot(fjx); fjy = rot(fjy); fjz = rot(fjz);
                    if (MODE == 0) daj_acc = rot(daj_acc);
                    const int jlane = (l + k * ldir) & 63;
                    const R dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    if (diag_take(D, jlane, k, l, diag) && r2 <= cut2) {
                        R dai = 0, daj = 0;
                        const R g = pair_term<R, MODE>(P, A, B, diag_bo(D, jlane), r2, E, dai, daj);
                        const R gx = g * dx, gy = g * dy, gz = g * dz;
                        fix += (double)gx; fiy += (double)gy; fiz += (double)gz;
                        fjx -= gx; fjy -= gy; fjz -= gz;
                        if (MODE == 0) { dai_acc += (double)dai; daj_acc += daj; }
                    }
                }
                int *st = S.inst_status + inst;
                fjx = rot_home(fjx, k1, ldir); fjy = rot_home(fjy, k1, ldir); fjz = rot_home(fjz, k1, ldir);
                if (MODE == 0) daj_acc = rot_home(daj_acc, k1, ldir);
                if (diag) {
                    const double gx = fix + (double)fjx, gy = fiy + (double)fjy, gz = fiz + (double)fjz;
                    const double gd = MODE == 0 ? dai_acc + (double)daj_acc : 0.0;
                    if (l < S.blk_n[I]) {
                        acc_add<R>(S.acc_fx, a, gx, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, gy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, gz, kForceScale, st);
                        if (MODE == 0) acc_add<R>(S.acc_deda, a, gd, kForceScale, st);
                    }
                } else {
                    if (l < S.blk_n[I]) {
                        acc_add<R>(S.acc_fx, a, fix, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, fiy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, fiz, kForceScale, st);
                        if (MODE == 0) acc_add<R>(S.acc_deda, a, dai_acc, kForceScale, st);
                    }
                    if (l < S.blk_n[J]) {
                        acc_add<R>(S.acc_fx, b, (double)fjx, kForceScale, st);
                        acc_add<R>(S.acc_fy, b, (double)fjy, kForceScale, st);
                        acc_add<R>(S.acc_fz, b, (double)fjz, kForceScale, st);
                        if (MODE == 0) acc_add<R>(S.acc_deda, b, (double)daj_acc, kForceScale, st);
                    }
                }
                sink.take(S, I, inst, E);
                continue;
            }
            const int ntot = cnt.x + cnt.y;
            if (ntot == 0) continue;
            // list entries are fetched one chunk ahead of their use
            const int capm1 = (kTileCap >> lg) - 1;
            const uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
            int e_nx = l < ntot ? (int)L[entry_pos(l, cnt.x, capm1)] : 0;
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            R Ri, ei, Rj, ej;
            if constexpr (Prec<R>::fixed) {
                const float2 u = S.lj32[a], v = S.lj32[b];
                Ri = u.x; ei = u.y; Rj = v.x; ej = v.y;
            } else {
                const double2 u = S.lj64[a], v = S.lj64[b];
                Ri = u.x; ei = u.y; Rj = v.x; ej = v.y;
            }
            si[w][l] = {pi.x, pi.y, pi.z, pi.q, Ri, ei, ld_born<R>(S, a)};
            sj[w][l] = {pj.x, pj.y, pj.z, pj.q, Rj, ej, ld_born<R>(S, b)};
            for (int k = 0; k < 4; ++k) { fi[w][k][l] = 0; fj[w][k][l] = 0; }
            wave_lds_sync();
            LaneEnergy<R> E;
            for (int q0 = 0; q0 < ntot; q0 += kBlk) {
                const int q = q0 + l;
                const int e = e_nx;
                e_nx = q + kBlk < ntot ? (int)L[entry_pos(q + kBlk, cnt.x, capm1)] : 0;
                if (q < ntot) {
                    const int ii = e & 63, jj = (e >> 6) & 63, bo = (e >> 12) & 7;
                    const PairAtom<R> A = si[w][ii], B = sj[w][jj];
                    const R dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= cut2) { // born-side entries may reach past Cutoff when GbsaStillCutoff is larger
                        R dai = 0, daj = 0;
                        const R g = pair_term<R, MODE>(P, A, B, bo, r2, E, dai, daj);
                        const R gx = g * dx, gy = g * dy, gz = g * dz;
                        atomicAdd(&fi[w][0][ii], gx);
                        atomicAdd(&fi[w][1][ii], gy);
                        atomicAdd(&fi[w][2][ii], gz);
                        atomicAdd(&fj[w][0][jj], -gx);
                        atomicAdd(&fj[w][1][jj], -gy);
                        atomicAdd(&fj[w][2][jj], -gz);
                        if (MODE == 0) {
                            atomicAdd(&fi[w][3][ii], dai);
                            atomicAdd(&fj[w][3][jj], daj);
                        }
                    }
                }
            }
            wave_lds_sync();
            int *st = S.inst_status + inst;
            if (l < S.blk_n[I]) {
                acc_add<R>(S.acc_fx, a, fi[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, a, fi[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, a, fi[w][2][l], kForceScale, st);
                if (MODE == 0) acc_add<R>(S.acc_deda, a, fi[w][3][l], kForceScale, st);
            }
            if (l < S.blk_n[J]) {
                acc_add<R>(S.acc_fx, b, fj[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, b, fj[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, b, fj[w][2][l], kForceScale, st);
                if (MODE == 0) acc_add<R>(S.acc_deda, b, fj[w][3][l], kForceScale, st);
            }
            sink.take(S, I, inst, E);
            wave_lds_sync();
        }
    }
    sink.flush(S);
}

// ---------------------------------------------------------- chain rule A6
// dE/dx through the Born radii (forcefields/gbff.cpp:1751-1808): for each
// unordered pair, both ordered terms share CCF; w_i = a_i^2/166 * dE/da_i
// (the reference's literal 166.0, gbff.cpp:1755).
template <typename R>
__device__ __forceinline__ R chain_weight(const DevState &S, size_t a) {
    const double ai = (double)ld_born<R>(S, a);
    return (R)(ai * ai / 166.0 * acc_read<R>(S.acc_deda, a, kForceScale));
}

template <typename R> struct ChainAtom { R x, y, z, rad, vj, wt; };

// d/dr of the Still term over r, without the atom weights:
// P4 (4 CCF - dCCF) / r^6 (forcefields/gbff.cpp:1770-1795)
template <typename R>
__device__ __forceinline__ R chain_coef(R r2, R sR, R invP5, R P4, R P5) {
    // one reciprocal serves both quotients: t = 1/(sR^2 r^6)
    const R s2 = sR * sR, r6 = r2 * r2 * r2;
    const R t = Prec<R>::rcp(s2 * r6);
    const R x = (r2 * r6) * t; // r^2 / sR^2
    R ccf = (R)1, dccf = (R)0;
    if (x <= invP5) {
        const R f = x * P5;
        const R h = (R)0.5 - (R)0.5 * Prec<R>::cospi_(f);
        ccf = h * h;
        dccf = (R)2 * h * Prec<R>::sinpi_(f) * (f * (R)3.14159265358979323846);
    }
    return P4 * ((R)4 * ccf - dccf) * (s2 * t);
}

template <typename R> struct ChainLds {
    ChainAtom<R> si[kWavesPerGroup][kBlk], sj[kWavesPerGroup][kBlk];
    double fi[kWavesPerGroup][3][kBlk], fj[kWavesPerGroup][3][kBlk];
};

// The chain-rule pass over the tiles, by the first ngroups workgroups of the
// grid (k_chain; with GB forces the bonded terms run as further workgroups
// of the same launch, k_chain_bonded).
template <typename R>
__device__ __forceinline__ void chain_work(const DevState &S, const PairParams &P, int blk_begin, int blk_end, int tpw,
                                           int ngroups, ChainLds<R> &Ls) {
    auto &si = Ls.si;
    auto &sj = Ls.sj;
    auto &fi = Ls.fi;
    auto &fj = Ls.fj;
    const int t_begin = S.row_off[blk_begin], t_end = min(S.row_off[blk_end], S.plist_cap);
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = ngroups * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    const R invP5 = (R)P.invP5, P4 = (R)P.P4, P5 = (R)P.P5, still2 = (R)P.still2;
    for (int c = wave_global_id(ngroups) * tpw; c < u_end; c += nw * tpw) {
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
            const int tslot = S.tile_slot[t];
            if (diag || cnt.x < 0) {
                const DiagBits D = diag_bits(S, tslot);
                const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
                const bool vi = l < S.blk_n[I], vjb = l < S.blk_n[J];
                const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
                const AtomXYZ<R> pj = diag ? pi : load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
                R rad, vjv, rj, vjj;
                if constexpr (Prec<R>::fixed) {
                    const float2 g = S.gb32[a], h = S.gb32[b];
                    rad = g.x; vjv = g.y; rj = h.x; vjj = h.y;
                } else {
                    const double2 g = S.gb64[a], h = S.gb64[b];
                    rad = g.x; vjv = g.y; rj = h.x; vjj = h.y;
                }
                const ChainAtom<R> A = {pi.x, pi.y, pi.z, rad, vjv, vi ? chain_weight<R>(S, a) : (R)0};
                ChainAtom<R> B = {pj.x, pj.y, pj.z, rj, vjj, vjb ? chain_weight<R>(S, b) : (R)0};
                // i-side sums in the pass's precision: at most 64 terms per
                // lane before the flush
                R fjx = 0, fjy = 0, fjz = 0;
                R fix = 0, fiy = 0, fiz = 0;
                const int ldir = rot_dir();
                // candidates that are not excluded, by step (bit k-1: step k)
                unsigned long long T = step_bits(D.mc, l, ldir) & ~step_bits(D.mx, l, ldir);
                if (diag && l >= 32) T &= ~(1ull << 31); // step 32 meets each pair twice: the lower lane takes it
                const int steps = diag ? 32 : kBlk;
                const int k0 = seg_step(it.s, steps, lg), k1 = seg_step(it.s + 1, steps, lg);
                B.x = rot_to(B.x, k0, ldir); B.y = rot_to(B.y, k0, ldir); B.z = rot_to(B.z, k0, ldir);
                B.rad = rot_to(B.rad, k0, ldir); B.vj = rot_to(B.vj, k0, ldir); B.wt = rot_to(B.wt, k0, ldir);
                for (int h = k0 >> 5; h < (k1 + 31) >> 5; ++h) {
                    const unsigned int Tw = (unsigned int)(T >> (32 * h));
                    const int kk1 = min(32, k1 - 32 * h);
                    for (int kk = max(0, k0 - 32 * h); kk < kk1; ++kk) {
                        B.x = rot(B.x); B.y = rot(B.y); B.z = rot(B.z);
                        B.rad = rot(B.rad); B.vj = rot(B.vj); B.wt = rot(B.wt);
                        fjx = rot(fjx); fjy = rot(fjy); fjz = rot(fjz);
                        const R dx = B.x - A.x, dy = B.y - A.y, dz = B.z - A.z;
                        const R r2 = dx * dx + dy * dy + dz * dz;
                        if (((Tw >> kk) & 1u) && r2 <= still2) {
                            const R coef = chain_coef<R>(r2, A.rad + B.rad, invP5, P4, P5) *
                                           (A.wt * B.vj + B.wt * A.vj);
                            const R gx = coef * dx, gy = coef * dy, gz = coef * dz;
                            fix += gx; fiy += gy; fiz += gz;
                            fjx -= gx; fjy -= gy; fjz -= gz;
                        }
                    }
                }
                int *st = S.inst_status + inst;
                fjx = rot_home(fjx, k1, ldir); fjy = rot_home(fjy, k1, ldir); fjz = rot_home(fjz, k1, ldir);
                if (diag) {
                    const double gx = (double)fix + (double)fjx, gy = (double)fiy + (double)fjy,
                                 gz = (double)fiz + (double)fjz;
                    if (vi) {
                        acc_add<R>(S.acc_fx, a, gx, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, gy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, gz, kForceScale, st);
                    }
                } else {
                    if (vi) {
                        acc_add<R>(S.acc_fx, a, (double)fix, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, (double)fiy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, (double)fiz, kForceScale, st);
                    }
                    if (vjb) {
                        acc_add<R>(S.acc_fx, b, (double)fjx, kForceScale, st);
                        acc_add<R>(S.acc_fy, b, (double)fjy, kForceScale, st);
                        acc_add<R>(S.acc_fz, b, (double)fjz, kForceScale, st);
                    }
                }
                continue;
            }
            const int nb = cnt.x;
            if (nb == 0) continue;
            // list entries are fetched one chunk ahead of their use
            const uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
            int e_nx = l < nb ? (int)L[l] : 0;
            const bool vi = l < S.blk_n[I], vj = l < S.blk_n[J];
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            R ri, vji, rj, vjj;
            if constexpr (Prec<R>::fixed) {
                const float2 gi = S.gb32[a], gj = S.gb32[b];
                ri = gi.x; vji = gi.y; rj = gj.x; vjj = gj.y;
            } else {
                const double2 gi = S.gb64[a], gj = S.gb64[b];
                ri = gi.x; vji = gi.y; rj = gj.x; vjj = gj.y;
            }
            si[w][l] = {pi.x, pi.y, pi.z, ri, vji, vi ? chain_weight<R>(S, a) : (R)0};
            sj[w][l] = {pj.x, pj.y, pj.z, rj, vjj, vj ? chain_weight<R>(S, b) : (R)0};
            for (int k = 0; k < 3; ++k) { fi[w][k][l] = 0; fj[w][k][l] = 0; }
            wave_lds_sync();
            for (int q0 = 0; q0 < nb; q0 += kBlk) {
                const int q = q0 + l;
                const int e = e_nx;
                e_nx = q + kBlk < nb ? (int)L[q + kBlk] : 0;
                if (q < nb) {
                    const int ii = e & 63, jj = (e >> 6) & 63;
                    const ChainAtom<R> A = si[w][ii], B = sj[w][jj];
                    const R dx = B.x - A.x, dy = B.y - A.y, dz = B.z - A.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    const R coef = chain_coef<R>(r2, A.rad + B.rad, invP5, P4, P5) * (A.wt * B.vj + B.wt * A.vj);
                    const R gx = coef * dx, gy = coef * dy, gz = coef * dz;
                    atomicAdd(&fi[w][0][ii], gx);
                    atomicAdd(&fi[w][1][ii], gy);
                    atomicAdd(&fi[w][2][ii], gz);
                    atomicAdd(&fj[w][0][jj], -gx);
                    atomicAdd(&fj[w][1][jj], -gy);
                    atomicAdd(&fj[w][2][jj], -gz);
                }
            }
            wave_lds_sync();
            int *st = S.inst_status + inst;
            if (vi) {
                acc_add<R>(S.acc_fx, a, fi[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, a, fi[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, a, fi[w][2][l], kForceScale, st);
            }
            if (vj) {
                acc_add<R>(S.acc_fx, b, fj[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, b, fj[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, b, fj[w][2][l], kForceScale, st);
            }
            wave_lds_sync();
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_chain(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ ChainLds<R> Ls;
    chain_work<R>(S, P, blk_begin, blk_end, tpw, gridDim.x, Ls);
}

// ------------------------------------------------ FF_SoftVDW soft core
// FF_SoftVDW::calcForces_Soft (forcefields/ffsoftvdw.cpp:71-145): each pair
// of bond order above 1-4 with r < s = R_i + R_j - OlapOffset and
// r < VdwCutoff takes E = H (1 - r^2/s^2)^2, so (dE/dr)/r =
// 4 H / s^2 (r^2/s^2 - 1). Returns (dE/dr)/r and the energy in e (both 0
// outside the range).
template <typename R>
__device__ __forceinline__ R soft_term(const PairParams &P, R rsum, R r2, R &e) {
    const R s = rsum - (R)P.soft_olap;
    const R s2 = s * s;
    if (!(r2 < s2 && r2 < (R)P.soft_cut2)) {
        e = 0;
        return 0;
    }
    const R is2 = Prec<R>::rcp(s2);
    const R x = r2 * is2;
    const R h = (R)P.soft_h;
    e = h * ((R)1 - x) * ((R)1 - x);
    return (R)4 * h * is2 * (x - (R)1);
}

template <typename R> struct SoftAtom { R x, y, z, rad; };

template <typename R> __device__ __forceinline__ R ld_radius(const DevState &S, size_t a) {
    if constexpr (Prec<R>::fixed) return S.lj32[a].x;
    else return S.lj64[a].x;
}

// Same tile walk as the pair pass: filtered lists with LDS sums, dense
// tiles by lane rotation from their bit rows (1-2, 1-3 and 1-4 pairs are
// excluded through the excluded and 1-4 rows).
template <typename R>
__global__ void __launch_bounds__(256) k_soft(DevState S, PairParams P, int blk_begin, int blk_end, int tpw) {
    __shared__ SoftAtom<R> si[kWavesPerGroup][kBlk], sj[kWavesPerGroup][kBlk];
    __shared__ double fi[kWavesPerGroup][3][kBlk], fj[kWavesPerGroup][3][kBlk];
    const int t_begin = S.row_off[blk_begin], t_end = min(S.row_off[blk_end], S.plist_cap);
    const int lg = P.lg_split;
    const int nt = t_end - t_begin, u_end = max(nt, 0) << lg;
    const int nw = gridDim.x * kWavesPerGroup;
    if (tpw <= 0) tpw = max(1, (u_end + nw - 1) / nw);
    const int l = lane_id(), w = wave_in_group();
    InstSink<R> sink;
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
            const int tslot = S.tile_slot[t];
            const size_t a = (size_t)I * kBlk + l, b = (size_t)J * kBlk + l;
            int *st = S.inst_status + inst;
            LaneEnergy<R> E;
            if (diag || cnt.x < 0) {
                const DiagBits D = diag_bits(S, tslot);
                const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
                const AtomXYZ<R> pj = diag ? pi : load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
                const R ri = ld_radius<R>(S, a);
                R xj = pj.x, yj = pj.y, zj = pj.z, rj = diag ? ri : ld_radius<R>(S, b);
                R fjx = 0, fjy = 0, fjz = 0;
                double fix = 0, fiy = 0, fiz = 0;
                const unsigned long long skip = D.mx | D.m4;
                const int ldir = rot_dir();
                const int steps = diag ? 32 : kBlk;
                const int k0 = seg_step(it.s, steps, lg), k1 = seg_step(it.s + 1, steps, lg);
                xj = rot_to(xj, k0, ldir); yj = rot_to(yj, k0, ldir); zj = rot_to(zj, k0, ldir);
                rj = rot_to(rj, k0, ldir);
                for (int k = k0 + 1; k <= k1; ++k) {
                    xj = rot(xj); yj = rot(yj); zj = rot(zj); rj = rot(rj);
                    fjx = rot(fjx); fjy = rot(fjy); fjz = rot(fjz);
                    const int jlane = (l + k * ldir) & 63;
                    const R dx = pi.x - xj, dy = pi.y - yj, dz = pi.z - zj;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    if (diag_take(D, jlane, k, l, diag) && !((skip >> jlane) & 1ull)) {
                        R e;
                        const R g = soft_term<R>(P, ri + rj, r2, e);
                        const R gx = g * dx, gy = g * dy, gz = g * dz;
                        fix += (double)gx; fiy += (double)gy; fiz += (double)gz;
                        fjx -= gx; fjy -= gy; fjz -= gz;
                        E.vdw += e;
                    }
                }
                fjx = rot_home(fjx, k1, ldir); fjy = rot_home(fjy, k1, ldir); fjz = rot_home(fjz, k1, ldir);
                if (diag) {
                    const double gx = fix + (double)fjx, gy = fiy + (double)fjy, gz = fiz + (double)fjz;
                    if (l < S.blk_n[I]) {
                        acc_add<R>(S.acc_fx, a, gx, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, gy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, gz, kForceScale, st);
                    }
                } else {
                    if (l < S.blk_n[I]) {
                        acc_add<R>(S.acc_fx, a, fix, kForceScale, st);
                        acc_add<R>(S.acc_fy, a, fiy, kForceScale, st);
                        acc_add<R>(S.acc_fz, a, fiz, kForceScale, st);
                    }
                    if (l < S.blk_n[J]) {
                        acc_add<R>(S.acc_fx, b, (double)fjx, kForceScale, st);
                        acc_add<R>(S.acc_fy, b, (double)fjy, kForceScale, st);
                        acc_add<R>(S.acc_fz, b, (double)fjz, kForceScale, st);
                    }
                }
                sink.take(S, I, inst, E);
                continue;
            }
            const int ntot = cnt.x + cnt.y;
            if (ntot == 0) continue;
            const int capm1 = (kTileCap >> lg) - 1;
            const uint16_t *L = S.plist + (size_t)t * kTileCap + (size_t)it.s * (kTileCap >> lg);
            int e_nx = l < ntot ? (int)L[entry_pos(l, cnt.x, capm1)] : 0;
            const AtomXYZ<R> pi = load_xyzq<R>(S, a, make_double4(0, 0, 0, 0));
            const AtomXYZ<R> pj = load_xyzq<R>(S, b, tile_shift<R>(S, I, J));
            si[w][l] = {pi.x, pi.y, pi.z, ld_radius<R>(S, a)};
            sj[w][l] = {pj.x, pj.y, pj.z, ld_radius<R>(S, b)};
            for (int k = 0; k < 3; ++k) { fi[w][k][l] = 0; fj[w][k][l] = 0; }
            wave_lds_sync();
            for (int q0 = 0; q0 < ntot; q0 += kBlk) {
                const int q = q0 + l;
                const int e = e_nx;
                e_nx = q + kBlk < ntot ? (int)L[entry_pos(q + kBlk, cnt.x, capm1)] : 0;
                if (q < ntot && ((e >> 12) & 7) > 3) {
                    const int ii = e & 63, jj = (e >> 6) & 63;
                    const SoftAtom<R> A = si[w][ii], B = sj[w][jj];
                    const R dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    R en;
                    const R g = soft_term<R>(P, A.rad + B.rad, r2, en);
                    if (g != (R)0 || en != (R)0) {
                        const R gx = g * dx, gy = g * dy, gz = g * dz;
                        atomicAdd(&fi[w][0][ii], (double)gx);
                        atomicAdd(&fi[w][1][ii], (double)gy);
                        atomicAdd(&fi[w][2][ii], (double)gz);
                        atomicAdd(&fj[w][0][jj], (double)-gx);
                        atomicAdd(&fj[w][1][jj], (double)-gy);
                        atomicAdd(&fj[w][2][jj], (double)-gz);
                        E.vdw += en;
                    }
                }
            }
            wave_lds_sync();
            if (l < S.blk_n[I]) {
                acc_add<R>(S.acc_fx, a, fi[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, a, fi[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, a, fi[w][2][l], kForceScale, st);
            }
            if (l < S.blk_n[J]) {
                acc_add<R>(S.acc_fx, b, fj[w][0][l], kForceScale, st);
                acc_add<R>(S.acc_fy, b, fj[w][1][l], kForceScale, st);
                acc_add<R>(S.acc_fz, b, fj[w][2][l], kForceScale, st);
            }
            sink.take(S, I, inst, E);
            wave_lds_sync();
        }
    }
    sink.flush(S);
}

// Forces f = -dE/dx in kcal/mol/A (double), dE/da for export.
template <typename R>
// With f2 given, also adds each instance's sum of |f|^2 to f2[inst *
// f2_stride] (one wavefront is one block of one instance, atom_begin being
// block-aligned).
__global__ void k_finalize(DevState S, int atom_begin, int atom_end, double *f2, int f2_stride) {
    const int a = atom_begin + blockIdx.x * blockDim.x + threadIdx.x;
    const int b = a / kBlk;
    const int inst = a < atom_end ? S.blk_inst[b] : -1;
    blk_fold<R>(S, b, inst);
    double sq = 0;
    if (inst >= 0 && (a % kBlk) < S.blk_n[b]) {
        const double fx = -acc_read<R>(S.acc_fx, a, kForceScale);
        const double fy = -acc_read<R>(S.acc_fy, a, kForceScale);
        const double fz = -acc_read<R>(S.acc_fz, a, kForceScale);
        S.fx[a] = fx; S.fy[a] = fy; S.fz[a] = fz;
        if (!(isfinite(fx) && isfinite(fy) && isfinite(fz))) atomicOr(S.inst_status + inst, kStatNonFinite);
        sq = fx * fx + fy * fy + fz * fz;
    }
    if (f2) { // uniform over the launch
        sq = wave_sum(sq);
        if (lane_id() == 0 && inst >= 0) unsafeAtomicAdd(f2 + (size_t)inst * f2_stride, sq);
    }
}

// Copies the records of slots [s0, s0 + ns) into host-mapped pinned memory
// laid out as [ns][kENum] energies, [ns][kCNum] counters, [ns] status words,
// as the last launch of the stream before the host waits. A copy-engine
// transfer queued behind the closing kernel started ~0.1 ms after that
// kernel ended; this launch follows it within a few microseconds.
__global__ void k_publish(DevState S, int s0, int ns, unsigned long long *out) {
    const size_t ne = (size_t)ns * kENum, nc = (size_t)ns * kCNum;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long *e = (const unsigned long long *)S.inst_energy + (size_t)s0 * kENum;
    const unsigned long long *c = S.inst_count + (size_t)s0 * kCNum;
    for (size_t k = t; k < ne; k += stride) out[k] = e[k];
    for (size_t k = t; k < nc; k += stride) out[ne + k] = c[k];
    int *st = (int *)(out + ne + nc);
    for (size_t k = t; k < (size_t)ns; k += stride) st[k] = S.inst_status[s0 + k];
}

// ------------------------------------------------------ PD-format export
// Packed list of one instance exactly as calcNewList_InfiniteSpace writes
// it (workspace/neighbourlist.cpp:263-315): j ascending, r^2 in FP64 with the
// reference's operation order (no contraction), r^2 <= lim included.
template <int PASS>
__global__ void k_export(DevState S, PairParams P, int blk0, int nblk, int shadow,
                         int *counts, const long long *offs, uint32_t *entries) {
#pragma clang fp contract(off)
    const int I = blk0 + blockIdx.x * (blockDim.x / 64) + wave_in_group();
    if (I >= blk0 + nblk) return;
    const int l = lane_id();
    const int inst = S.blk_inst[I];
    const int il = (I - blk0) * kBlk + l;
    const bool vi = l < S.blk_n[I];
    const size_t a = (size_t)I * kBlk + l;
    const double xi = vi ? S.px[a] : 0, yi = vi ? S.py[a] : 0, zi = vi ? S.pz[a] : 0;
    long long pos = (PASS == 1 && vi) ? offs[il] : 0;
    int cnt = 0;
    const int Jend = shadow ? blk0 + nblk - 1 : I;
    for (int J = blk0; J <= Jend; ++J) {
        if (!boxes_within(S, I, J, P.list_r)) continue;
        const int nj = S.blk_n[J];
        for (int k = 0; k < nj; ++k) {
            const int jl = (J - blk0) * kBlk + k;
            if (!shadow && jl >= il) break;
            const size_t bj = (size_t)J * kBlk + k;
            const double dx = S.px[bj] - xi;
            const double dy = S.py[bj] - yi;
            const double dz = S.pz[bj] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (!vi || d2 > P.list_lim2 || jl == il) continue;
            if (PASS == 0) {
                ++cnt;
            } else {
                const int lo = min(il, jl), hi = max(il, jl);
                const int bo = (hi - lo < S.inst_mid[inst])
                                   ? pair_bondorder(S, inst, hi, lo, 0)
                                   : pair_bondorder(S, inst, hi, lo, kTileSparse);
                entries[pos++] = ((uint32_t)jl & 0x00FFFFFFu) | ((uint32_t)(bo & 7) << 24);
            }
        }
    }
    if (PASS == 0 && vi) counts[il] = cnt;
}

} // namespace pdgpu
