// pdgpu.hip -- context, device memory layout, launch sequences and the C ABI
// declared in include/pdgpu.h.
#include "../../include/pdgpu.h"
#include "pdgpu_device.h"
#include "pdgpu_pair32.h"
#include "pdgpu_bonded.h"
#include "pdgpu_integrate.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <map>
#include <vector>

using namespace pdgpu;

// ------------------------------------------- neighbour-list export, no allocation
// pdgpu_nlist_export_noalloc stages the per-atom counts, row offsets and the
// packed entries in one static device array instead of allocating them per
// call. Each GPU holds its own copy until the process ends; host code uses
// the per-device address from hipGetSymbolAddress, never the symbol itself.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race (one context or one thread per GPU at a time).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
unsigned char *g_scratch[kMaxDevices];     // device address of scratch per ordinal
int g_scratch_rc[kMaxDevices];             // 0 unset, 1 ready, -1 lookup failed

// Called once per device with that device current (context creation).
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return PDGPU_ERR_NODEVICE;
    if (g_scratch_rc[device] != 0) return g_scratch_rc[device] > 0 ? PDGPU_OK : PDGPU_ERR_HIP;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch[device] = (unsigned char *)ptr;
    g_scratch_rc[device] = (e == hipSuccess && ptr) ? 1 : -1;
    return g_scratch_rc[device] > 0 ? PDGPU_OK : PDGPU_ERR_HIP;
}
} // namespace


namespace {

constexpr double kKcal2J = 4184.0;
constexpr double kNa = 6.0221418E23;
constexpr double kAngstrom = 1E-10;
constexpr double kPI = 3.14159265358979323846;
constexpr double kBoltzmann = 1.38065E-23;        // PhysicsConst::kB, J/K (forcefields/forcefield.cpp:18-38)
const double kJPerKcal = kKcal2J / kNa;          // J/molecule per kcal/mol
const double kForceToN = kJPerKcal / kAngstrom;  // N per kcal/mol/A

// PD's electrostatic constant, kcal/mol*A (forcefields/forcefield.cpp:32-36)
double econv_kcal() {
    const double e = 1.60217645E-19, e0 = 8.854187817E-12;
    const double four_pi_e0 = (kPI * 4.0) * e0;
    return ((1.0 / kKcal2J) * kNa * (e * e)) / (four_pi_e0 * kAngstrom);
}

// First-fit allocator over [0, cap) with coalescing of freed ranges.
struct Ranges {
    std::vector<std::pair<int64_t, int64_t>> free_; // (begin, length), sorted
    int64_t cap = 0;
    void init(int64_t c) { cap = c; free_.clear(); if (c > 0) free_.push_back({0, c}); }
    int64_t alloc(int64_t n) {
        if (n == 0) return 0;
        for (size_t k = 0; k < free_.size(); ++k) {
            if (free_[k].second >= n) {
                const int64_t b = free_[k].first;
                free_[k].first += n;
                free_[k].second -= n;
                if (free_[k].second == 0) free_.erase(free_.begin() + k);
                return b;
            }
        }
        return -1;
    }
    void release(int64_t b, int64_t n) {
        if (n == 0) return;
        auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(b, (int64_t)0));
        it = free_.insert(it, {b, n});
        size_t k = it - free_.begin();
        if (k + 1 < free_.size() && free_[k].first + free_[k].second == free_[k + 1].first) {
            free_[k].second += free_[k + 1].second;
            free_.erase(free_.begin() + k + 1);
        }
        if (k > 0 && free_[k - 1].first + free_[k - 1].second == free_[k].first) {
            free_[k - 1].second += free_[k].second;
            free_.erase(free_.begin() + k);
        }
    }
    // extend capacity to c (new space appended to the free list)
    void grow(int64_t c) {
        if (c <= cap) return;
        release(cap, c - cap);
        cap = c;
    }
    int64_t tail_free() const {
        if (free_.empty()) return 0;
        const auto &b = free_.back();
        return (b.first + b.second == cap) ? b.second : 0;
    }
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Instance {
    bool active = false;
    int natoms = 0, nblocks = 0, block0 = 0, mid = 0, nsparse = 0;
    int64_t band_off = 0, band_len = 0, sp_off = 0;
    int64_t bd_off = 0, bd_len = 0, an_off = 0, an_len = 0, dh_off = 0, dh_len = 0; // FF_Bonded records
    int64_t cs_off = 0, cs_len = 0; // candidate tile slots
    bool listed = false;  // tile rows built
    bool has_vel = false;
    bool sd = false;      // a minimisation was started and not finished
    std::vector<double> mass;
};

// One destination of a staged upload: bytes from blob offset src (or
// zeros) to the device address dst.
struct UpSec {
    unsigned long long dst;
    long long src, bytes, zero;
};

// Scatters the sections of an uploaded blob: grid (sections, parts), each
// part copying a strided share of its section in the widest words that the
// section's alignment allows.
__global__ void k_scatter(const unsigned char *blob, const UpSec *secs) {
    const UpSec u = secs[blockIdx.x];
    const long long tid = (long long)blockIdx.y * blockDim.x + threadIdx.x, nth = (long long)gridDim.y * blockDim.x;
    const unsigned long long al = u.dst | (unsigned long long)u.src | (unsigned long long)u.bytes;
    if ((al & 7) == 0) {
        unsigned long long *d = reinterpret_cast<unsigned long long *>(u.dst);
        const unsigned long long *q = reinterpret_cast<const unsigned long long *>(blob + u.src);
        for (long long k = tid; k < u.bytes / 8; k += nth) d[k] = u.zero ? 0ull : q[k];
    } else if ((al & 3) == 0) {
        unsigned int *d = reinterpret_cast<unsigned int *>(u.dst);
        const unsigned int *q = reinterpret_cast<const unsigned int *>(blob + u.src);
        for (long long k = tid; k < u.bytes / 4; k += nth) d[k] = u.zero ? 0u : q[k];
    } else {
        unsigned char *d = reinterpret_cast<unsigned char *>(u.dst);
        for (long long k = tid; k < u.bytes; k += nth) d[k] = u.zero ? (unsigned char)0 : blob[u.src + k];
    }
}

} // namespace

struct pdgpu_ctx {
    pdgpu_options opt;
    PairParams P;
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cus = 0;
    bool mixed = false;
    int64_t budget = 0, used = 0;

    // capacities
    int64_t cap_blocks = 0, cap_band = 0, cap_sparse = 0, cap_tiles = 0;
    int max_inst = 0;
    Ranges blocks, bandr, spr, bdr, anr, dhr, csr;
    int64_t cap_bd = 0, cap_an = 0, cap_dh = 0, cap_cs = 0;
    std::vector<Instance> inst;
    int64_t tri_sum = 0; // sum over instances of nb*(nb+1)/2 (tile bound)
    int64_t ntiles = 0;  // tiles after the last build

    // device buffers
    // per atom
    DevBuf px, py, pz, q, lj64, gb64, t123, xyzq32, lj32, gb32, pp32, born, acc_g, acc_deda, acc_fx, acc_fy, acc_fz;
    DevBuf fx, fy, fz, vx, vy, vz, minv, opx, opy, opz, rx, ry, rz;
    // per block
    DevBuf blk_inst, blk_n, blk_first, blk_ctr, blk_half, blk_box, blk_acc, row_count, row_off, tiles, tile_slot;
    // per instance
    DevBuf any_need, inst_need, inst_moved, inst_nat, inst_mid, inst_band_off, inst_sp_off, inst_sp_n, inst_md;
    // per-instance records, one allocation: energies [max_inst][kENum], then
    // counters [max_inst][kCNum], then status [max_inst], so that the records
    // of every slot come back in one copy
    DevBuf inst_rec;
    DevBuf fox, foy, foz; // Beeman's old forces (N), per atom
    DevBuf is_h;          // Particle::isHydrogen per atom (Langevin without hydrogens)
    double thermo_tau = 0, thermo_temp = 300; // Berendsen thermostat of the PD MD calls (tau 0: none)
    double thermo_dt = 0;                     // time step of the running MD call
    DevBuf rng_ctr;       // Langevin device-noise step counter
    DevBuf stage;         // packing buffer of the output formats
    // DCD trajectories attached to instances (pdgpu_traj_dcd): frames are
    // packed on the device into a ring and written in batches
    struct TrajSink {
        std::string path;
        int every = 0;
        bool created = false;
        int frames = 0, cap = 0;
        DevBuf ring;
    };
    std::map<int, TrajSink> traj;
    DevBuf band, sparse;
    DevBuf plist, tile_cnt, overflow;
    DevBuf cand, cand_n, inst_cslot; // candidate pairs of the last rebuild, per tile slot
    DevBuf bd_idx, bd_par, an_idx, an_par, dh_idx, dh_meta, dh_vn, dh_n, dh_gamma; // FF_Bonded records
    int64_t plist_cap = 0; // tiles the pair-list buffer holds
    int grid_wpc[4] = {16, 16, 16, 16}; // wavefronts per CU of the tile kernels (enum, born, pair, chain)

    // host mirrors of per-block metadata
    std::vector<int> h_blk_inst, h_blk_n, h_blk_first;
    // pinned host staging for per-instance results (energies, counts, status)
    void *h_stage = nullptr, *d_stage = nullptr; // pinned, host and device addresses
    size_t h_stage_bytes = 0;
    // staged uploads: host data of an instance (topology, term records,
    // positions) is gathered into one blob, copied once through pinned
    // memory and scattered on the device by k_scatter
    std::vector<unsigned char> up_blob;
    std::vector<UpSec> up_secs;
    void *h_up = nullptr;
    size_t h_up_bytes = 0;
    hipEvent_t up_done = nullptr;
    DevBuf upl;

    // captured step graphs, keyed by loop kind and instance, each with the
    // launch arguments it was captured with
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<char> sig;
    };
    std::map<std::string, StepGraph> graphs;

    // optional per-kernel timing with HIP events on the context stream
    bool prof = false;
    struct ProfRec { int kind; hipEvent_t a, b; };
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;

    DevState S() const {
        DevState s;
        std::memset(&s, 0, sizeof(s)); // padding too: graph signatures compare the bytes
        s.px = (const double *)px.p; s.py = (const double *)py.p; s.pz = (const double *)pz.p;
        s.q = (const double *)q.p;
        s.lj64 = (const double2 *)lj64.p; s.gb64 = (const double2 *)gb64.p;
        s.t123 = (const double *)t123.p;
        s.xyzq32 = (float4 *)xyzq32.p; s.lj32 = (const float2 *)lj32.p; s.gb32 = (const float2 *)gb32.p;
        s.pp32 = (float4 *)pp32.p;
        s.born = born.p; s.acc_g = acc_g.p; s.acc_deda = acc_deda.p;
        s.acc_fx = acc_fx.p; s.acc_fy = acc_fy.p; s.acc_fz = acc_fz.p;
        s.fx = (double *)fx.p; s.fy = (double *)fy.p; s.fz = (double *)fz.p;
        s.blk_inst = (const int *)blk_inst.p; s.blk_n = (const int *)blk_n.p; s.blk_first = (const int *)blk_first.p;
        s.blk_acc = (unsigned long long *)blk_acc.p;
        s.blk_ctr = (double4 *)blk_ctr.p; s.blk_half = (float4 *)blk_half.p; s.blk_box = (double4 *)blk_box.p;
        s.row_count = (int *)row_count.p; s.row_off = (int *)row_off.p; s.tiles = (int2 *)tiles.p;
        s.tile_slot = (int *)tile_slot.p;
        s.inst_mid = (const int *)inst_mid.p; s.inst_band_off = (const long long *)inst_band_off.p;
        s.inst_sp_off = (const int *)inst_sp_off.p; s.inst_sp_n = (const int *)inst_sp_n.p;
        s.band = (const int8_t *)band.p; s.sparse = (const int4 *)sparse.p;
        s.inst_energy = inst_rec.p;
        s.inst_count = inst_rec.p ? (unsigned long long *)((char *)inst_rec.p + (size_t)max_inst * kENum * 8) : nullptr;
        s.inst_status = inst_rec.p ? (int *)((char *)s.inst_count + (size_t)max_inst * kCNum * 8) : nullptr;
        s.inst_need = nullptr;
        s.any_need = nullptr;
        s.plist = (uint16_t *)plist.p;
        s.tile_cnt = (int2 *)tile_cnt.p;
        s.plist_cap = (int)std::min<int64_t>(plist_cap, INT32_MAX);
        s.overflow = (int *)overflow.p;
        s.rx = (double *)rx.p; s.ry = (double *)ry.p; s.rz = (double *)rz.p;
        s.inst_moved = (int *)inst_moved.p;
        s.cand = (uint16_t *)cand.p;
        s.cand_n = (int *)cand_n.p;
        s.inst_cslot = (const long long *)inst_cslot.p;
        return s;
    }

    // FF_Bonded records of instance id (PDGPU_ALL: every record slot)
    BondedState B(int id) const {
        BondedState b;
        std::memset(&b, 0, sizeof(b));
        b.bd_idx = (const int4 *)bd_idx.p; b.bd_par = (const double2 *)bd_par.p;
        b.an_idx = (const int4 *)an_idx.p; b.an_par = (const double2 *)an_par.p;
        b.dh_idx = (const int4 *)dh_idx.p; b.dh_meta = (const int *)dh_meta.p;
        b.dh_vn = (const double4 *)dh_vn.p; b.dh_n = (const double4 *)dh_n.p; b.dh_gamma = (const double4 *)dh_gamma.p;
        if (id == PDGPU_ALL) {
            b.b0 = 0; b.b1 = (int)cap_bd; b.a0 = 0; b.a1 = (int)cap_an; b.d0 = 0; b.d1 = (int)cap_dh;
        } else {
            const Instance &I = inst[id];
            b.b0 = (int)I.bd_off; b.b1 = (int)(I.bd_off + I.bd_len);
            b.a0 = (int)I.an_off; b.a1 = (int)(I.an_off + I.an_len);
            b.d0 = (int)I.dh_off; b.d1 = (int)(I.dh_off + I.dh_len);
        }
        b.comps = opt.components;
        return b;
    }
};

namespace {

int fail(pdgpu_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    return code;
}

#define HIPCHECK(c, x)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fail((c), PDGPU_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// (Re)allocate buf to hold n elements of esz bytes, preserving the old
// contents; new bytes are zeroed.
int resize(pdgpu_ctx *c, DevBuf &b, size_t n, size_t esz) {
    const size_t nb = std::max<size_t>(n * esz, 256);
    if (nb <= b.bytes) return PDGPU_OK;
    if (c->budget > 0 && c->used + (int64_t)(nb - b.bytes) > c->budget)
        return fail(c, PDGPU_ERR_NOMEM, "device memory budget exceeded");
    void *p = nullptr;
    if (hipMalloc(&p, nb) != hipSuccess) return fail(c, PDGPU_ERR_NOMEM, "hipMalloc failed");
    HIPCHECK(c, hipMemsetAsync(p, 0, nb, c->stream));
    if (b.p) {
        HIPCHECK(c, hipMemcpyAsync(p, b.p, b.bytes, hipMemcpyDeviceToDevice, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(b.p);
    }
    c->used += (int64_t)(nb - b.bytes);
    b.p = p;
    b.bytes = nb;
    return PDGPU_OK;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

#define TRY(x)                          \
    do {                                \
        int r_ = (x);                   \
        if (r_ != PDGPU_OK) return r_;  \
    } while (0)

// Staged upload: up_put / up_zero queue sections (the host bytes are
// copied into the blob at once), up_commit sends the blob in one copy and
// scatters it in one kernel, in stream order; the host does not wait (the
// pinned buffer is reused only after the previous upload's event).
void up_put(pdgpu_ctx *c, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    const size_t off = (c->up_blob.size() + 15) & ~(size_t)15;
    c->up_blob.resize(off + bytes);
    std::memcpy(c->up_blob.data() + off, src, bytes);
    c->up_secs.push_back({(unsigned long long)dst, (long long)off, (long long)bytes, 0});
}

void up_zero(pdgpu_ctx *c, void *dst, size_t bytes) {
    if (bytes) c->up_secs.push_back({(unsigned long long)dst, 0, (long long)bytes, 1});
}

int up_commit(pdgpu_ctx *c) {
    if (c->up_secs.empty()) return PDGPU_OK;
    const size_t toff = (c->up_blob.size() + 15) & ~(size_t)15;
    const size_t total = toff + c->up_secs.size() * sizeof(UpSec);
    c->up_blob.resize(total);
    std::memcpy(c->up_blob.data() + toff, c->up_secs.data(), c->up_secs.size() * sizeof(UpSec));
    const int nsec = (int)c->up_secs.size();
    c->up_secs.clear();
    if (c->up_done) HIPCHECK(c, hipEventSynchronize(c->up_done));
    if (c->h_up_bytes < total) {
        if (c->h_up) (void)hipHostFree(c->h_up);
        c->h_up = nullptr;
        c->h_up_bytes = 0;
        const size_t cap = std::max<size_t>(total, 1 << 20);
        HIPCHECK(c, hipHostMalloc(&c->h_up, cap, hipHostMallocDefault));
        c->h_up_bytes = cap;
    }
    std::memcpy(c->h_up, c->up_blob.data(), total);
    c->up_blob.clear();
    TRY(resize(c, c->upl, total, 1));
    HIPCHECK(c, hipMemcpyAsync(c->upl.p, c->h_up, total, hipMemcpyHostToDevice, c->stream));
    const unsigned char *blob = (const unsigned char *)c->upl.p;
    hipLaunchKernelGGL(k_scatter, dim3(nsec, 8), dim3(256), 0, c->stream, blob, (const UpSec *)(blob + toff));
    HIPCHECK(c, hipGetLastError());
    if (!c->up_done) HIPCHECK(c, hipEventCreateWithFlags(&c->up_done, hipEventDisableTiming));
    HIPCHECK(c, hipEventRecord(c->up_done, c->stream));
    return PDGPU_OK;
}

int grow_atoms(pdgpu_ctx *c, int64_t blocks) {
    const size_t na = (size_t)blocks * kBlk;
    DevBuf *d8[] = {&c->px, &c->py, &c->pz, &c->q, &c->t123, &c->acc_g, &c->acc_deda, &c->acc_fx, &c->acc_fy,
                    &c->acc_fz, &c->fx, &c->fy, &c->fz, &c->vx, &c->vy, &c->vz, &c->minv, &c->opx,
                    &c->opy, &c->opz, &c->rx, &c->ry, &c->rz, &c->fox, &c->foy, &c->foz};
    for (DevBuf *b : d8) TRY(resize(c, *b, na, 8));
    TRY(resize(c, c->is_h, na, 1));
    TRY(resize(c, c->lj64, na, 16));
    TRY(resize(c, c->gb64, na, 16));
    TRY(resize(c, c->born, na, c->mixed ? 4 : 8));
    if (c->mixed) {
        TRY(resize(c, c->xyzq32, na, 16));
        TRY(resize(c, c->lj32, na, 8));
        TRY(resize(c, c->gb32, na, 8));
        TRY(resize(c, c->pp32, na, 16));
    }
    TRY(resize(c, c->blk_inst, blocks, 4));
    TRY(resize(c, c->blk_n, blocks, 4));
    TRY(resize(c, c->blk_first, blocks, 4));
    TRY(resize(c, c->blk_ctr, blocks, 32));
    TRY(resize(c, c->blk_box, blocks, 32));
    TRY(resize(c, c->blk_half, blocks, 16));
    TRY(resize(c, c->blk_acc, (size_t)blocks * kBAcc, 8));
    TRY(resize(c, c->row_count, blocks, 4));
    TRY(resize(c, c->row_off, blocks + 1, 4));
    const int64_t old = c->cap_blocks;
    c->h_blk_inst.resize(blocks, -1);
    c->h_blk_n.resize(blocks, 0);
    c->h_blk_first.resize(blocks, 0);
    if (blocks > old) {
        // new blocks are inactive
        std::vector<int> neg(blocks - old, -1);
        HIPCHECK(c, hipMemcpyAsync((int *)c->blk_inst.p + old, neg.data(), neg.size() * 4, hipMemcpyHostToDevice,
                                   c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    c->blocks.grow(blocks);
    c->cap_blocks = blocks;
    return PDGPU_OK;
}

void setup_params(pdgpu_ctx *c) {
    const pdgpu_options &o = c->opt;
    PairParams &P = c->P;
    std::memset(&P, 0, sizeof(P));
    const double econv = econv_kcal();
    P.cut = o.cutoff;
    P.cut2 = o.cutoff * o.cutoff;
    P.inner = o.inner_cutoff;
    P.inner2 = o.inner_cutoff * o.inner_cutoff;
    // a switching width of zero (InnerCutoff >= Cutoff) switches nothing
    // inside the cutoff; 0 keeps the branch-free switch finite there
    P.invSw = o.cutoff > o.inner_cutoff ? 1.0 / (o.cutoff - o.inner_cutoff) : 0.0;
    P.vcut = o.vdw_cutoff;
    P.vinner = o.vdw_inner_cutoff;
    P.vinvSw = o.vdw_cutoff > o.vdw_inner_cutoff ? 1.0 / (o.vdw_cutoff - o.vdw_inner_cutoff) : 0.0;
    P.same_switch = P.inner == P.vinner && P.invSw == P.vinvSw;
    P.s14v = o.vdw14_scaling;
    P.s14e = o.elec14_scaling;
    P.econv_in = econv * (1.0 / o.dielectric_solute);
    P.econv_out = econv * (1.0 / o.dielectric_solvent);
    P.soft_h = o.softvdw_hardness;
    P.soft_olap = o.softvdw_olap_offset;
    P.soft_cut2 = o.softvdw_cutoff * o.softvdw_cutoff;
    P.econv_in32 = (float)P.econv_in;
    P.econv_out32 = (float)P.econv_out;
    P.premul = -(econv * 0.5) * (1 / o.dielectric_solute - 1 / o.dielectric_solvent);
    // the list radius is max(Cutoff, VdwCutoff) + Padding
    // (forcefields/nonbonded.cpp:1013, workspace/neighbourlist.cpp:258)
    const double lc = std::max(o.cutoff, o.vdw_cutoff);
    P.list_r = lc + o.nlist_padding;
    P.list_lim2 = (lc + o.nlist_padding) * (lc + o.nlist_padding);
    const double sc = std::min(o.gbsa_still_cutoff, P.list_r);
    P.still2 = sc * sc;
    P.invP5 = 1.0 / o.p5;
    P.P4 = o.p4;
    P.P5 = o.p5;
    P.exp_thresh = o.exp_approx_threshold;
    P.halfeconv = econv * 0.5;
    P.born_off = o.born_radius_offset;
    // Steinbach-Brooks force switch constants (forcefields/nonbonded.cpp:601-616)
    const double rc = o.cutoff, ri = o.inner_cutoff, rc2 = rc * rc, ri2 = ri * ri;
    P.fs_sA = 1.0 / ((rc2 - ri2) * (rc2 - ri2) * (rc2 - ri2));
    P.fs_sB = -(rc2 * rc2 * rc2 - 3.0 * rc2 * rc2 * ri2);
    P.fs_sC = 6.0 * rc2 * ri2;
    P.fs_sD = -(rc2 + ri2);
    P.fs_sE = 2.0 / 5.0;
    const double fin = P.fs_sA * (P.fs_sB * (1.0 / ri) + P.fs_sC * ri + P.fs_sD * ri * ri * ri +
                                  P.fs_sE * ri * ri * ri * ri2);
    P.fs_cutV = P.fs_sA * (P.fs_sB * (1.0 / rc) + P.fs_sC * rc + P.fs_sD * rc * rc * rc +
                           P.fs_sE * rc * rc * rc * rc2);
    P.fs_eshift = (ri < rc) ? 1 / ri + (fin - P.fs_cutV) : 0;
    // default: tiles whose candidates are expected to be at least 7/8 of
    // the 4,096 slots inside the cutoff, i.e. candidates >= 3584 times the
    // list-to-cutoff volume ratio (all-pairs GB, where the ratio is ~1); at
    // lower fill the filtered list with LDS accumulation is faster
    if (o.dense_tile_pairs > 0) {
        P.dense_pairs = o.dense_tile_pairs;
    } else {
        const double vr = std::pow(P.list_r / std::max(o.cutoff, 1e-6), 3.0);
        P.dense_pairs = (int)std::min<double>(kTileCap + 1, std::ceil(kTileCap / 8 * 7 * vr));
    }
    P.elec_mode = o.elec_mode;
    P.vdw_mode = o.vdw_mode;
    P.lg_split = 0;
    while ((1 << P.lg_split) < o.tile_split) ++P.lg_split;
}

bool valid_id(const pdgpu_ctx *c, int id) {
    return id >= 0 && id < c->max_inst && c->inst[id].active;
}

// block / atom / slot ranges of an evaluation
struct Span {
    int b0, b1, s0, s1;
};

Span span_of(const pdgpu_ctx *c, int id) {
    if (id == PDGPU_ALL) return {0, (int)c->cap_blocks, 0, c->max_inst};
    const Instance &I = c->inst[id];
    return {I.block0, I.block0 + I.nblocks, id, id + 1};
}

// Tile kernels whose grid shape the autotuner chooses.
enum { kGridEnum = 0, kGridBorn, kGridPair, kGridChain, kGridKinds };

int waves_cap(const pdgpu_ctx *c, int kind) { return std::max(64, c->num_cus * c->grid_wpc[kind]); }

hipEvent_t prof_event(pdgpu_ctx *c) {
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
    pdgpu_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    ProfScope(pdgpu_ctx *c_, int k) : c(c_), kind(k) {
        if (c->prof) {
            a = prof_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope() {
        if (c->prof && a) {
            hipEvent_t b = prof_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->prof_recs.push_back({kind, a, b});
            }
        }
    }
};

// Grid of a tile kernel for an estimated tile count. Unless the options fix
// tiles per wavefront, the kernel is passed tpw = 0 and divides the tile
// count it reads on the device over the waves it has, so a stale or bounding
// estimate only changes the grid size, never how many tiles one wavefront
// walks in sequence (the critical path of small, latency-bound launches).
void pair_grid(const pdgpu_ctx *c, int kind, int64_t tiles, int *groups, int *tpw) {
    const int wcap = waves_cap(c, kind);
    int t = c->opt.tiles_per_wave;
    if (t <= 0) t = (int)std::min<int64_t>(16, std::max<int64_t>(1, (tiles + wcap - 1) / wcap));
    int64_t waves = std::max<int64_t>(1, (tiles + t - 1) / t);
    waves = std::min<int64_t>(waves, wcap);
    int g = (int)((waves + kWavesPerGroup - 1) / kWavesPerGroup);
    if (g > 8) g = (g + 7) / 8 * 8;
    *groups = g;
    *tpw = c->opt.tiles_per_wave > 0 ? t : 0;
}

// Grows the per-tile pair-list buffer (kTileCap 2-byte entries per tile);
// its contents are rewritten by every evaluation, so nothing is copied.
int ensure_plist(pdgpu_ctx *c, int64_t tiles) {
    if (tiles <= c->plist_cap) return PDGPU_OK;
    const int64_t cap = std::max<int64_t>(tiles, c->plist_cap * 5 / 4 + 64);
    c->used -= (int64_t)(c->plist.bytes + c->tile_cnt.bytes);
    release(c->plist);
    release(c->tile_cnt);
    c->plist_cap = 0;
    TRY(resize(c, c->plist, (size_t)cap * kTileCap, 2));
    TRY(resize(c, c->tile_cnt, (size_t)cap << c->P.lg_split, 8)); // one record per segment
    c->plist_cap = cap;
    return PDGPU_OK;
}

// Tiles an evaluation of id walks, for sizing launch grids only (kernels
// read the true range from row_off): the count read at the last counted
// rebuild, else the upper bound.
// Tile count used to shape launch grids: the last counted total for all
// instances (or the triangle bound before any count), the instance's
// triangle bound for one instance. Kernels loop until the device-side count
// is covered, so an estimate only affects balance.
int64_t span_tiles(const pdgpu_ctx *c, int id) {
    if (id == PDGPU_ALL) return c->ntiles > 0 ? c->ntiles : c->tri_sum;
    const int nb = c->inst[id].nblocks;
    return (int64_t)nb * (nb + 1) / 2;
}

// Rebuilds the tile rows and candidate pairs of instance id (or all). With
// flagged, only the instances whose device flag inst_need is set are rebuilt
// (the flags are written by a kernel). Every instance owns a candidate region
// for its largest possible tile count and the pair-list buffer covers all
// tiles, so a rebuild needs nothing from the host; count = true reads the new
// tile total back to size later launch grids (a host round trip).
template <typename R>
int launch_build(pdgpu_ctx *c, int id, bool flagged = false, bool count = true, const int *gate = nullptr) {
    DevState S = c->S();
    if (flagged) S.inst_need = (const int *)c->inst_need.p;
    S.any_need = gate; // with a gate, the kernels return at once when nothing is flagged
    const Span sp = span_of(c, flagged ? PDGPU_ALL : id);
    const int nb = sp.b1 - sp.b0;
    ProfScope ps(c, PDGPU_KERNEL_NLIST);
    if (nb > 0) {
        hipLaunchKernelGGL(k_block_bounds, dim3((nb + 3) / 4), dim3(256), 0, c->stream, S, sp.b0, sp.b1);
        hipLaunchKernelGGL(k_tile_rows<0>, dim3((nb + 3) / 4), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1);
    }
    const int all = (int)c->cap_blocks;
    hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, c->stream, S, all);
    hipLaunchKernelGGL(k_tile_rows<1>, dim3((all + 3) / 4), dim3(256), 0, c->stream, S, c->P, 0, all);
    {
        // The new tile count is only known on the device here: size the grid
        // for the upper bound (every instance's full triangle, capped by the
        // grid limit) and let the kernel split the real count (tpw = 0).
        int g, t;
        const int64_t bound = (id == PDGPU_ALL || flagged) ? std::max<int64_t>(c->ntiles, c->tri_sum)
                                                           : span_tiles(c, id);
        pair_grid(c, kGridEnum, bound, &g, &t);
        // the mixed path enumerates from FP32 positions relative to the block
        // centres k_block_bounds has just set
        const int a0 = sp.b0 * kBlk, na = nb * kBlk;
        if (c->mixed && na > 0)
            hipLaunchKernelGGL(k_xyzq32, dim3((na + 255) / 256), dim3(256), 0, c->stream, S, a0, a0 + na);
        hipLaunchKernelGGL(k_cand<R>, dim3(g), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1, 0);
    }
    HIPCHECK(c, hipGetLastError());
    if (count) {
        int total = 0;
        HIPCHECK(c, hipMemcpyAsync(&total, (int *)c->row_off.p + all, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->ntiles = total;
    }
    if (flagged) return PDGPU_OK; // flagged builds only run on instances listed before
    if (id == PDGPU_ALL) {
        for (auto &I : c->inst)
            if (I.active) I.listed = true;
    } else {
        c->inst[id].listed = true;
    }
    return PDGPU_OK;
}

// Builds the lists of instances that have none yet, leaving the lists of
// the others as they are (a list is only rebuilt when its owner asks).
int ensure_listed(pdgpu_ctx *c, int id) {
    if (id != PDGPU_ALL) {
        if (c->inst[id].listed) return PDGPU_OK;
        return c->mixed ? launch_build<float>(c, id) : launch_build<double>(c, id);
    }
    std::vector<int> need(c->max_inst, 0);
    int n = 0;
    for (int s = 0; s < c->max_inst; ++s)
        if (c->inst[s].active && !c->inst[s].listed) { need[s] = 1; ++n; }
    if (n == 0) return PDGPU_OK;
    HIPCHECK(c, hipMemcpyAsync(c->inst_need.p, need.data(), need.size() * 4, hipMemcpyHostToDevice, c->stream));
    TRY(c->mixed ? launch_build<float>(c, PDGPU_ALL, true) : launch_build<double>(c, PDGPU_ALL, true));
    HIPCHECK(c, hipStreamSynchronize(c->stream)); // the host flag vector goes out of scope
    for (int s = 0; s < c->max_inst; ++s)
        if (need[s]) c->inst[s].listed = true;
    return PDGPU_OK;
}

// The mixed pair pass variant for the context's switching ranges and
// cutoffs (all list entries lie within Cutoff when GbsaStillCutoff does not
// exceed it).
using PairKernel = void (*)(DevState, PairParams, int, int, int);
PairKernel pair_mixed_kernel(const PairParams &P) {
    const bool allin = P.still2 <= P.cut2;
    if (P.same_switch) return allin ? k_pair_mixed<true, true> : k_pair_mixed<true, false>;
    return allin ? k_pair_mixed<false, true> : k_pair_mixed<false, false>;
}

// The GB kernels of one evaluation: enumeration, Born sums, Born radii,
// pair pass and chain rule (MODE 0) or the energy-path pair pass (MODE 1).
// With nbonded > 0 (MODE 0) the bonded terms B run as extra workgroups of
// the chain-rule launch.
template <typename R>
int launch_gb(pdgpu_ctx *c, const DevState &S, const Span &sp, int64_t tiles, int mode, const BondedState &B,
              int nbonded) {
    const int a0 = sp.b0 * kBlk, a1 = sp.b1 * kBlk, na = a1 - a0;
    int g[kGridKinds], t[kGridKinds];
    for (int k = 0; k < kGridKinds; ++k) pair_grid(c, k, tiles << c->P.lg_split, &g[k], &t[k]); // work items
    {
        ProfScope ps(c, PDGPU_KERNEL_ENUM);
        hipLaunchKernelGGL(k_filter<R>, dim3(g[kGridEnum]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                           t[kGridEnum]);
    }
    {
        ProfScope ps(c, PDGPU_KERNEL_BORN);
        hipLaunchKernelGGL(k_born<R>, dim3(g[kGridBorn]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                           t[kGridBorn]);
    }
    {
        ProfScope ps(c, PDGPU_KERNEL_ATOM);
        hipLaunchKernelGGL(k_born_final<R>, dim3((na + 255) / 256), dim3(256), 0, c->stream, S, c->P, a0, a1);
    }
    if (mode == 0) {
        {
            ProfScope ps(c, PDGPU_KERNEL_PAIR);
            if constexpr (std::is_same<R, float>::value)
                hipLaunchKernelGGL(pair_mixed_kernel(c->P), dim3(g[kGridPair]),
                                   dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1, t[kGridPair]);
            else
                hipLaunchKernelGGL((k_pair<R, 0>), dim3(g[kGridPair]), dim3(256), 0, c->stream, S, c->P, sp.b0,
                                   sp.b1, t[kGridPair]);
        }
        ProfScope ps(c, PDGPU_KERNEL_CHAIN);
        if (nbonded > 0)
            hipLaunchKernelGGL(k_chain_bonded<R>, dim3(g[kGridChain] + (nbonded + 255) / 256), dim3(256), 0,
                               c->stream, S, c->P, sp.b0, sp.b1, t[kGridChain], g[kGridChain], B);
        else
            hipLaunchKernelGGL(k_chain<R>, dim3(g[kGridChain]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                               t[kGridChain]);
    } else {
        ProfScope ps(c, PDGPU_KERNEL_PAIR);
        hipLaunchKernelGGL((k_pair<R, 1>), dim3(g[kGridPair]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                           t[kGridPair]);
    }
    return PDGPU_OK;
}

// One evaluation: MODE 0 = calcForces, MODE 1 = calcEnergies. Without
// finalize the caller launches a kernel that finalises the forces itself
// (the velocity-Verlet kick). prepared: the set-up kernel's work was done by
// the previous launch (a velocity-Verlet closing kernel), so the evaluation
// starts without it.
template <typename R>
int launch_eval(pdgpu_ctx *c, int id, int mode, bool finalize = true, double *f2 = nullptr, int f2_stride = 0,
                bool prepared = false) {
    const DevState S = c->S();
    const Span sp = span_of(c, id);
    const int a0 = sp.b0 * kBlk, a1 = sp.b1 * kBlk;
    const int na = a1 - a0;
    const int ns = sp.s1 - sp.s0;
    const bool gb = c->opt.components & PDGPU_FF_GB;
    const bool soft = c->opt.components & PDGPU_FF_SOFTVDW;
    const bool nonb = c->opt.components & PDGPU_FF_NONBONDED;
    {
        ProfScope ps(c, PDGPU_KERNEL_ATOM);
        const int nt = std::max(na, ns);
        if (gb) {
            if (!prepared)
                hipLaunchKernelGGL(k_prepare<R>, dim3((nt + 255) / 256), dim3(256), 0, c->stream, S, a0, a1, sp.s0,
                                   sp.s1);
        } else
            hipLaunchKernelGGL(k_clear_grad<R>, dim3((nt + 255) / 256), dim3(256), 0, c->stream, S, a0, a1, sp.s0,
                               sp.s1);
        if ((soft || nonb) && c->mixed) // block-relative FP32 positions (k_prepare's job when GB runs)
            hipLaunchKernelGGL(k_xyzq32, dim3((na + 255) / 256), dim3(256), 0, c->stream, S, a0, a1);
    }
    const BondedState B = c->B(id);
    const int nterms = (B.b1 - B.b0) + (B.a1 - B.a0) + (B.d1 - B.d0);
    const bool bonded = (c->opt.components & PDGPU_FF_BONDED) && nterms > 0;
    // with GB forces the bonded terms ride along with the chain-rule launch
    const bool fused = bonded && gb && mode == 0;
    if (gb) TRY(launch_gb<R>(c, S, sp, span_tiles(c, id), mode, B, fused ? nterms : 0));
    if (nonb) {
        // FF_NonBonded alone: the energy path's pair term without the GB
        // energy (MODE 2); its calcEnergies is the same evaluation
        int g[kGridKinds], t[kGridKinds];
        for (int k = 0; k < kGridKinds; ++k) pair_grid(c, k, span_tiles(c, id) << c->P.lg_split, &g[k], &t[k]);
        {
            ProfScope ps(c, PDGPU_KERNEL_ENUM);
            hipLaunchKernelGGL(k_filter<R>, dim3(g[kGridEnum]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                               t[kGridEnum]);
        }
        ProfScope ps(c, PDGPU_KERNEL_PAIR);
        hipLaunchKernelGGL((k_pair<R, 2>), dim3(g[kGridPair]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                           t[kGridPair]);
    }
    if (soft) {
        // FF_SoftVDW over the cutoff-filtered lists (calcEnergies is its
        // calcForces, ffsoftvdw.cpp:56-59)
        int g[kGridKinds], t[kGridKinds];
        for (int k = 0; k < kGridKinds; ++k) pair_grid(c, k, span_tiles(c, id) << c->P.lg_split, &g[k], &t[k]);
        {
            ProfScope ps(c, PDGPU_KERNEL_ENUM);
            hipLaunchKernelGGL(k_filter<R>, dim3(g[kGridEnum]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                               t[kGridEnum]);
        }
        ProfScope ps(c, PDGPU_KERNEL_PAIR);
        hipLaunchKernelGGL(k_soft<R>, dim3(g[kGridPair]), dim3(256), 0, c->stream, S, c->P, sp.b0, sp.b1,
                           t[kGridPair]);
    }
    if (bonded && !fused) {
        ProfScope ps(c, PDGPU_KERNEL_BONDED);
        hipLaunchKernelGGL(k_bonded<R>, dim3((nterms + 255) / 256), dim3(256), 0, c->stream, S, B);
    }
    if (finalize) {
        ProfScope ps(c, PDGPU_KERNEL_ATOM);
        hipLaunchKernelGGL(k_finalize<R>, dim3((na + 255) / 256), dim3(256), 0, c->stream, S, a0, a1, f2,
                           f2_stride);
    }
    HIPCHECK(c, hipGetLastError());
    return PDGPU_OK;
}


// Converts device per-instance records into PD's units.
int read_energies(pdgpu_ctx *c, int id, pdgpu_energies *out, int mode) {
    const Span sp = span_of(c, id);
    const int ns = sp.s1 - sp.s0;
    // host-mapped pinned staging, filled by k_publish on the stream
    const size_t eb = (size_t)ns * kENum * 8, cb = (size_t)ns * kCNum * 8, sb = (size_t)ns * 4;
    if (c->h_stage_bytes < eb + cb + sb) {
        if (c->h_stage) (void)hipHostFree(c->h_stage);
        c->h_stage = c->d_stage = nullptr;
        c->h_stage_bytes = 0;
        HIPCHECK(c, hipHostMalloc(&c->h_stage, eb + cb + sb, hipHostMallocMapped));
        HIPCHECK(c, hipHostGetDevicePointer(&c->d_stage, c->h_stage, 0));
        c->h_stage_bytes = eb + cb + sb;
    }
    const long long *e = (const long long *)c->h_stage;
    const unsigned long long *cnt = (const unsigned long long *)((char *)c->h_stage + eb);
    const int *st = (const int *)((char *)c->h_stage + eb + cb);
    const int nblk = (int)std::min<size_t>(256, ((size_t)ns * kENum + 255) / 256);
    if (ns > 0) {
        hipLaunchKernelGGL(k_publish, dim3(nblk), dim3(256), 0, c->stream, c->S(), sp.s0, ns,
                           (unsigned long long *)c->d_stage);
        HIPCHECK(c, hipGetLastError());
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    int rc = PDGPU_OK;
    for (int k = 0; k < ns; ++k) {
        const int slot = sp.s0 + k;
        if (!c->inst[slot].active) continue;
        double v[kENum];
        for (int q = 0; q < kENum; ++q) {
            const long long raw = e[(size_t)k * kENum + q];
            double x;
            if (c->mixed) x = (double)raw / kEnergyScale;
            else std::memcpy(&x, &raw, 8);
            v[q] = x * kJPerKcal;
        }
        pdgpu_energies E;
        std::memset(&E, 0, sizeof(E));
        E.epot_vdw = v[kEVdw];
        E.epot_elec = v[kEElec];
        E.epot_pol_cross = v[kEPolCross];
        E.epot_pol_self = v[kEPolSelf];
        E.epot_pol = v[kEPolCross] + v[kEPolSelf];
        // the GB force path adds no virial; the energy path and bond
        // stretching do (ffbonded.cpp:924)
        E.virial = v[kEVirial];
        E.epot_bond = v[kEBond];
        E.epot_angle = v[kEAngle];
        E.epot_torsion = v[kETorsion] + v[kEImproper]; // calcImproperForces adds to epot_torsion
        E.epot_improper = v[kEImproper];
        E.epot = E.epot_vdw + E.epot_elec + E.epot_pol + E.epot_bond + E.epot_angle + E.epot_torsion;
        E.pairs = (int64_t)cnt[(size_t)k * kCNum + kCPairs];
        E.born_pairs = (int64_t)cnt[(size_t)k * kCNum + kCBornPairs];
        const bool bad = st[k] != 0 || !std::isfinite(E.epot);
        E.status = bad ? PDGPU_ERR_NONFINITE : PDGPU_OK;
        if (bad) {
            E.epot = NAN; // in-band NaN, as the reference signals instability
            rc = PDGPU_ERR_NONFINITE;
        }
        if (out) {
            if (id == PDGPU_ALL) out[slot] = E;
            else *out = E;
        }
    }
    if (rc != PDGPU_OK) c->err = "non-finite or out-of-range result";
    return rc;
}

// Gather per-atom doubles of one instance from the padded layout.
int gather3(pdgpu_ctx *c, int id, const DevBuf &x, const DevBuf &y, const DevBuf &z, double *out, double scale) {
    const Instance &I = c->inst[id];
    const size_t a0 = (size_t)I.block0 * kBlk;
    std::vector<double> hx(I.natoms), hy(I.natoms), hz(I.natoms);
    HIPCHECK(c, hipMemcpyAsync(hx.data(), (double *)x.p + a0, I.natoms * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(hy.data(), (double *)y.p + a0, I.natoms * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(hz.data(), (double *)z.p + a0, I.natoms * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < I.natoms; ++i) {
        out[3 * i] = hx[i] * scale;
        out[3 * i + 1] = hy[i] * scale;
        out[3 * i + 2] = hz[i] * scale;
    }
    return PDGPU_OK;
}

int scatter3(pdgpu_ctx *c, int id, DevBuf &x, DevBuf &y, DevBuf &z, const double *in, double scale) {
    const Instance &I = c->inst[id];
    const size_t a0 = (size_t)I.block0 * kBlk;
    std::vector<double> hx(I.natoms), hy(I.natoms), hz(I.natoms);
    for (int i = 0; i < I.natoms; ++i) {
        hx[i] = in[3 * i] * scale;
        hy[i] = in[3 * i + 1] * scale;
        hz[i] = in[3 * i + 2] * scale;
    }
    up_put(c, (double *)x.p + a0, hx.data(), I.natoms * 8);
    up_put(c, (double *)y.p + a0, hy.data(), I.natoms * 8);
    up_put(c, (double *)z.p + a0, hz.data(), I.natoms * 8);
    return up_commit(c);
}

// Reserves n records in one FF_Bonded record family (bufs / esz: its
// arrays and their record sizes), growing the arrays when needed. New
// array space is zeroed, which marks its records free.
int term_alloc(pdgpu_ctx *c, Ranges &r, int64_t &cap, int64_t n, DevBuf *const *bufs, const size_t *esz, int nbuf,
               int64_t *off) {
    *off = 0;
    if (n == 0) return PDGPU_OK;
    int64_t o = r.alloc(n);
    if (o < 0) {
        const int64_t need = cap - r.tail_free() + n;
        const int64_t nc = std::max<int64_t>(std::max<int64_t>(need, cap * 3 / 2), 1024);
        if (nc > INT32_MAX) return fail(c, PDGPU_ERR_NOMEM, "too many bonded terms");
        for (int k = 0; k < nbuf; ++k) TRY(resize(c, *bufs[k], (size_t)nc, esz[k]));
        r.grow(nc);
        cap = nc;
        o = r.alloc(n);
        if (o < 0) return fail(c, PDGPU_ERR_NOMEM, "bonded term allocation failed");
    }
    *off = o;
    return PDGPU_OK;
}

// Frees the FF_Bonded records of an instance (zeroed records are skipped).
int bonded_release(pdgpu_ctx *c, Instance &I) {
    if (I.bd_len) HIPCHECK(c, hipMemsetAsync((int4 *)c->bd_idx.p + I.bd_off, 0, I.bd_len * 16, c->stream));
    if (I.an_len) HIPCHECK(c, hipMemsetAsync((int4 *)c->an_idx.p + I.an_off, 0, I.an_len * 16, c->stream));
    if (I.dh_len) HIPCHECK(c, hipMemsetAsync((int *)c->dh_meta.p + I.dh_off, 0, I.dh_len * 4, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->bdr.release(I.bd_off, I.bd_len);
    c->anr.release(I.an_off, I.an_len);
    c->dhr.release(I.dh_off, I.dh_len);
    I.bd_off = I.bd_len = I.an_off = I.an_len = I.dh_off = I.dh_len = 0;
    return PDGPU_OK;
}

// Bytes of everything a captured step passes to its kernels: device
// pointers, ranges, grid shapes and parameters. A graph is replayed only
// while these are unchanged (buffers move when they grow, spans when
// instances come and go).
std::vector<char> graph_sig(const pdgpu_ctx *c, int id) {
    const DevState S = c->S();
    const BondedState B = c->B(id);
    const Span sp = span_of(c, id);
    const int64_t nt = span_tiles(c, id);
    std::vector<char> v(sizeof(S) + sizeof(B) + sizeof(sp) + sizeof(nt) + sizeof(c->grid_wpc) + sizeof(c->P));
    char *p = v.data();
    std::memcpy(p, &S, sizeof(S)); p += sizeof(S);
    std::memcpy(p, &B, sizeof(B)); p += sizeof(B);
    std::memcpy(p, &sp, sizeof(sp)); p += sizeof(sp);
    std::memcpy(p, &nt, sizeof(nt)); p += sizeof(nt);
    std::memcpy(p, c->grid_wpc, sizeof(c->grid_wpc)); p += sizeof(c->grid_wpc);
    std::memcpy(p, &c->P, sizeof(c->P));
    return v;
}

// Launches the kernels `issue` puts on the context stream as one HIP graph,
// captured on first use and again whenever the signature changes. Only
// kernel launches are captured (no memset or copy nodes).
template <typename F>
int graph_run(pdgpu_ctx *c, const std::string &key, const std::vector<char> &sig, F &&issue) {
    pdgpu_ctx::StepGraph &g = c->graphs[key];
    if (!g.exec || g.sig != sig) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.sig.clear();
        hipGraph_t graph = nullptr;
        HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = issue();
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc != PDGPU_OK) {
            if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIPCHECK(c, e);
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g.exec = nullptr;
            return fail(c, PDGPU_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        g.sig = sig;
    }
    HIPCHECK(c, hipGraphLaunch(g.exec, c->stream));
    return PDGPU_OK;
}

bool use_graphs(const pdgpu_ctx *c) { return c->opt.use_graphs != 0 && !c->prof; }

// One evaluation of `id` (or every instance). Lists are brought up to date
// eagerly (a rebuild reads counts back); the kernels of the evaluation are
// then replayed as a HIP graph when graphs are on.
int eval(pdgpu_ctx *c, int id, int mode) {
    if (c->opt.components & (PDGPU_FF_GB | PDGPU_FF_SOFTVDW | PDGPU_FF_NONBONDED)) TRY(ensure_listed(c, id));
    auto issue = [&] { return c->mixed ? launch_eval<float>(c, id, mode) : launch_eval<double>(c, id, mode); };
    if (!use_graphs(c)) return issue();
    return graph_run(c, "calc:" + std::to_string(id) + ":" + std::to_string(mode), graph_sig(c, id), issue);
}

} // namespace

// =================================================================== C ABI

extern "C" const char *pdgpu_last_error(const pdgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int pdgpu_ctx_create(const pdgpu_options *o, pdgpu_ctx **out) {
    if (!o || !out) return PDGPU_ERR_INVALID;
    *out = nullptr;
    if (o->fast_mode != 1) return PDGPU_ERR_UNSUPPORTED;
    if (o->precision != PDGPU_FP64 && o->precision != PDGPU_MIXED) return PDGPU_ERR_INVALID;
    if ((o->components & ~(PDGPU_FF_ALL | PDGPU_FF_SOFTVDW | PDGPU_FF_NONBONDED)) != 0) return PDGPU_ERR_INVALID;
    {
        const int pairwise = !!(o->components & PDGPU_FF_GB) + !!(o->components & PDGPU_FF_SOFTVDW) +
                             !!(o->components & PDGPU_FF_NONBONDED);
        if (pairwise > 1) return PDGPU_ERR_UNSUPPORTED; // one pair component per context
    }
    if ((o->components & PDGPU_FF_NONBONDED) && o->elec_mode != 0 && o->elec_mode != 1 && o->elec_mode != 3 &&
        o->elec_mode != 4)
        return PDGPU_ERR_INVALID;
    // instance slots are stored in 24-bit fields (term records, noise counters)
    if (o->max_instances >= (1 << 24)) return PDGPU_ERR_INVALID;
    if (!(o->cutoff > 0) || !(o->vdw_cutoff > 0) || o->inner_cutoff > o->cutoff ||
        o->vdw_inner_cutoff > o->vdw_cutoff || !(o->p5 > 0) || !(o->dielectric_solute > 0) ||
        !(o->dielectric_solvent > 0) || o->nlist_padding < 0)
        return PDGPU_ERR_INVALID;
    // segments per tile: 0 (= 1) or a power of two up to kMaxSplit
    if (o->tile_split < 0 || o->tile_split > kMaxSplit || (o->tile_split & (o->tile_split - 1)) != 0)
        return PDGPU_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= o->device || o->device < 0) return PDGPU_ERR_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o->device) != hipSuccess) return PDGPU_ERR_NODEVICE;
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return PDGPU_ERR_NODEVICE;
    if (hipSetDevice(o->device) != hipSuccess) return PDGPU_ERR_NODEVICE;
    // the device address of the static scratch array, once per device
    if (scratch_setup(o->device) != PDGPU_OK) return PDGPU_ERR_HIP;
    pdgpu_ctx *c = new pdgpu_ctx();
    c->opt = *o;
    c->device = o->device;
    c->num_cus = prop.multiProcessorCount;
    c->mixed = o->precision == PDGPU_MIXED;
    c->max_inst = o->max_instances > 0 ? o->max_instances : 8192;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PDGPU_ERR_HIP;
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    c->budget = o->device_memory_budget > 0 ? o->device_memory_budget : (int64_t)(0.8 * (double)fr);
    setup_params(c);
    c->inst.resize(c->max_inst);
    int rc = PDGPU_OK;
    const int ms = c->max_inst;
    if (rc == PDGPU_OK) rc = resize(c, c->inst_nat, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_need, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_moved, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->any_need, 1, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_mid, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_band_off, ms, 8);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_sp_off, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_sp_n, ms, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_rec, (size_t)ms * (kENum * 8 + kCNum * 8 + 4), 1);
    if (rc == PDGPU_OK) rc = resize(c, c->inst_md, (size_t)ms, sizeof(InstMD));
    if (rc == PDGPU_OK) rc = resize(c, c->inst_cslot, (size_t)ms, 8);
    if (rc == PDGPU_OK) rc = grow_atoms(c, 64);
    if (rc == PDGPU_OK) rc = resize(c, c->band, 4096, 1);
    if (rc == PDGPU_OK) rc = resize(c, c->sparse, 64, 16);
    if (rc == PDGPU_OK) rc = resize(c, c->tiles, 4096, 8);
    if (rc == PDGPU_OK) rc = resize(c, c->tile_slot, 4096, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->overflow, 1, 4);
    if (rc == PDGPU_OK) rc = resize(c, c->rng_ctr, 1, 8);
    if (rc != PDGPU_OK) {
        pdgpu_ctx_destroy(c);
        return rc;
    }
    c->cap_band = c->band.bytes;
    c->bandr.init(c->cap_band);
    c->cap_sparse = c->sparse.bytes / 16;
    c->spr.init(c->cap_sparse);
    c->cap_tiles = c->tiles.bytes / 8;
    c->bdr.init(0);
    c->anr.init(0);
    c->dhr.init(0);
    c->csr.init(0);
    *out = c;
    return PDGPU_OK;
}

extern "C" void pdgpu_ctx_destroy(pdgpu_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf *all[] = {&c->px, &c->py, &c->pz, &c->q, &c->lj64, &c->gb64, &c->t123, &c->xyzq32, &c->lj32, &c->gb32,
                     &c->pp32, &c->born, &c->acc_g, &c->acc_deda, &c->acc_fx, &c->acc_fy, &c->acc_fz, &c->fx, &c->fy, &c->fz,
                     &c->vx, &c->vy, &c->vz, &c->minv, &c->opx, &c->opy, &c->opz, &c->blk_inst,
                     &c->blk_n, &c->blk_first, &c->blk_ctr, &c->blk_half, &c->blk_box, &c->blk_acc, &c->row_count, &c->row_off,
                     &c->tiles, &c->tile_slot, &c->inst_need, &c->inst_nat, &c->inst_mid, &c->inst_band_off, &c->inst_sp_off, &c->inst_sp_n, &c->inst_rec,
                     &c->inst_md, &c->band, &c->sparse, &c->plist, &c->tile_cnt, &c->overflow,
                     &c->bd_idx, &c->bd_par, &c->an_idx, &c->an_par, &c->dh_idx, &c->dh_meta, &c->dh_vn, &c->dh_n,
                     &c->dh_gamma, &c->cand, &c->cand_n, &c->inst_cslot, &c->rx, &c->ry, &c->rz,
                     &c->inst_moved, &c->any_need, &c->fox, &c->foy, &c->foz, &c->rng_ctr, &c->stage, &c->upl, &c->is_h};
    for (DevBuf *b : all) release(*b);
    for (auto &t : c->traj) release(t.second.ring);
    for (auto &g : c->graphs)
        if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    for (auto &r : c->prof_recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_up) (void)hipHostFree(c->h_up);
    if (c->up_done) (void)hipEventDestroy(c->up_done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int64_t pdgpu_instance_bytes(const pdgpu_ctx *c, int32_t natoms, int32_t mid, int32_t nsparse) {
    if (!c || natoms <= 0) return -1;
    const int64_t nb = (natoms + kBlk - 1) / kBlk;
    // the 26 double arrays of grow_atoms, the hydrogen flags, the FP64 LJ and
    // GB records, the Born radii and the mixed path's FP32 records
    const int64_t per_atom = 8 * 26 + 1 + 16 * 2 + (c->mixed ? 4 + 16 + 8 + 8 + 16 : 8);
    const int64_t per_block = 4 * 3 + 32 * 2 + 16 + 8 * kBAcc + 4 * 2;
    // per possible tile: the tile record, a candidate slot, a pair-list slot
    // and its segment counts
    const int64_t per_tile = 8 + 2 * kTileCap + 4 + 2 * kTileCap + (8 << c->P.lg_split);
    return nb * kBlk * per_atom + nb * per_block + (int64_t)mid * nb * kBlk + nsparse * 16 +
           nb * (nb + 1) / 2 * per_tile;
}

extern "C" int64_t pdgpu_bytes_free(const pdgpu_ctx *c) { return c ? c->budget - c->used : 0; }

extern "C" int pdgpu_instance_natoms(const pdgpu_ctx *c, int32_t id) {
    if (!c || !valid_id(c, id)) return PDGPU_ERR_INVALID;
    return c->inst[id].natoms;
}

extern "C" int pdgpu_instance_add(pdgpu_ctx *c, const pdgpu_topology *t, int32_t *id_out) {
    if (!c || !t || !id_out) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    const int n = t->natoms;
    if (n <= 0 || n > (1 << 24) - 1 || !t->charge || !t->radius || !t->epsilon || !t->gb_radius || !t->vj ||
        !t->terms123 || t->max_index_delta < 1 || (n > 1 && !t->band) || t->nsparse < 0 ||
        (t->nsparse > 0 && (!t->sparse_i || !t->sparse_j || !t->sparse_order)))
        return fail(c, PDGPU_ERR_INVALID, "invalid topology");
    int slot = -1;
    for (int k = 0; k < c->max_inst; ++k)
        if (!c->inst[k].active) { slot = k; break; }
    if (slot < 0) return fail(c, PDGPU_ERR_NOMEM, "no free instance slot");
    const int nb = (n + kBlk - 1) / kBlk;
    // blocks
    int64_t b0 = c->blocks.alloc(nb);
    if (b0 < 0) {
        const int64_t need = c->cap_blocks - c->blocks.tail_free() + nb;
        TRY(grow_atoms(c, std::max<int64_t>(need, c->cap_blocks * 3 / 2)));
        b0 = c->blocks.alloc(nb);
        if (b0 < 0) return fail(c, PDGPU_ERR_NOMEM, "block allocation failed");
    }
    // band: mid*(n-1) entries are uploaded; the range is allocated for every
    // padded atom so that the enumeration can stage 64 whole rows per block
    const int64_t bcopy = (int64_t)t->max_index_delta * (n - 1);
    const int64_t blen = (int64_t)t->max_index_delta * nb * kBlk;
    int64_t boff = c->bandr.alloc(blen);
    if (boff < 0) {
        const int64_t need = c->cap_band - c->bandr.tail_free() + blen;
        const int64_t nc = std::max<int64_t>(need, c->cap_band * 3 / 2);
        int rc = resize(c, c->band, nc, 1);
        if (rc != PDGPU_OK) { c->blocks.release(b0, nb); return rc; }
        c->bandr.grow(nc);
        c->cap_band = nc;
        boff = c->bandr.alloc(blen);
    }
    // sparse
    int64_t soff = c->spr.alloc(t->nsparse);
    if (soff < 0) {
        const int64_t need = c->cap_sparse - c->spr.tail_free() + t->nsparse;
        const int64_t nc = std::max<int64_t>(need, c->cap_sparse * 3 / 2);
        int rc = resize(c, c->sparse, nc, 16);
        if (rc != PDGPU_OK) { c->blocks.release(b0, nb); c->bandr.release(boff, blen); return rc; }
        c->spr.grow(nc);
        c->cap_sparse = nc;
        soff = c->spr.alloc(t->nsparse);
    }
    // tiles: bound of sum nb(nb+1)/2
    const int64_t tri = c->tri_sum + (int64_t)nb * (nb + 1) / 2;
    if (tri > c->cap_tiles) {
        const int64_t nc = std::max<int64_t>(tri, c->cap_tiles * 3 / 2);
        int rc = resize(c, c->tiles, nc, 8);
        if (rc == PDGPU_OK) rc = resize(c, c->tile_slot, nc, 4);
        if (rc != PDGPU_OK) {
            c->blocks.release(b0, nb); c->bandr.release(boff, blen); c->spr.release(soff, t->nsparse);
            return rc;
        }
        c->cap_tiles = nc;
    }
    // pair lists: a candidate region of nb(nb+1)/2 tile slots (the largest
    // tile count the instance can have) and pair-list slots for every tile
    const int64_t tslots = (int64_t)nb * (nb + 1) / 2;
    int64_t csoff = c->csr.alloc(tslots);
    int rcl = PDGPU_OK;
    if (csoff < 0) {
        const int64_t need_cap = c->cap_cs - c->csr.tail_free() + tslots;
        const int64_t nc = std::max<int64_t>(need_cap, c->cap_cs * 5 / 4 + 64);
        rcl = resize(c, c->cand, (size_t)nc * kTileCap, 2);
        if (rcl == PDGPU_OK) rcl = resize(c, c->cand_n, (size_t)nc, 4);
        if (rcl == PDGPU_OK) {
            c->csr.grow(nc);
            c->cap_cs = nc;
            csoff = c->csr.alloc(tslots);
        }
    }
    if (rcl == PDGPU_OK) rcl = ensure_plist(c, tri);
    if (rcl != PDGPU_OK || csoff < 0) {
        c->blocks.release(b0, nb); c->bandr.release(boff, blen); c->spr.release(soff, t->nsparse);
        if (csoff >= 0) c->csr.release(csoff, tslots);
        return rcl != PDGPU_OK ? rcl : fail(c, PDGPU_ERR_NOMEM, "candidate slot allocation failed");
    }
    c->tri_sum = tri;

    // upload per-atom data in the padded layout
    const size_t na = (size_t)nb * kBlk;
    const size_t a0 = (size_t)b0 * kBlk;
    const double sqrt_j = std::sqrt(kJPerKcal);
    std::vector<double> q(na, 0.0), t123(na, 0.0), minv(na, 0.0);
    std::vector<double2> lj(na, make_double2(0, 0)), gb(na, make_double2(1.0, 0.0));
    for (int i = 0; i < n; ++i) {
        q[i] = t->charge[i];
        lj[i] = make_double2(t->radius[i], t->epsilon[i] / sqrt_j);
        gb[i] = make_double2(t->gb_radius[i], t->vj[i]);
        t123[i] = t->terms123[i];
        minv[i] = (t->mass && t->mass[i] > 0) ? 1.0 / t->mass[i] : 0.0;
    }
    // one staged upload of everything below (no host wait)
    up_put(c, (double *)c->q.p + a0, q.data(), na * 8);
    up_put(c, (double *)c->t123.p + a0, t123.data(), na * 8);
    up_put(c, (double *)c->minv.p + a0, minv.data(), na * 8);
    up_put(c, (double2 *)c->lj64.p + a0, lj.data(), na * 16);
    up_put(c, (double2 *)c->gb64.p + a0, gb.data(), na * 16);
    if (c->mixed) {
        std::vector<float2> lj32(na), gb32(na);
        for (size_t i = 0; i < na; ++i) {
            lj32[i] = make_float2((float)lj[i].x, (float)lj[i].y);
            gb32[i] = make_float2((float)gb[i].x, (float)gb[i].y);
        }
        up_put(c, (float2 *)c->lj32.p + a0, lj32.data(), na * 8);
        up_put(c, (float2 *)c->gb32.p + a0, gb32.data(), na * 8);
    }
    DevBuf *zs[] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->fox, &c->foy, &c->foz};
    for (DevBuf *b : zs) up_zero(c, (double *)b->p + a0, na * 8);
    up_zero(c, (uint8_t *)c->is_h.p + a0, na);
    // per block
    std::vector<int> bi(nb, slot), bn(nb), bf(nb, (int)b0);
    for (int k = 0; k < nb; ++k) bn[k] = std::min(kBlk, n - k * kBlk);
    up_put(c, (int *)c->blk_inst.p + b0, bi.data(), nb * 4);
    up_put(c, (int *)c->blk_n.p + b0, bn.data(), nb * 4);
    up_put(c, (int *)c->blk_first.p + b0, bf.data(), nb * 4);
    for (int k = 0; k < nb; ++k) {
        c->h_blk_inst[b0 + k] = slot;
        c->h_blk_n[b0 + k] = bn[k];
        c->h_blk_first[b0 + k] = (int)b0;
    }
    // bond orders
    if (bcopy > 0) up_put(c, (int8_t *)c->band.p + boff, t->band, bcopy);
    std::vector<int4> sp(t->nsparse);
    for (int k = 0; k < t->nsparse; ++k) {
        int i = t->sparse_i[k], j = t->sparse_j[k];
        if (i > j) std::swap(i, j);
        sp[k] = make_int4(i, j, t->sparse_order[k], 0);
    }
    if (t->nsparse > 0) up_put(c, (int4 *)c->sparse.p + soff, sp.data(), sp.size() * 16);
    const int mid = t->max_index_delta;
    const long long boffl = boff;
    const int soffi = (int)soff, sn = t->nsparse;
    up_put(c, (int *)c->inst_mid.p + slot, &mid, 4);
    up_put(c, (int *)c->inst_nat.p + slot, &n, 4);
    up_put(c, (long long *)c->inst_band_off.p + slot, &boffl, 8);
    up_put(c, (int *)c->inst_sp_off.p + slot, &soffi, 4);
    up_put(c, (int *)c->inst_sp_n.p + slot, &sn, 4);
    const long long csl = csoff;
    up_put(c, (long long *)c->inst_cslot.p + slot, &csl, 8);
    TRY(up_commit(c));

    Instance &I = c->inst[slot];
    I = Instance();
    I.active = true;
    I.natoms = n;
    I.nblocks = nb;
    I.block0 = (int)b0;
    I.mid = mid;
    I.nsparse = t->nsparse;
    I.band_off = boff;
    I.band_len = blen;
    I.sp_off = soff;
    I.cs_off = csoff;
    I.cs_len = tslots;
    I.listed = false;
    if (t->mass) I.mass.assign(t->mass, t->mass + n);
    *id_out = slot;
    return PDGPU_OK;
}

extern "C" int pdgpu_instance_remove(pdgpu_ctx *c, int32_t id) {
    if (!c || !valid_id(c, id)) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    Instance &I = c->inst[id];
    std::vector<int> neg(I.nblocks, -1);
    up_put(c, (int *)c->blk_inst.p + I.block0, neg.data(), I.nblocks * 4);
    up_zero(c, (int *)c->row_count.p + I.block0, I.nblocks * 4);
    up_zero(c, (int *)c->inst_nat.p + id, 4);
    TRY(up_commit(c));
    TRY(bonded_release(c, I));
    auto tj = c->traj.find(id);
    if (tj != c->traj.end()) { // frames are written at the end of every MD call
        c->used -= (int64_t)tj->second.ring.bytes;
        release(tj->second.ring);
        c->traj.erase(tj);
    }
    c->csr.release(I.cs_off, I.cs_len);
    for (int k = 0; k < I.nblocks; ++k) c->h_blk_inst[I.block0 + k] = -1;
    c->blocks.release(I.block0, I.nblocks);
    c->bandr.release(I.band_off, I.band_len);
    c->spr.release(I.sp_off, I.nsparse);
    c->tri_sum -= (int64_t)I.nblocks * (I.nblocks + 1) / 2;
    I = Instance();
    // re-pack the remaining rows (their bounds are kept) so no tile of the
    // removed instance is reachable
    const DevState S = c->S();
    const int all = (int)c->cap_blocks;
    hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(1024), 0, c->stream, S, all);
    hipLaunchKernelGGL(k_tile_rows<1>, dim3((all + 3) / 4), dim3(256), 0, c->stream, S, c->P, 0, all);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return PDGPU_OK;
}

extern "C" int pdgpu_bonded_set(pdgpu_ctx *c, int32_t id, const pdgpu_bonded *b) {
    if (!c || !valid_id(c, id) || !b) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    Instance &I = c->inst[id];
    const int n = I.natoms;
    const int nb = b->nbonds, na = b->nangles, nt = b->ntorsions, ni = b->nimpropers;
    if (nb < 0 || na < 0 || nt < 0 || ni < 0 || (nb && (!b->bond_ij || !b->bond_lk)) ||
        (na && (!b->angle_iaj || !b->angle_tk)) ||
        (nt && (!b->torsion_iabj || !b->torsion_terms || !b->torsion_vn || !b->torsion_n || !b->torsion_gamma)) ||
        (ni && (!b->improper_iabj || !b->improper_terms || !b->improper_vn || !b->improper_n || !b->improper_gamma)))
        return fail(c, PDGPU_ERR_INVALID, "invalid bonded term lists");
    auto in_range = [n](const int32_t *p, int64_t cnt) {
        for (int64_t k = 0; k < cnt; ++k)
            if (p[k] < 0 || p[k] >= n) return false;
        return true;
    };
    if (!in_range(b->bond_ij, 2 * (int64_t)nb) || !in_range(b->angle_iaj, 3 * (int64_t)na) ||
        !in_range(b->torsion_iabj, 4 * (int64_t)nt) || !in_range(b->improper_iabj, 4 * (int64_t)ni))
        return fail(c, PDGPU_ERR_INVALID, "bonded term refers to an atom outside the instance");
    for (int k = 0; k < nt; ++k)
        if (b->torsion_terms[k] < 0 || b->torsion_terms[k] > 4) return fail(c, PDGPU_ERR_INVALID, "torsion terms");
    for (int k = 0; k < ni; ++k)
        if (b->improper_terms[k] < 0 || b->improper_terms[k] > 4) return fail(c, PDGPU_ERR_INVALID, "improper terms");
    TRY(bonded_release(c, I));
    const int64_t nd = (int64_t)nt + ni;
    {
        DevBuf *bb[] = {&c->bd_idx, &c->bd_par};
        const size_t be[] = {16, 16};
        TRY(term_alloc(c, c->bdr, c->cap_bd, nb, bb, be, 2, &I.bd_off));
        I.bd_len = nb;
        DevBuf *ab[] = {&c->an_idx, &c->an_par};
        TRY(term_alloc(c, c->anr, c->cap_an, na, ab, be, 2, &I.an_off));
        I.an_len = na;
        DevBuf *db[] = {&c->dh_idx, &c->dh_meta, &c->dh_vn, &c->dh_n, &c->dh_gamma};
        const size_t de[] = {16, 4, 32, 32, 32};
        TRY(term_alloc(c, c->dhr, c->cap_dh, nd, db, de, 5, &I.dh_off));
        I.dh_len = nd;
    }
    const int g0 = I.block0 * kBlk, s1 = id + 1;
    const double inv = 1.0 / kJPerKcal; // J/molecule -> kcal/mol
    std::vector<int4> bi(nb), ai(na), di(nd);
    std::vector<double2> bp(nb), ap(na);
    std::vector<int> dm(nd);
    std::vector<double4> dv(nd), dn(nd), dg(nd);
    for (int k = 0; k < nb; ++k) {
        bi[k] = make_int4(g0 + b->bond_ij[2 * k], g0 + b->bond_ij[2 * k + 1], s1, 0);
        bp[k] = make_double2(b->bond_lk[2 * k], b->bond_lk[2 * k + 1] * inv);
    }
    for (int k = 0; k < na; ++k) {
        ai[k] = make_int4(g0 + b->angle_iaj[3 * k], g0 + b->angle_iaj[3 * k + 1], g0 + b->angle_iaj[3 * k + 2], s1);
        ap[k] = make_double2(b->angle_tk[2 * k], b->angle_tk[2 * k + 1] * inv);
    }
    for (int64_t k = 0; k < nd; ++k) {
        const bool im = k >= nt;
        const int64_t r = im ? k - nt : k;
        const int32_t *q = (im ? b->improper_iabj : b->torsion_iabj) + 4 * r;
        const double *v = (im ? b->improper_vn : b->torsion_vn) + 4 * r;
        const double *nn = (im ? b->improper_n : b->torsion_n) + 4 * r;
        const double *g = (im ? b->improper_gamma : b->torsion_gamma) + 4 * r;
        const int terms = (im ? b->improper_terms : b->torsion_terms)[r];
        di[k] = make_int4(g0 + q[0], g0 + q[1], g0 + q[2], g0 + q[3]);
        dm[k] = s1 | (terms << 24) | ((im ? 1 : 0) << 28);
        double vv[4], ww[4], gg[4];
        for (int t = 0; t < 4; ++t) {
            const bool used = t < terms;
            vv[t] = used ? v[t] * inv : 0.0;
            ww[t] = used ? nn[t] : 0.0;
            gg[t] = used ? g[t] : 0.0;
        }
        dv[k] = make_double4(vv[0], vv[1], vv[2], vv[3]);
        dn[k] = make_double4(ww[0], ww[1], ww[2], ww[3]);
        dg[k] = make_double4(gg[0], gg[1], gg[2], gg[3]);
    }
    if (nb) {
        up_put(c, (int4 *)c->bd_idx.p + I.bd_off, bi.data(), nb * 16);
        up_put(c, (double2 *)c->bd_par.p + I.bd_off, bp.data(), nb * 16);
    }
    if (na) {
        up_put(c, (int4 *)c->an_idx.p + I.an_off, ai.data(), na * 16);
        up_put(c, (double2 *)c->an_par.p + I.an_off, ap.data(), na * 16);
    }
    if (nd) {
        up_put(c, (int4 *)c->dh_idx.p + I.dh_off, di.data(), nd * 16);
        up_put(c, (double4 *)c->dh_vn.p + I.dh_off, dv.data(), nd * 32);
        up_put(c, (double4 *)c->dh_n.p + I.dh_off, dn.data(), nd * 32);
        up_put(c, (double4 *)c->dh_gamma.p + I.dh_off, dg.data(), nd * 32);
        up_put(c, (int *)c->dh_meta.p + I.dh_off, dm.data(), nd * 4);
    }
    return up_commit(c);
}

extern "C" int64_t pdgpu_bonded_bytes(int32_t nbonds, int32_t nangles, int32_t ntorsions, int32_t nimpropers) {
    if (nbonds < 0 || nangles < 0 || ntorsions < 0 || nimpropers < 0) return -1;
    return 32 * (int64_t)nbonds + 32 * (int64_t)nangles + 116 * ((int64_t)ntorsions + nimpropers);
}

extern "C" int pdgpu_set_hydrogens(pdgpu_ctx *c, int32_t id, const uint8_t *flags) {
    if (!c || !valid_id(c, id) || !flags) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    const Instance &I = c->inst[id];
    std::vector<uint8_t> h(flags, flags + I.natoms);
    for (auto &v : h) v = v ? 1 : 0;
    up_put(c, (uint8_t *)c->is_h.p + (size_t)I.block0 * kBlk, h.data(), h.size());
    return up_commit(c);
}

extern "C" int pdgpu_set_positions(pdgpu_ctx *c, int32_t id, const double *xyz) {
    if (!c || !valid_id(c, id) || !xyz) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return scatter3(c, id, c->px, c->py, c->pz, xyz, 1.0);
}

extern "C" int pdgpu_get_positions(pdgpu_ctx *c, int32_t id, double *xyz) {
    if (!c || !valid_id(c, id) || !xyz) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return gather3(c, id, c->px, c->py, c->pz, xyz, 1.0);
}

extern "C" int pdgpu_set_velocities(pdgpu_ctx *c, int32_t id, const double *v) {
    if (!c || !valid_id(c, id) || !v) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    c->inst[id].has_vel = true;
    return scatter3(c, id, c->vx, c->vy, c->vz, v, 1.0);
}

extern "C" int pdgpu_get_velocities(pdgpu_ctx *c, int32_t id, double *v) {
    if (!c || !valid_id(c, id) || !v) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return gather3(c, id, c->vx, c->vy, c->vz, v, 1.0);
}

extern "C" int pdgpu_nlist_build(pdgpu_ctx *c, int32_t id) {
    if (!c || (id != PDGPU_ALL && !valid_id(c, id))) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return c->mixed ? launch_build<float>(c, id) : launch_build<double>(c, id);
}

extern "C" int pdgpu_nlist_export(pdgpu_ctx *c, int32_t id, int32_t shadow, int32_t *row_start, uint32_t *entries,
                                  int64_t capacity, int64_t *count) {
    if (!c || !valid_id(c, id) || !row_start || !count || capacity < 0 || (capacity > 0 && !entries))
        return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    TRY(ensure_listed(c, id));
    const Instance &I = c->inst[id];
    const DevState S = c->S();
    int *d_cnt = nullptr;
    long long *d_off = nullptr;
    uint32_t *d_ent = nullptr;
    HIPCHECK(c, hipMalloc(&d_cnt, (size_t)I.natoms * 4 + 4));
    HIPCHECK(c, hipMemsetAsync(d_cnt, 0, (size_t)I.natoms * 4 + 4, c->stream));
    hipLaunchKernelGGL(k_export<0>, dim3((I.nblocks + 3) / 4), dim3(256), 0, c->stream, S, c->P, I.block0, I.nblocks,
                       shadow ? 1 : 0, d_cnt, (const long long *)nullptr, (uint32_t *)nullptr);
    std::vector<int> cnt(I.natoms);
    hipError_t e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)I.natoms * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) { (void)hipFree(d_cnt); return fail(c, PDGPU_ERR_HIP, hipGetErrorString(e)); }
    std::vector<long long> off(I.natoms + 1, 0);
    for (int i = 0; i < I.natoms; ++i) off[i + 1] = off[i] + cnt[i];
    const int64_t total = off[I.natoms];
    *count = total;
    for (int i = 0; i <= I.natoms; ++i) row_start[i] = (int32_t)std::min<long long>(off[i], INT32_MAX);
    if (total > capacity) {
        (void)hipFree(d_cnt);
        return fail(c, PDGPU_ERR_NOMEM, "export capacity too small");
    }
    e = hipMalloc(&d_off, (size_t)(I.natoms + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&d_ent, (size_t)std::max<int64_t>(total, 1) * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_export<1>, dim3((I.nblocks + 3) / 4), dim3(256), 0, c->stream, S, c->P, I.block0,
                           I.nblocks, shadow ? 1 : 0, (int *)nullptr, (const long long *)d_off, d_ent);
        e = hipGetLastError();
    }
    if (e == hipSuccess && total > 0)
        e = hipMemcpyAsync(entries, d_ent, (size_t)total * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_cnt);
    if (d_off) (void)hipFree(d_off);
    if (d_ent) (void)hipFree(d_ent);
    if (e != hipSuccess) return fail(c, PDGPU_ERR_HIP, hipGetErrorString(e));
    return PDGPU_OK;
}

namespace pdgpu {
// Exclusive scan of counts[0, n) into offs (int64) and row32 (int32,
// clamped), total at offs[n]; one workgroup of 1024 threads, deterministic.
__global__ void __launch_bounds__(1024) k_scan_counts(const int *counts, int n, long long *offs, int32_t *row32) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = t * per, b1 = min(n, b0 + per);
    long long sum = 0;
    for (int i = b0; i < b1; ++i) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - sum;
    for (int i = b0; i < b1; ++i) {
        offs[i] = run;
        row32[i] = (int32_t)min(run, (long long)INT32_MAX);
        run += counts[i];
    }
    if (t == 1023) {
        offs[n] = part[1023];
        row32[n] = (int32_t)min(part[1023], (long long)INT32_MAX);
    }
}

// Zero fill (instead of hipMemsetAsync, so a captured graph replays it).
__global__ void k_zero_i32(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
} // namespace pdgpu

extern "C" int pdgpu_nlist_export_noalloc(pdgpu_ctx *c, int32_t id, int32_t shadow, int32_t *row_start,
                                          uint32_t *entries, int64_t capacity, int64_t *count) {
    if (!c || !valid_id(c, id) || !row_start || !count || capacity < 0 || (capacity > 0 && !entries))
        return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    if (!c->inst[id].listed)
        return fail(c, PDGPU_ERR_INVALID, "nlist_export_noalloc: build the list first (pdgpu_nlist_build)");
    TRY(scratch_setup(c->device));
    unsigned char *base = g_scratch[c->device];
    const Instance &I = c->inst[id];
    const size_t n = (size_t)I.natoms;
    // layout: counts int32[n] | row32 int32[n+1] | offs int64[n+1] | entries
    const size_t o_row = (n * 4 + 255) / 256 * 256;
    const size_t o_off = (o_row + (n + 1) * 4 + 255) / 256 * 256;
    const size_t o_ent = (o_off + (n + 1) * 8 + 255) / 256 * 256;
    if (o_ent >= kScratchBytes)
        return fail(c, PDGPU_ERR_NOMEM, "nlist_export_noalloc: instance beyond the 64 MiB scratch array");
    int *d_cnt = (int *)base;
    int32_t *d_row = (int32_t *)(base + o_row);
    long long *d_off = (long long *)(base + o_off);
    uint32_t *d_ent = (uint32_t *)(base + o_ent);
    const DevState S = c->S();
    // the scratch array keeps the last call's contents: counts start at zero
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_cnt, (int)n);
    hipLaunchKernelGGL(k_export<0>, dim3((I.nblocks + 3) / 4), dim3(256), 0, c->stream, S, c->P, I.block0, I.nblocks,
                       shadow ? 1 : 0, d_cnt, (const long long *)nullptr, (uint32_t *)nullptr);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, c->stream, (const int *)d_cnt, (int)n, d_off, d_row);
    long long total = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&total, d_off + n, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(row_start, d_row, (n + 1) * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, PDGPU_ERR_HIP, hipGetErrorString(e));
    *count = total;
    if (total > capacity) return fail(c, PDGPU_ERR_NOMEM, "export capacity too small");
    if ((size_t)total * 4 > kScratchBytes - o_ent)
        return fail(c, PDGPU_ERR_NOMEM, "nlist_export_noalloc: list beyond the 64 MiB scratch array");
    hipLaunchKernelGGL(k_export<1>, dim3((I.nblocks + 3) / 4), dim3(256), 0, c->stream, S, c->P, I.block0, I.nblocks,
                       shadow ? 1 : 0, (int *)nullptr, (const long long *)d_off, d_ent);
    e = hipGetLastError();
    if (e == hipSuccess && total > 0)
        e = hipMemcpyAsync(entries, d_ent, (size_t)total * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, PDGPU_ERR_HIP, hipGetErrorString(e));
    return PDGPU_OK;
}

static int calc_common(pdgpu_ctx *c, int32_t id, pdgpu_energies *en, double *forces, int mode) {
    if (!c || (id != PDGPU_ALL && !valid_id(c, id))) return PDGPU_ERR_INVALID;
    if (id == PDGPU_ALL && forces) return fail(c, PDGPU_ERR_INVALID, "forces output needs a single instance");
    (void)hipSetDevice(c->device);
    TRY(eval(c, id, mode));
    const int rc = read_energies(c, id, en, mode);
    if (forces) TRY(gather3(c, id, c->fx, c->fy, c->fz, forces, kForceToN));
    return rc;
}

extern "C" int pdgpu_calc_forces(pdgpu_ctx *c, int32_t id, pdgpu_energies *en, double *forces) {
    return calc_common(c, id, en, forces, 0);
}

extern "C" int pdgpu_calc_energies(pdgpu_ctx *c, int32_t id, pdgpu_energies *en, double *forces) {
    return calc_common(c, id, en, forces, 1);
}

extern "C" int pdgpu_get_forces(pdgpu_ctx *c, int32_t id, double *forces) {
    if (!c || !valid_id(c, id) || !forces) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return gather3(c, id, c->fx, c->fy, c->fz, forces, kForceToN);
}

extern "C" int pdgpu_get_born_radii(pdgpu_ctx *c, int32_t id, double *out) {
    if (!c || !valid_id(c, id) || !out) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    const Instance &I = c->inst[id];
    const size_t a0 = (size_t)I.block0 * kBlk;
    if (c->mixed) {
        std::vector<float> h(I.natoms);
        HIPCHECK(c, hipMemcpyAsync(h.data(), (float *)c->born.p + a0, I.natoms * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < I.natoms; ++i) out[i] = h[i];
    } else {
        HIPCHECK(c, hipMemcpyAsync(out, (double *)c->born.p + a0, I.natoms * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    return PDGPU_OK;
}

// Grid-shape autotuner: times each tile kernel of a force evaluation of
// `id` (or all instances) at several wavefronts-per-CU settings and keeps
// the fastest per kernel. Results do not depend on the grid (every tile's
// sums are rounded to fixed point on their own), so tuning changes speed
// only.
extern "C" int pdgpu_autotune(pdgpu_ctx *c, int32_t id, int32_t reps, int32_t *chosen) {
    if (!c || (id != PDGPU_ALL && !valid_id(c, id)) || reps <= 0) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    static const int cand[] = {8, 16, 32, 64, 128, 256, 512, 1024};
    const int kinds[kGridKinds] = {PDGPU_KERNEL_ENUM, PDGPU_KERNEL_BORN, PDGPU_KERNEL_PAIR, PDGPU_KERNEL_CHAIN};
    double best[kGridKinds];
    int pick[kGridKinds];
    for (int k = 0; k < kGridKinds; ++k) { best[k] = 1e300; pick[k] = c->grid_wpc[k]; }
    const bool was = c->prof;
    TRY(eval(c, id, 0)); // lists built, code loaded
    for (int w : cand) {
        for (int k = 0; k < kGridKinds; ++k) c->grid_wpc[k] = w;
        TRY(pdgpu_profile(c, 1));
        for (int r = 0; r < reps; ++r) TRY(eval(c, id, 0));
        double ms[PDGPU_KERNEL_NGROUPS];
        int64_t n[PDGPU_KERNEL_NGROUPS];
        TRY(pdgpu_profile_read(c, ms, n, PDGPU_KERNEL_NGROUPS));
        for (int k = 0; k < kGridKinds; ++k)
            if (ms[kinds[k]] < best[k]) { best[k] = ms[kinds[k]]; pick[k] = w; }
    }
    for (int k = 0; k < kGridKinds; ++k) {
        c->grid_wpc[k] = pick[k];
        if (chosen) chosen[k] = pick[k];
    }
    TRY(pdgpu_profile(c, was ? 1 : 0));
    return PDGPU_OK;
}

extern "C" int pdgpu_profile(pdgpu_ctx *c, int32_t enable) {
    if (!c) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (auto &r : c->prof_recs) {
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_recs.clear();
    c->prof = enable != 0;
    return PDGPU_OK;
}

extern "C" int pdgpu_profile_read(pdgpu_ctx *c, double *ms, int64_t *launches, int32_t n) {
    if (!c || !ms || !launches || n <= 0) return PDGPU_ERR_INVALID;
    (void)hipSetDevice(c->device);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k) { ms[k] = 0; launches[k] = 0; }
    for (auto &r : c->prof_recs) {
        float t = 0;
        HIPCHECK(c, hipEventElapsedTime(&t, r.a, r.b));
        if (r.kind < n) { ms[r.kind] += t; launches[r.kind] += 1; }
    }
    return PDGPU_OK;
}

// Exposes the internals the integrators need.
namespace pdgpu_internal {
DevState state(pdgpu_ctx *c) { return c->S(); }
hipStream_t stream(pdgpu_ctx *c) { return c->stream; }
}

#include "pdgpu_io.inc"

#include "pdgpu_integrate.inc"
