"""Benchmark of the MI355X GB / nonbonded hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--sections batch,md,serving]

Headline (`value`): the batched path. 4,096 independent extended
conformers of 10-150 residues (~5.3 M atoms), each with its own tile list
(10 A cutoff, 7 A inner, 2 A skin) and Still GB (eps_out 78.5), all
evaluated together. A step is one energy + force evaluation of every
conformer: pair enumeration, Born radii, fused LJ + Coulomb + GB pair
pass, chain rule. `value` is atom-pair evaluations per second (unordered
pairs within the cutoff, counted on the device) in mixed precision; the
FP64 path runs the same steps beside it.

Sections added to the same JSON line:
  md       one 500-residue chain (~8,200 atoms, phi/psi from the alpha /
           beta basins, bonded terms + GB), 4,000 steepest-descent steps of
           clash relief, 1,000 steps of velocity-rescaling equilibration to
           300 K, then velocity-Verlet MD at 1 fs, 12 A
           cutoff, 2 A skin, list rebuilt every 10 steps, in mixed
           precision; every 100 steps the FP64 path evaluates the mixed
           positions and the mixed energy and forces at the same positions
           are compared with it; FP64 MD timed beside it for the ratio.
  serving  minimisation requests with Poisson arrivals at several rates,
           sizes log-uniform in 20-400 residues (extended build, bonded +
           GB, 10 A cutoff); steepest descent until the force RMS is below
           0.1 kcal/mol/A or 1,000 steps; continuous batching with
           admission by device-memory budget; p50/p99 latency and
           minimisations/s.
  allpairs one 2,000-residue chain (~32,000 atoms), all-pairs Coulomb +
           Still GB, LJ at 12 A, 1,000 mixed velocity-Verlet steps with the
           same FP64 check every 100 steps.

The workloads (molecules, options) are pd.arcus_amd/pdgpu/workloads.py,
the same inputs the configuration tests run. Inputs are synthetic
(pd.arcus_amd/pdgpu/synth.py) and resident in HBM when the timed region
starts. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pd.arcus_amd"))

KB = 1.38065e-23
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6    # MI355X published FP64 vector peak
# algorithmic operation counts per pair (add/mul/cmp = 1, FMA = 2,
# transcendental = 1), counted from the reference arithmetic
# (forcefields/gbff.cpp:1549-1719 pair pass, :432-465 Born sum, :1751-1808 chain rule)
FLOPS_PAIR, FLOPS_BORN, FLOPS_CHAIN = 86, 20, 30
# the compiled reference's pair rate over the oracle port's on the same
# 3,362 PD-built batch conformers, one and eight processes alike (measured
# by oracle/ref_cpu_baseline.py in the build host, where the reference tree
# exists; it never reaches a GPU host, so the port is timed there)
REF_OVER_PORT = 1.63


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_md(precision, mol, xyz0, vel0, steps, warmup, profile, pdgpu):
    """Runs warmup + steps MD steps; returns the timed steps' wall time, the
    energies before and after them and a profile of a further short run."""
    from pdgpu import workloads as W
    ctx = pdgpu.Context(W.options("chain", precision))
    i = ctx.add_system(mol, xyz=xyz0, bonded=True)
    ctx.set_velocities(i, vel0)
    e0, ek0 = ctx.md_vv(i, max(1, warmup), 1e-15, 10)
    t0 = time.perf_counter()
    e, ek = ctx.md_vv(i, steps, 1e-15, 10)
    dt = time.perf_counter() - t0
    prof = None
    if profile:
        ctx.profile(True)
        ctx.md_vv(i, min(steps, 1000), 1e-15, 10)
        prof = ctx.profile_read()
        ctx.profile(False)
    ctx.close()
    return dt, (e0, ek0), (e, ek), prof


def energy_contract(e_mixed, e_ref, pairs):
    """The mixed path's per-instance contract against FP64 (tests/gpu_helpers.py):
    |dE| <= max(2e-5 |E|, 2e-5 kcal/mol x sqrt(pairs))."""
    import pdgpu
    d = abs(e_mixed - e_ref)
    bound = max(2e-5 * abs(e_ref), 2e-5 * np.sqrt(max(pairs, 1)) * pdgpu.J_PER_KCALMOL)
    return d, bound


def _oracle_pairs(O, s, xyz, rs, ent, cutoff):
    """Unordered pairs within the cutoff in an oracle neighbour list."""
    j = (ent & 0xFFFFFF).astype(np.int64)
    i = np.repeat(np.arange(len(rs) - 1), np.diff(rs))
    keep = j > i
    d2 = ((xyz[i[keep]] - xyz[j[keep]]) ** 2).sum(1)
    return int((d2 <= cutoff * cutoff).sum())


def _cpu_worker(job):
    """Oracle evaluations of conformers job[0] (list of (mol, ...)) for
    job[1] seconds: list build + Born radii + pair pass + chain rule each."""
    mols, budget = job
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    p = O.default_params(cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0, vdw_inner_cutoff=7.0, vdw14_scaling=0.5,
                         elec14_scaling=0.83333, dielectric_solvent=78.5, padding=2.0)
    done, npairs, busy = 0, 0, 0.0
    t_end = time.perf_counter() + budget
    for m in mols:
        s = O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                      m["mid"], m["band"], m["sp_i"], m["sp_j"], m["sp_o"])
        t0 = time.perf_counter()
        rs, ent = O.nlist_build(s, m["xyz"], 10.0, 2.0, True)
        O.gb_calc_forces(s, m["xyz"], rs, ent, p)
        busy += time.perf_counter() - t0
        done += 1
        npairs += _oracle_pairs(O, s, m["xyz"], rs, ent, 10.0)
        if time.perf_counter() > t_end:
            break
    return done, npairs, busy


def cpu_baseline_batch(mols, single_s=10.0, multi_s=6.0):
    """The oracle (scalar C port of the reference path) on a bounded sample of
    the batch, on the host: one thread for ~single_s seconds, then one
    process per core (up to 16, the host share of one GPU) for ~multi_s
    seconds of wall time. Runs before any GPU context exists (forks)."""
    import multiprocessing as mp
    d1, p1, b1 = _cpu_worker((mols, single_s))
    cores = min(16, os.cpu_count() or 1)
    chunks = [mols[k::cores] for k in range(cores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(c, multi_s) for c in chunks])
    wall = time.perf_counter() - t0
    dn, pn = sum(r[0] for r in res), sum(r[1] for r in res)
    return {"value": pn / wall, "unit": "atom-pair evals/s", "cores": cores, "kind": "port",
            "single_thread_value": p1 / b1,
            "reference_estimate": REF_OVER_PORT * pn / wall,
            "single_thread_reference_estimate": REF_OVER_PORT * p1 / b1,
            "tie": f"reference = {REF_OVER_PORT} x port, the ratio of the compiled reference (its own -O3 "
                   f"-ffast-math flags) to this -O2 port on the same 3,362 PD-built batch conformers on one "
                   f"build host, 1 and 8 processes",
            "sample": f"oracle C port (tests/oracle_lib.py), one evaluation (list build + Born + pair pass + "
                      f"chain rule) per conformer of the batch: {d1} conformers on 1 thread in {b1:.1f} s; "
                      f"{dn} conformers on {cores} processes in {wall:.1f} s wall"}


KERNEL_NAMES = {"pair": "pdgpu::k_pair<{R}, 0>", "pair_mixed": "pdgpu::k_pair_mixed<{S}, {A}>",
                "born": "pdgpu::k_born<{R}>", "chain": "pdgpu::k_chain<{R}>",
                "enum": "pdgpu::k_filter<{R}>"}


def roofline(prof, pairs, born_pairs, precision, traffic=None, same_switch=True, all_in=True):
    """Roofline of the dominant kernel. Algorithmic work per launch: 86 FLOP
    per unordered pair within the cutoff (pair pass), 20 and 30 FLOP per
    ordered Still pair (Born sum, chain rule). The enumeration kernel has no
    algorithmic FLOPs of its own (it selects the pairs), so the FLOP roofline
    is taken on the dominant of the three pair passes; the enumeration's
    share of the step is reported beside it. `step` applies the same count
    to the whole evaluation (all kernels)."""
    flops = {"pair": FLOPS_PAIR * pairs, "born": FLOPS_BORN * born_pairs, "chain": FLOPS_CHAIN * born_pairs}
    name = max(flops, key=lambda k: prof[k][0])
    ms, launches = prof[name]
    avg_s = ms / 1e3 / max(1, launches)
    achieved = flops[name] / avg_s / 1e12
    peak = PEAK_FP32_TFLOPS if precision == "mixed" else PEAK_FP64_TFLOPS
    key = "pair_mixed" if (name == "pair" and precision == "mixed") else name
    kname = KERNEL_NAMES[key].format(R="float" if precision == "mixed" else "double",
                                     S="true" if same_switch else "false", A="true" if all_in else "false")
    total_ms = sum(v[0] for k, v in prof.items() if k in ("atom", "born", "pair", "chain", "enum", "bonded"))
    evals = max(1, prof["pair"][1])
    step_s = total_ms / 1e3 / evals
    step_flops = sum(flops.values())
    return {"kernel": kname, "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": (traffic or {}).get(kname), "avg_launch_us": avg_s * 1e6,
            "flops_per_launch": flops[name],
            "step": {"flops": step_flops, "kernel_time_us": step_s * 1e6, "achieved": step_flops / step_s / 1e12,
                     "frac": step_flops / step_s / 1e12 / peak},
            "enum_share_of_step": prof["enum"][0] / max(total_ms, 1e-12),
            "groups_ms": {k: round(v[0], 4) for k, v in prof.items()},
            "groups_launches": {k: v[1] for k, v in prof.items()}}


def load_traffic(name="kernels_batch4096_traffic.json"):
    """HBM bytes per dispatch by kernel, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def build_md():
    from pdgpu import workloads as W
    log("md: building the 500-residue chain")
    return W.chain_molecule()


def run_md(args, pdgpu, mol):
    """Mixed-precision velocity Verlet on the relaxed 500-residue chain with
    the FP64 path evaluated at the mixed positions every 100 steps (the two
    precisions' own trajectories part chaotically within a few hundred
    steps, so their end energies are not comparable); FP64 MD timed over
    1,000 steps for the ratio."""
    from pdgpu import workloads as W
    n = mol["n"]
    t0 = time.perf_counter()
    xyz = W.relax(lambda: pdgpu.Context(W.options("chain", pdgpu.MIXED)), mol)
    log(f"md: relaxed by 4000 SD steps in {time.perf_counter() - t0:.1f} s")
    vel = W.maxwell(mol["mass"], 300.0, seed=7)
    steps, every = args.md_steps, 100
    cm = pdgpu.Context(W.options("chain", pdgpu.MIXED))
    c64 = pdgpu.Context(W.options("chain", pdgpu.FP64))
    i = cm.add_system(mol, xyz=xyz, bonded=True)
    j = c64.add_system(mol, xyz=xyz, bonded=True)
    cm.set_velocities(i, vel)
    W.equilibrate(cm, i, mol["mass"], 300.0)  # 1,000 steps with rescaling to 300 K
    e0, ek0 = cm.md_vv(i, 10, 1e-15, 10)  # warm-up
    checks, t_md = [], 0.0
    for k in range(max(1, steps // every)):
        t0 = time.perf_counter()
        e, ek = cm.md_vv(i, every, 1e-15, 10)
        t_md += time.perf_counter() - t0
        x = cm.get_positions(i)
        em, fm = cm.calc_forces(i)
        c64.set_positions(j, x)
        c64.nlist_build(j)  # the list of the mixed positions, as the mixed run keeps it
        ed, fd = c64.calc_forces(j)
        d, bound = energy_contract(em["epot"], ed["epot"], ed["pairs"])
        checks.append({"step": (k + 1) * every, "epot_fp64_kcal": ed["epot"] / pdgpu.J_PER_KCALMOL,
                       "energy_rel_err": d / abs(ed["epot"]), "energy_err_over_bound": d / bound,
                       "force_rms_rel_err": float(np.sqrt(np.mean((fm - fd) ** 2)) / np.sqrt(np.mean(fd ** 2))),
                       "temperature_K": 2.0 * ek / (3.0 * n * KB)})
    done = len(checks) * every
    pairs = em["pairs"]
    cm.profile(True)
    cm.md_vv(i, 1000, 1e-15, 10)
    prof = cm.profile_read()
    cm.close()
    c64.set_velocities(j, vel)
    c64.md_vv(j, 10, 1e-15, 10)
    t0 = time.perf_counter()
    c64.md_vv(j, 1000, 1e-15, 10)
    t64 = (time.perf_counter() - t0) / 1000
    c64.profile(True)
    e64, _ = c64.md_vv(j, 1000, 1e-15, 10)
    prof64 = c64.profile_read()
    c64.close()
    temps = [c["temperature_K"] for c in checks]
    log(f"md mixed: {done} steps in {t_md:.3f} s; T {min(temps):.0f}-{max(temps):.0f} K; "
        f"worst energy error / bound {max(c['energy_err_over_bound'] for c in checks):.2f}")
    ms = t_md / done * 1e3
    return {"workload": f"single 500-residue chain ({n} atoms), 4000 SD steps of clash relief, 1000 steps of "
                        "velocity-rescaling equilibration to 300 K, then NVE velocity-Verlet MD, dt 1 fs, 12 A "
                        "cutoff + Still GB + bonded terms, mixed precision, two wavefronts per tile; FP64 "
                        "evaluated at the mixed positions every 100 steps",
            "steps": done, "atoms": n, "unit": "atom-pair evals/s",
            "mixed": {"value": pairs / (ms * 1e-3), "ms_per_step": ms, "pairs_per_step": pairs,
                      "roofline": roofline(prof, em["pairs"], em["born_pairs"], "mixed")},
            "fp64": {"value": e64["pairs"] / t64, "ms_per_step": t64 * 1e3,
                     "roofline": roofline(prof64, e64["pairs"], e64["born_pairs"], "fp64")},
            "mixed_over_fp64": t64 * 1e3 / ms,
            "epot_start_kcal": e0["epot"] / pdgpu.J_PER_KCALMOL,
            "temperature_K": {"min": min(temps), "max": max(temps), "end": temps[-1]},
            "max_energy_rel_err": max(c["energy_rel_err"] for c in checks),
            "max_energy_err_over_contract": max(c["energy_err_over_bound"] for c in checks),
            "max_force_rms_rel_err": max(c["force_rms_rel_err"] for c in checks),
            "checks": checks[:: max(1, len(checks) // 10)]}


def build_allpairs(args):
    from pdgpu import workloads as W
    log(f"allpairs: building a {args.allpairs_residues}-residue chain")
    return W.allpairs_molecule(args.allpairs_residues)


def run_allpairs(args, pdgpu, mol):
    """Mixed-precision velocity Verlet with all-pairs GB; every `every`
    steps the mixed trajectory's positions are evaluated by the FP64 path
    (the path held to the oracle at 1e-10) and the mixed energy and forces
    at the same positions are compared with it."""
    from pdgpu import workloads as W
    n = mol["n"]
    steps, every = args.allpairs_steps, 100
    # clash relief as for the single-chain MD (12 A cutoff, bonded + GB); the
    # extended build leaves a few side-chain contacts near 0.9 A
    xyz = W.relax(lambda: pdgpu.Context(W.options("chain", pdgpu.MIXED)), mol)
    vel = W.maxwell(mol["mass"], 300.0, seed=9)
    cm = pdgpu.Context(W.options("allpairs", pdgpu.MIXED))
    c64 = pdgpu.Context(W.options("allpairs", pdgpu.FP64))
    i = cm.add_system(mol, xyz=xyz, bonded=True)
    j = c64.add_system(mol, xyz=xyz, bonded=True)
    cm.set_velocities(i, vel)
    grids = {"mixed": cm.autotune(), "fp64": c64.autotune()}
    log(f"allpairs: autotuned wavefronts per CU {grids}")
    cm.md_vv(i, 2, 1e-15, 0)  # warm-up (graph capture); the list holds every pair, no rebuilds
    checks, t_md = [], 0.0
    for k in range(max(1, steps // every)):
        t0 = time.perf_counter()
        e, ek = cm.md_vv(i, every, 1e-15, 0)
        t_md += time.perf_counter() - t0
        x = cm.get_positions(i)
        em, fm = cm.calc_forces(i)
        c64.set_positions(j, x)
        c64.nlist_build(j)  # the list of the mixed positions, as the mixed run keeps it
        ed, fd = c64.calc_forces(j)
        d, bound = energy_contract(em["epot"], ed["epot"], ed["pairs"])
        checks.append({"step": (k + 1) * every, "epot_fp64_kcal": ed["epot"] / pdgpu.J_PER_KCALMOL,
                       "energy_rel_err": d / abs(ed["epot"]), "energy_err_over_bound": d / bound,
                       "force_rms_rel_err": float(np.sqrt(np.mean((fm - fd) ** 2)) / np.sqrt(np.mean(fd ** 2))),
                       "temperature_K": 2.0 * ek / (3.0 * n * KB)})
        log(f"allpairs step {(k + 1) * every}: epot {checks[-1]['epot_fp64_kcal']:.1f} kcal/mol, "
            f"mixed rel err {checks[-1]['energy_rel_err']:.2e}, T {checks[-1]['temperature_K']:.0f} K")
    done = len(checks) * every
    pairs = em["pairs"]
    # FP64 speed on the same system, for the mixed/FP64 ratio
    c64.set_velocities(j, vel)
    c64.md_vv(j, 2, 1e-15, 0)
    t0 = time.perf_counter()
    c64.md_vv(j, every, 1e-15, 0)
    t64 = (time.perf_counter() - t0) / every
    cm.profile(True)
    cm.md_vv(i, 20, 1e-15, 0)
    prof = cm.profile_read()
    cm.close()
    c64.close()
    v = pairs * done / t_md
    return {"workload": f"single {args.allpairs_residues}-residue chain ({n} atoms), velocity-Verlet MD, "
                        "dt 1 fs, 300 K, all-pairs Still GB (Born radii, GB pair term and Coulomb over every "
                        "pair), LJ 12 A, bonded terms, mixed precision; FP64 check every 100 steps",
            "unit": "atom-pair evals/s", "atoms": n, "steps": done, "pairs_per_step": pairs,
            "value": v, "ms_per_step": t_md / done * 1e3, "fp64_ms_per_step": t64 * 1e3,
            "mixed_over_fp64": t64 / (t_md / done),
            "max_energy_rel_err": max(c["energy_rel_err"] for c in checks),
            "max_energy_err_over_contract": max(c["energy_err_over_bound"] for c in checks),
            "max_force_rms_rel_err": max(c["force_rms_rel_err"] for c in checks),
            "autotuned_waves_per_cu": grids,
            "checks": checks, "roofline": roofline(prof, em["pairs"], em["born_pairs"], "mixed", same_switch=False)}


def build_serving(args):
    from pdgpu import workloads as W
    nreq = args.serve_requests
    log(f"serving: building {nreq} requests (20-400 residues, bonded terms)")
    t0 = time.perf_counter()
    mols = W.serving_molecules(nreq, workers=args.build_workers)
    log(f"serving: built in {time.perf_counter() - t0:.1f} s")
    return mols


def run_serving(args, pdgpu, mols):
    """Continuous-batching minimisation server at several Poisson rates."""
    from pdgpu import serve
    nreq = len(mols)
    from pdgpu import workloads as W
    opts = W.options("serving", pdgpu.MIXED, max_instances=nreq)
    out = {"workload": f"{nreq} minimisation requests per rate, sizes log-uniform 20-400 residues "
                       f"(mean {int(np.mean([m['n'] for m in mols]))} atoms), steepest descent until force "
                       "RMS < 0.1 kcal/mol/A or 1000 steps, 10 A cutoff + Still GB + bonded, mixed precision",
           "memory_budget_bytes": args.serve_budget, "steps_per_tick": 25,
           "list_update": "rebuild per instance when an atom moved > padding/2", "rates": []}
    for rate in args.serve_rates:
        arrivals = serve.poisson_arrivals(rate, nreq, seed=int(rate))
        reqs = [serve.Request(k, m, float(t)) for k, (m, t) in enumerate(zip(mols, arrivals))]
        srv = serve.MinimisationServer(opts, memory_budget=args.serve_budget, max_steps=1000, grms_tol=0.1,
                                       steps_per_tick=25, keep_xyz=False)
        t0 = time.perf_counter()
        done = srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.close()
        summ = serve.summarize(done, wall)
        summ["ticks"] = srv.ticks
        summ["rate_req_per_s"] = rate
        summ["wall_s"] = wall
        out["rates"].append(summ)
        log(f"serving {rate}/s: {summ['minimisations_per_s']:.1f} min/s, p50 {summ['p50_latency_s']:.3f} s, "
            f"p99 {summ['p99_latency_s']:.3f} s")
    return out


def build_batch(args):
    from pdgpu import workloads as W
    log(f"building {args.conformers} conformers")
    t0 = time.perf_counter()
    mols = W.batch_molecules(args.conformers, workers=args.build_workers)
    log(f"built in {time.perf_counter() - t0:.1f} s")
    return mols


def run_batch(args, pdgpu, mols):
    from pdgpu import workloads as W
    opt = lambda prec: W.options("batch", prec, max_instances=args.conformers)
    natoms = sum(m["n"] for m in mols)
    res, grids = {}, {}
    for name, prec in (("mixed", pdgpu.MIXED), ("fp64", pdgpu.FP64)):
        ctx = pdgpu.Context(opt(prec))
        t0 = time.perf_counter()
        for m in mols:
            ctx.add_system(m)
        ctx.nlist_build()
        log(f"{name}: {len(mols)} instances resident in {time.perf_counter() - t0:.1f} s")
        grid = ctx.autotune(pdgpu.PDGPU_ALL, 2)
        grids[name] = grid
        log(f"{name}: autotuned wavefronts per CU {grid}")
        for _ in range(args.warmup):
            ctx.calc_forces(pdgpu.PDGPU_ALL, raw=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.calc_forces(pdgpu.PDGPU_ALL, raw=True)
        dt = time.perf_counter() - t0
        ctx.profile(True)
        for _ in range(args.steps):
            ctx.calc_forces(pdgpu.PDGPU_ALL, raw=True)
        prof = ctx.profile_read()
        r = ctx.calc_forces(pdgpu.PDGPU_ALL)
        ctx.close()
        pairs = sum(e["pairs"] for e in r.values())
        born = sum(e["born_pairs"] for e in r.values())
        res[name] = (dt, pairs, born, prof, r)
        log(f"{name}: {args.steps} evaluations in {dt:.3f} s, {pairs} pairs")
    dt, pairs, born, prof, r = res["mixed"]
    dt64, pairs64, born64, prof64, r64 = res["fp64"]
    errs = [abs(r[k]["epot"] - r64[k]["epot"]) / abs(r64[k]["epot"]) for k in r]
    over = [energy_contract(r[k]["epot"], r64[k]["epot"], r64[k]["pairs"]) for k in r]
    abs_sqrt = [abs(r[k]["epot"] - r64[k]["epot"]) / pdgpu.J_PER_KCALMOL / np.sqrt(max(1, r64[k]["pairs"]))
                for k in r]
    worst = max(r, key=lambda k: abs(r[k]["epot"] - r64[k]["epot"]) / abs(r64[k]["epot"]))
    tot_err = abs(sum(r[k]["epot"] for k in r) - sum(r64[k]["epot"] for k in r)) / abs(sum(r64[k]["epot"] for k in r))
    cpu = args.cpu_baseline
    traffic = load_traffic()
    rl = roofline(prof, pairs, born, "mixed", traffic)
    return {
        "metric": "atom-pair evals/s", "value": pairs * args.steps / dt, "unit": "atom-pair evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 pair maths, int64 fixed-point accumulation",
        "data": "synthetic" if args.ff is None else f"synthetic sequences built by NewProtein from {args.ff}",
        "config": {"workload": f"{args.conformers} extended conformers, 10-150 residues, one batched energy+force "
                               "evaluation per step, 10 A cutoff + Still GB, mixed precision",
                   "atoms": natoms, "pairs_per_step": pairs, "parallelism": "single GPU",
                   "autotuned_waves_per_cu": grids},
        "fp64": {"value": pairs64 * args.steps / dt64, "ms_per_step": dt64 / args.steps * 1e3,
                 "roofline": roofline(prof64, pairs64, born64, "fp64", traffic)},
        "mixed_over_fp64": (pairs * args.steps / dt) / (pairs64 * args.steps / dt64),
        "accuracy": {
            "contract": "per conformer |E_mixed - E_fp64| <= max(2e-5 |E_fp64|, 2e-5 kcal/mol x sqrt(pairs))",
            "conformers_within_contract": sum(1 for d, b in over if d <= b),
            "conformers": len(over),
            "max_err_over_contract": max(d / b for d, b in over),
            "fraction_within_1e-5_relative": float(np.mean([e <= 1e-5 for e in errs])),
            "max_energy_rel_err": max(errs),
            "median_energy_rel_err": float(np.median(errs)),
            "max_abs_err_kcal_per_sqrt_pair": max(abs_sqrt),
            "batch_total_energy_rel_err": tot_err},
        "worst_instance": {"epot_fp64_kcal": r64[worst]["epot"] / pdgpu.J_PER_KCALMOL,
                           "epot_mixed_kcal": r[worst]["epot"] / pdgpu.J_PER_KCALMOL,
                           "components_fp64_kcal": {k: r64[worst][k] / pdgpu.J_PER_KCALMOL for k in
                                                    ("epot_vdw", "epot_elec", "epot_pol_self", "epot_pol_cross")},
                           "components_mixed_kcal": {k: r[worst][k] / pdgpu.J_PER_KCALMOL for k in
                                                     ("epot_vdw", "epot_elec", "epot_pol_self", "epot_pol_cross")}},
        "roofline": rl,
        **({"cpu_baseline": cpu} if cpu else {}),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched evaluations")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--conformers", type=int, default=4096)
    ap.add_argument("--sections", default="batch,md,serving,allpairs")
    ap.add_argument("--allpairs-residues", type=int, default=2000)
    ap.add_argument("--allpairs-steps", type=int, default=1000)
    ap.add_argument("--md-steps", type=int, default=10000)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", type=lambda v: [float(x) for x in v.split(",")],
                    default=[50, 100, 200, 500, 1000, 2000])
    ap.add_argument("--serve-budget", type=int, default=4 << 30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--build-workers", type=int, default=0,
                    help="processes building the synthetic inputs (0: up to 16; 1 under a profiler)")
    ap.add_argument("--ff", default=None,
                    help="a PD parameter file: build every workload with PD's NewProtein from it "
                         "(the same sequences) instead of the synthetic builder")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances shard by molecule across GPUs without exchange)", file=sys.stderr)
        sys.exit(2)
    sections = args.sections.split(",")
    import pdgpu
    from pdgpu import workloads as W
    W.use_parameter_file(args.ff)
    # every input is built before the first GPU context: the builders fork
    batch_mols = build_batch(args) if "batch" in sections else None
    md_mol = build_md() if "md" in sections else None
    serve_mols = build_serving(args) if "serving" in sections else None
    ap_mol = build_allpairs(args) if "allpairs" in sections else None
    args.cpu_baseline = None
    if not args.no_cpu_baseline and batch_mols is not None:
        log("cpu baseline (oracle port)")
        args.cpu_baseline = cpu_baseline_batch(batch_mols)
    # the headline line needs the batch section; without it (profiling a
    # single section) only that section's record is printed
    out = run_batch(args, pdgpu, batch_mols) if batch_mols is not None else {}
    if md_mol is not None:
        out["md"] = run_md(args, pdgpu, md_mol)
    if serve_mols is not None:
        out["serving"] = run_serving(args, pdgpu, serve_mols)
    if ap_mol is not None:
        out["allpairs"] = run_allpairs(args, pdgpu, ap_mol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
