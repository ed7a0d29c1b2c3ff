"""Benchmark of the MI355X GB / nonbonded hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload md|batch]

Default workload (md): one 500-residue peptide-like chain (~8,000 atoms),
backbone phi/psi drawn 60/40 from the alpha / beta basins, clashes relieved
by 100 steepest-descent steps, then velocity-Verlet MD at dt = 1 fs from
300 K Maxwell velocities: 12 A cutoff (9 A inner), 2 A list skin with a
rebuild every 10 steps, Still GB with eps_out 78.5. A step is one MD step:
half kick + drift, Born radii, fused LJ + Coulomb + GB pair pass, chain
rule, half kick. The mixed-precision path is timed for `value`; the FP64
path runs the same steps for comparison.

batch: 4,096 independent extended conformers of 10-150 residues
(~2.6 M atoms, 10 A cutoff), one energy + force evaluation of all of them
per step.

Prints one JSON line. Inputs are synthetic (pd.arcus_amd/pdgpu/synth.py).
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def md_options(precision, pdgpu):
    return pdgpu.default_options(precision=precision, cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0,
                                 vdw_inner_cutoff=9.0, vdw14_scaling=0.5, elec14_scaling=0.83333,
                                 dielectric_solvent=78.5, nlist_padding=2.0)


def maxwell(mass, T, seed):
    rng = np.random.default_rng(seed)
    sig = np.sqrt(KB * T / mass)[:, None]
    v = rng.normal(size=(len(mass), 3)) * sig
    v -= (mass[:, None] * v).sum(0) / mass.sum()  # no net momentum
    return v


def gpu_md(precision, mol, xyz0, vel0, steps, warmup, profile, pdgpu):
    ctx = pdgpu.Context(md_options(precision, pdgpu))
    i = ctx.add_system(mol, xyz=xyz0)
    ctx.set_velocities(i, vel0)
    ctx.md_vv(i, warmup, 1e-15, 10)
    t0 = time.perf_counter()
    e, ek = ctx.md_vv(i, steps, 1e-15, 10)
    dt = time.perf_counter() - t0
    prof = None
    if profile:
        ctx.profile(True)
        e2, _ = ctx.md_vv(i, steps, 1e-15, 10)
        prof = ctx.profile_read()
        ctx.profile(False)
    ctx.close()
    return dt, e, ek, prof


def cpu_baseline_md(mol, xyz, params_kw, pairs_per_eval, budget_s=12.0):
    """The oracle (tests/oracle_lib.py, a scalar C port) timed on the host on a
    bounded sample: repeated single-point evaluations (list build + Born radii
    + pair pass + chain rule) of the same structure for ~budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    p = O.default_params(**params_kw)
    s = O.OSystem(mol["charge"], mol["radius"], mol["epsilon"], mol["gb_radius"], mol["vj"], mol["terms123"],
                  mol["mass"], mol["mid"], mol["band"], mol["sp_i"], mol["sp_j"], mol["sp_o"])
    n = 0
    t0 = time.perf_counter()
    while True:
        rs, ent = O.nlist_build(s, xyz, max(p.cutoff, p.vdw_cutoff), p.padding, True)
        O.gb_calc_forces(s, xyz, rs, ent, p)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": pairs_per_eval * n / dt, "unit": "atom-pair evals/s", "cores": 1, "kind": "port",
            "sample": f"{n} single-point evaluations (list build + Born + pair pass + chain rule) of the "
                      f"{len(mol['charge'])}-atom md structure, {dt:.1f} s, oracle C port, 1 thread"}


def cpu_baseline_batch(mols, pairs, budget_s=12.0):
    """The oracle (scalar C port, 1 thread) on a bounded sample of the batch:
    conformers in order, each a list build + Born + pair pass + chain rule,
    until ~budget_s seconds have passed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    p = O.default_params(cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0, vdw_inner_cutoff=7.0, vdw14_scaling=0.5,
                         elec14_scaling=0.83333, dielectric_solvent=78.5, padding=2.0)
    done = 0
    npairs = 0
    t0 = time.perf_counter()
    for m, pc in zip(mols, pairs):
        s = O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                      m["mid"], m["band"], m["sp_i"], m["sp_j"], m["sp_o"])
        rs, ent = O.nlist_build(s, m["xyz"], 10.0, 2.0, True)
        O.gb_calc_forces(s, m["xyz"], rs, ent, p)
        done += 1
        npairs += pc
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": npairs / dt, "unit": "atom-pair evals/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of {len(mols)} conformers, one evaluation each (list build + Born + pair "
                      f"pass + chain rule), {dt:.1f} s, oracle C port, 1 thread"}


def roofline(prof, pairs, born_pairs, precision):
    """Dominant kernel group and its achieved algorithmic FLOP rate."""
    flops = {"pair": FLOPS_PAIR * pairs, "born": FLOPS_BORN * born_pairs, "chain": FLOPS_CHAIN * born_pairs,
             "enum": 0}
    name = max(("pair", "born", "chain", "enum"), key=lambda k: prof[k][0])
    ms, launches = prof[name]
    avg_s = ms / 1e3 / max(1, launches)
    achieved = flops[name] / avg_s / 1e12
    peak = PEAK_FP32_TFLOPS if precision == "mixed" else PEAK_FP64_TFLOPS
    return {"kernel": f"k_{name}", "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None, "avg_launch_us": avg_s * 1e6,
            "flops_per_launch": flops[name],
            "groups_ms": {k: round(v[0], 4) for k, v in prof.items()},
            "groups_launches": {k: v[1] for k, v in prof.items()}}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(kernel)


def run_md(args, pdgpu):
    from pdgpu import synth
    log("building 500-residue chain")
    mol = synth.peptide(seed=2, residues=500, mode="basins")
    n = mol["n"]
    # clash relief: 100 steepest-descent steps (PD semantics) in mixed precision
    ctx = pdgpu.Context(md_options(pdgpu.MIXED, pdgpu))
    i = ctx.add_system(mol)
    ctx.minimise_sd(i, 100, 2e5, 1)
    xyz = ctx.get_positions(i)
    e0, f0 = ctx.calc_forces(i)
    ctx.close()
    vel = maxwell(mol["mass"], 300.0, seed=7)
    log(f"relaxed: {n} atoms, {e0['pairs']} pairs in cutoff, epot {e0['epot'] / pdgpu.J_PER_KCALMOL:.1f} kcal/mol")

    t_mix, e_mix, ek_mix, prof = gpu_md(pdgpu.MIXED, mol, xyz, vel, args.steps, args.warmup, True, pdgpu)
    log(f"mixed: {args.steps} steps in {t_mix:.3f} s")
    t64, e64, ek64, prof64 = gpu_md(pdgpu.FP64, mol, xyz, vel, args.steps, args.warmup, True, pdgpu)
    log(f"fp64:  {args.steps} steps in {t64:.3f} s")
    pairs = e_mix["pairs"]
    value = pairs * args.steps / t_mix
    value64 = e64["pairs"] * args.steps / t64
    rl = roofline(prof, pairs, e_mix["born_pairs"], "mixed")
    rl["traffic"] = load_traffic(rl["kernel"], "md")
    rl64 = roofline(prof64, e64["pairs"], e64["born_pairs"], "fp64")
    out = {
        "metric": "atom-pair evals/s", "value": value, "unit": "atom-pair evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_mix / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 pair maths, int64 "
        "fixed-point accumulation, fp64 reductions", "data": "synthetic",
        "config": {"workload": "single 500-residue chain, velocity-Verlet MD (dt 1 fs, 300 K), 12 A cutoff + "
                               "Still GB, mixed precision", "atoms": n, "pairs_per_step": pairs,
                   "born_pairs_per_step": e_mix["born_pairs"], "parallelism": "single GPU"},
        "fp64": {"value": value64, "ms_per_step": t64 / args.steps * 1e3, "roofline": rl64},
        "mixed_over_fp64": value / value64,
        "energy_rel_diff_mixed_vs_fp64_after_md": abs(e_mix["epot"] - e64["epot"]) / abs(e64["epot"]),
        "roofline": rl,
    }
    if not args.no_cpu_baseline:
        kw = dict(cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0, vdw_inner_cutoff=9.0, vdw14_scaling=0.5,
                  elec14_scaling=0.83333, dielectric_solvent=78.5, padding=2.0)
        log("cpu baseline (oracle port)")
        out["cpu_baseline"] = cpu_baseline_md(mol, xyz, kw, e0["pairs"])
    return out


def run_batch(args, pdgpu):
    from pdgpu import synth
    rng = np.random.default_rng(11)
    lengths = rng.integers(10, 151, size=args.conformers)
    opt = lambda prec: pdgpu.default_options(precision=prec, cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0,
                                             vdw_inner_cutoff=7.0, vdw14_scaling=0.5, elec14_scaling=0.83333,
                                             dielectric_solvent=78.5, nlist_padding=2.0,
                                             max_instances=args.conformers)
    log(f"building {args.conformers} conformers")
    t0 = time.perf_counter()
    mols = synth.many([1000 + k for k in range(args.conformers)], [int(L) for L in lengths])
    log(f"built in {time.perf_counter() - t0:.1f} s")
    natoms = sum(m["n"] for m in mols)
    res = {}
    for name, prec in (("mixed", pdgpu.MIXED), ("fp64", pdgpu.FP64)):
        ctx = pdgpu.Context(opt(prec))
        t0 = time.perf_counter()
        for m in mols:
            ctx.add_system(m)
        ctx.nlist_build()
        log(f"{name}: {len(mols)} instances resident in {time.perf_counter() - t0:.1f} s")
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
    comp = lambda e: abs(e["epot_vdw"]) + abs(e["epot_elec"]) + abs(e["epot_pol_self"]) + abs(e["epot_pol_cross"])
    errs_c = [abs(r[k]["epot"] - r64[k]["epot"]) / comp(r64[k]) for k in r]
    worst = max(r, key=lambda k: errs[k] if False else abs(r[k]["epot"] - r64[k]["epot"]) / abs(r64[k]["epot"]))
    tot_err = abs(sum(r[k]["epot"] for k in r) - sum(r64[k]["epot"] for k in r)) / abs(sum(r64[k]["epot"] for k in r))
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle port)")
        cpu = cpu_baseline_batch(mols, [r64[k]["pairs"] for k in sorted(r64)])
    rl = roofline(prof, pairs, born, "mixed")
    rl["traffic"] = load_traffic(rl["kernel"], "batch")
    return {
        "metric": "atom-pair evals/s", "value": pairs * args.steps / dt, "unit": "atom-pair evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 pair maths, int64 fixed-point accumulation",
        "data": "synthetic",
        "config": {"workload": f"{args.conformers} extended conformers, 10-150 residues, one batched energy+force "
                               "evaluation per step, 10 A cutoff + Still GB, mixed precision",
                   "atoms": natoms, "pairs_per_step": pairs, "parallelism": "single GPU"},
        "fp64": {"value": pairs64 * args.steps / dt64, "ms_per_step": dt64 / args.steps * 1e3,
                 "roofline": roofline(prof64, pairs64, born64, "fp64")},
        "mixed_over_fp64": (pairs * args.steps / dt) / (pairs64 * args.steps / dt64),
        "max_energy_rel_err_mixed_vs_fp64": max(errs),
        "median_energy_rel_err_mixed_vs_fp64": float(np.median(errs)),
        "max_energy_err_rel_to_component_sum": max(errs_c),
        "batch_total_energy_rel_err": tot_err,
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
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="batch", choices=("md", "batch"))
    ap.add_argument("--conformers", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances shard by molecule across GPUs without exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import pdgpu
    out = run_md(args, pdgpu) if args.workload == "md" else run_batch(args, pdgpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
