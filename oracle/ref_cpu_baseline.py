"""Times the compiled reference's own CPU path beside the oracle port on the
same molecules, on this host. Test infrastructure: runs only where the
reference tree and its parameter file exist (make -C oracle ref_bench).

  python3 oracle/ref_cpu_baseline.py [--single S] [--multi S] [--procs P]

Molecules: the batch workload's first conformers (lengths and sequences of
conformers 0, 1, 2, ... of the 4,096-conformer batch), built extended by
PD's NewProtein with PD's parameter file, one evaluation each (neighbour
list + Born radii + fused pair pass + chain rule) at 10 A cutoff, 2 A
padding. The reference (oracle/_ref/ref_bench, the reference's own sources
at its own optimisation flags) runs first on one process for ~S seconds,
dumping each molecule's inputs; then on P processes over disjoint molecule
ranges for up to ~S seconds each. The oracle port (oracle/liboracle.so, -O2)
then evaluates the same dumped molecules, one process, then P processes.
Both pair counts (unordered pairs within the cutoff) and energies are
checked to agree molecule by molecule; the printout is a markdown summary
with the reference/port time ratio that ties the port's figure on any host
to the reference's."""
import argparse
import multiprocessing as mp
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
sys.path.insert(0, os.path.join(ROOT, "pd.arcus_amd"))

PARAM = "/root/reference/param/amber99aa.ff"
BIN = os.path.join(HERE, "_ref", "ref_bench")
DUMP = os.path.join(HERE, "_ref", "bench")


def run_ref(first, count, budget, lengths, dump):
    out = subprocess.run([BIN, PARAM, dump, str(first), str(count), str(budget), lengths],
                         check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.splitlines():
        f = line.split()
        if len(f) == 5 and f[0].isdigit():
            rows.append((int(f[0]), int(f[1]), int(f[2]), float(f[3]), float(f[4])))
    return rows


def _ref_job(a):
    return run_ref(*a)


def _port_job(job):
    import oracle_lib as O
    from make_golden import read_records
    ks, budget = job
    p = O.default_params(cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0, vdw_inner_cutoff=7.0, vdw14_scaling=0.5,
                         elec14_scaling=0.83333, dielectric_solvent=78.5, padding=2.0)
    rows = []
    t_end = time.perf_counter() + budget
    for k in ks:
        m = read_records(os.path.join(DUMP, f"bench_{k}.bin"))
        xyz = m["xyz"].reshape(-1, 3)
        s = O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                      m["mid"][0], m["band"], m["sp_i"], m["sp_j"], m["sp_o"])
        t0 = time.perf_counter()
        rs, ent = O.nlist_build(s, xyz, 10.0, 2.0, True)
        e, _, _, _ = O.gb_calc_forces(s, xyz, rs, ent, p)
        dt = time.perf_counter() - t0
        j = (ent & 0xFFFFFF).astype(np.int64)
        i = np.repeat(np.arange(s.n), np.diff(rs))
        keep = j > i
        pairs = int((((xyz[i[keep]] - xyz[j[keep]]) ** 2).sum(1) <= 100.0).sum())
        rows.append((k, s.n, pairs, dt, e["epot"]))
        if time.perf_counter() > t_end:
            break
    return rows


def rate(rows):
    return sum(r[2] for r in rows) / sum(r[3] for r in rows)


def aggregate(groups):
    """Whole-host rate of concurrent processes: the sum of each process's
    pairs over its own timed seconds (molecule construction, parameter
    reading and file input are outside the timed regions)."""
    return sum(rate(g) for g in groups if g)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--single", type=float, default=20.0)
    ap.add_argument("--multi", type=float, default=10.0)
    ap.add_argument("--procs", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    from pdgpu import workloads
    os.makedirs(DUMP, exist_ok=True)
    lengths = os.path.join(DUMP, "lengths.txt")
    with open(lengths, "w") as f:
        f.write("\n".join(map(str, workloads.batch_lengths())) + "\n")

    ref1 = run_ref(0, 4096, a.single, lengths, DUMP)
    n1 = len(ref1)
    per = max(1, n1 // a.procs + 1)
    with mp.get_context("spawn").Pool(a.procs) as pool:
        refg = pool.map(_ref_job, [(q * per, per, a.multi, lengths, "-") for q in range(a.procs)])
    refp = [r for rows in refg for r in rows]

    port1 = _port_job(([r[0] for r in ref1], 1e9))
    ks = [r[0] for r in ref1]
    with mp.get_context("spawn").Pool(a.procs) as pool:
        portg = pool.map(_port_job, [(ks[q::a.procs], 1e9) for q in range(a.procs)])
    portp = [r for rows in portg for r in rows]

    bad_pairs = sum(1 for r, p in zip(ref1, port1) if r[2] != p[2])
    rel = max(abs(r[4] - p[4]) / abs(r[4]) for r, p in zip(ref1, port1))
    per_mol = np.array([r[3] / p[3] for r, p in zip(ref1, port1)])
    atoms = sum(r[1] for r in ref1)
    print("| leg | molecules | atoms | pairs | time (s) | atom-pair evals/s |")
    print("|---|---|---|---|---|---|")
    print(f"| reference, 1 process | {n1} | {atoms} | {sum(r[2] for r in ref1)} | "
          f"{sum(r[3] for r in ref1):.2f} | {rate(ref1):.4g} |")
    print(f"| reference, {a.procs} processes | {len(refp)} | {sum(r[1] for r in refp)} | "
          f"{sum(r[2] for r in refp)} | {max(sum(r[3] for r in g) for g in refg):.2f} max per process | "
          f"{aggregate(refg):.4g} |")
    print(f"| oracle port, 1 process | {len(port1)} | {sum(r[1] for r in port1)} | "
          f"{sum(r[2] for r in port1)} | {sum(r[3] for r in port1):.2f} | {rate(port1):.4g} |")
    print(f"| oracle port, {a.procs} processes | {len(portp)} | {sum(r[1] for r in portp)} | "
          f"{sum(r[2] for r in portp)} | {max(sum(r[3] for r in g) for g in portg):.2f} max per process | "
          f"{aggregate(portg):.4g} |")
    print()
    print(f"reference/port single-process time on the same {n1} molecules: "
          f"{sum(r[3] for r in ref1) / sum(r[3] for r in port1):.3f} "
          f"(per molecule: median {np.median(per_mol):.3f}, min {per_mol.min():.3f}, max {per_mol.max():.3f})")
    print(f"pair counts that differ: {bad_pairs} of {n1}; max |epot| relative difference: {rel:.2e}")
    print(f"host: {os.cpu_count()} logical CPUs; {a.procs} processes used for the multi-process legs")


if __name__ == "__main__":
    main()
