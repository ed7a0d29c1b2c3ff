"""The benchmark configurations at their own sizes (pdgpu.workloads, the
inputs bench.py runs), checked on the GPU:

- batch: all 4,096 conformers (10-150 residues, 10 A) in one context per
  precision; a seeded sample of 32 against the oracle (FP64 <= 1e-10
  energy, <= 1e-9 force RMS; mixed within the contract of gpu_helpers),
  every conformer mixed against FP64, and the sample re-evaluated alone,
  bit for bit equal to its batched result (mixed);
- chain: the 500-residue chain after clash relief, FP64 single point
  against the oracle, then 200 mixed velocity-Verlet steps with the FP64
  path evaluated at the mixed positions every 100 steps;
- serving: 40 Poisson-arriving requests of 20-400 residues minimised to
  force RMS < 0.1 kcal/mol/A or 1,000 steps; a sample minimised alone gives
  the same bits, and final energies agree with the oracle at the final
  positions;
- allpairs: the 32,034-atom all-pairs chain, FP64 single point against the
  oracle (one ~30 s oracle evaluation), then 200 mixed steps checked
  against FP64 every 100 steps.
The single-peptide configuration is the golden fixtures c1_seed1..3
(test_gpu_parity.py, test_gpu_bonded.py)."""
import numpy as np
import pytest

import oracle_lib as O
import pdgpu
from gpu_helpers import MIXED_FORCE_RMS, mixed_energy_bound, rel, rms_rel
from pdgpu import serve, workloads as W

pytestmark = pytest.mark.gpu


def osys(m, bonded=False):
    return O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                     m["mid"], m["band"], m["sp_i"], m["sp_j"], m["sp_o"], bonded=O.OBonded(m) if bonded else None)


def oracle_gb(m, name, xyz=None):
    """FF_GeneralizedBorn_Still::calcForces of the oracle (energies, forces
    in N) on the workload's knobs."""
    x = m["xyz"] if xyz is None else xyz
    p = O.default_params(**W.oracle_params(name))
    s = osys(m)
    rs, ent = O.nlist_build(s, x, max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e, f, _, _ = O.gb_calc_forces(s, x, rs, ent, p)
    return e, f


def gb_only(name, precision, **kw):
    return W.options(name, precision, components=pdgpu.FF_GB, **kw)


# ------------------------------------------------------------------ batch
@pytest.fixture(scope="module")
def batch():
    return W.batch_molecules(start="spawn")


def test_batch_4096_against_oracle_and_fp64(batch):
    n = len(batch)
    assert n == 4096
    sample = sorted(np.random.default_rng(2024).choice(n, size=32, replace=False).tolist())
    res = {}
    for prec in (pdgpu.FP64, pdgpu.MIXED):
        with pdgpu.Context(W.options("batch", prec, max_instances=n)) as ctx:
            ids = [ctx.add_system(m) for m in batch]
            r = ctx.calc_forces(pdgpu.PDGPU_ALL)
            forces = {k: ctx.forces(ids[k]) for k in sample}
            res[prec] = ([r[i] for i in ids], forces)
    e64, f64 = res[pdgpu.FP64]
    emx, fmx = res[pdgpu.MIXED]
    assert all(e["status"] == 0 for e in e64) and all(e["status"] == 0 for e in emx)
    # every conformer: mixed against FP64 within the contract
    for a, b in zip(emx, e64):
        assert a["pairs"] == b["pairs"]
        assert abs(a["epot"] - b["epot"]) <= mixed_energy_bound(b["epot"], b["pairs"])
    # the sample against the oracle
    for k in sample:
        eo, fo = oracle_gb(batch[k], "batch")
        assert rel(e64[k]["epot"], eo["epot"]) <= 1e-10
        assert rms_rel(f64[k], fo) <= 1e-9
        assert abs(emx[k]["epot"] - eo["epot"]) <= mixed_energy_bound(eo["epot"], e64[k]["pairs"])
        assert rms_rel(fmx[k], fo) <= MIXED_FORCE_RMS
    # the sample alone: the same bits as in the batch (fixed-point sums)
    for k in sample[:8]:
        with pdgpu.Context(W.options("batch", pdgpu.MIXED)) as ctx:
            i = ctx.add_system(batch[k])
            e, f = ctx.calc_forces(i)
        assert e["epot"] == emx[k]["epot"]
        assert np.array_equal(f, fmx[k])


# ------------------------------------------------------------------ chain
@pytest.fixture(scope="module")
def chain():
    m = W.chain_molecule()
    x = W.relax(lambda: pdgpu.Context(W.options("chain", pdgpu.MIXED)), m)
    return m, x


def test_chain_fp64_against_oracle_then_mixed_md(chain):
    m, x0 = chain
    eo, fo = oracle_gb(m, "chain", x0)
    with pdgpu.Context(gb_only("chain", pdgpu.FP64)) as c64:
        j = c64.add_system(m, xyz=x0)
        e, f = c64.calc_forces(j)
        assert rel(e["epot"], eo["epot"]) <= 1e-10
        assert rms_rel(f, fo) <= 1e-9
    vel = W.maxwell(m["mass"], 300.0, seed=7)
    with pdgpu.Context(W.options("chain", pdgpu.MIXED)) as cm, pdgpu.Context(W.options("chain", pdgpu.FP64)) as c64:
        i = cm.add_system(m, xyz=x0, bonded=True)
        j = c64.add_system(m, xyz=x0, bonded=True)
        cm.set_velocities(i, vel)
        for _ in range(2):
            e, ek = cm.md_vv(i, 100, 1e-15, 10)
            assert e["status"] == 0
            x = cm.get_positions(i)
            em, fm = cm.calc_forces(i)
            c64.set_positions(j, x)
            c64.nlist_build(j)  # the list of the mixed positions, as the mixed run keeps it
            ed, fd = c64.calc_forces(j)
            assert abs(em["epot"] - ed["epot"]) <= mixed_energy_bound(ed["epot"], ed["pairs"])
            assert rms_rel(fm, fd) <= MIXED_FORCE_RMS
            # the trajectory stays physical: no blow-up from the relaxed start
            temp = 2.0 * ek / (3.0 * m["n"] * W.KB)
            assert 100.0 < temp < 1000.0


# ---------------------------------------------------------------- serving
def test_serving_requests_alone_and_in_mix():
    sizes = W.serving_sizes(40)
    mols = W.serving_molecules(40, start="spawn")
    assert min(sizes) >= 20 and max(sizes) <= 400
    opts = W.options("serving", pdgpu.MIXED, max_instances=64)
    reqs = [serve.Request(k, mm, float(t)) for k, (mm, t) in
            enumerate(zip(mols, serve.poisson_arrivals(500.0, len(mols), seed=5)))]
    srv = serve.MinimisationServer(opts, memory_budget=4 << 30, max_steps=1000, grms_tol=0.1, steps_per_tick=25)
    try:
        out = srv.run(reqs)
    finally:
        srv.close()
    assert len(out) == len(reqs) and not any(r.failed for r in out)
    by_id = {r.rid: r for r in out}
    # the largest, the smallest and two more, alone
    order = np.argsort(sizes)
    for k in (int(order[-1]), int(order[0]), 7, 23):
        r = by_id[k]
        with pdgpu.Context(W.options("serving", pdgpu.MIXED)) as ctx:
            i = ctx.add_system(r.mol, bonded=True)
            st = ctx.minimise_sd(i, 1000, 2e5, -1, 0.1)
            e, _ = ctx.calc_forces(i)
            x = ctx.get_positions(i)
        assert st == r.steps
        assert e["epot"] == r.energies["epot"]
        assert np.array_equal(x, r.xyz)
        # the GB part of the final energy against the oracle at the final
        # positions (FF_Bonded is held to the reference by test_gpu_bonded)
        eo, _ = oracle_gb(r.mol, "serving", x)
        with pdgpu.Context(gb_only("serving", pdgpu.MIXED)) as ctx:
            i = ctx.add_system(r.mol, xyz=x)
            eg, _ = ctx.calc_forces(i)
        assert abs(eg["epot"] - eo["epot"]) <= mixed_energy_bound(eo["epot"], eg["pairs"])


# --------------------------------------------------------------- allpairs
def test_allpairs_32k_fp64_against_oracle_then_mixed_md():
    m = W.allpairs_molecule()
    n = m["n"]
    assert n > 30000
    x0 = W.relax(lambda: pdgpu.Context(W.options("chain", pdgpu.MIXED)), m)
    eo, fo = oracle_gb(m, "allpairs", x0)
    with pdgpu.Context(gb_only("allpairs", pdgpu.FP64)) as c64:
        j = c64.add_system(m, xyz=x0)
        e, f = c64.calc_forces(j)
        assert e["pairs"] == n * (n - 1) // 2
        assert rel(e["epot"], eo["epot"]) <= 1e-10
        assert rms_rel(f, fo) <= 1e-9
    vel = W.maxwell(m["mass"], 300.0, seed=9)
    with pdgpu.Context(W.options("allpairs", pdgpu.MIXED)) as cm, \
            pdgpu.Context(W.options("allpairs", pdgpu.FP64)) as c64:
        i = cm.add_system(m, xyz=x0, bonded=True)
        j = c64.add_system(m, xyz=x0, bonded=True)
        cm.set_velocities(i, vel)
        for _ in range(2):
            e, ek = cm.md_vv(i, 100, 1e-15, 0)
            assert e["status"] == 0
            x = cm.get_positions(i)
            em, fm = cm.calc_forces(i)
            c64.set_positions(j, x)
            c64.nlist_build(j)  # the list of the mixed positions, as the mixed run keeps it
            ed, fd = c64.calc_forces(j)
            assert abs(em["epot"] - ed["epot"]) <= mixed_energy_bound(ed["epot"], ed["pairs"])
            assert rms_rel(fm, fd) <= MIXED_FORCE_RMS
