"""GPU tests of the step drivers and of synthetic workloads at size:
batched synthetic conformers against the oracle, the textbook velocity
Verlet against the oracle's, and the continuous-batching minimiser (each
request's result is the same alone and in the mix)."""
import numpy as np
import pytest

import oracle_lib as O
import pdgpu
from gpu_helpers import rel, rms_rel
from pdgpu import serve, synth

pytestmark = pytest.mark.gpu

KW = dict(cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0, vdw_inner_cutoff=9.0, vdw14_scaling=0.5,
          elec14_scaling=0.83333, dielectric_solvent=78.5)


def opts(precision, **kw):
    d = dict(KW, nlist_padding=2.0)
    d.update(kw)
    return pdgpu.default_options(precision=precision, **d)


def oparams(**kw):
    d = dict(KW, padding=2.0)
    d.update(kw)
    return O.default_params(**d)


def osys(m):
    return O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                     m["mid"], m["band"], m["sp_i"], m["sp_j"], m["sp_o"])


def oracle_forces(m, xyz=None, **kw):
    x = m["xyz"] if xyz is None else xyz
    p = oparams(**kw)
    s = osys(m)
    rs, ent = O.nlist_build(s, x, max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e, f, born, _ = O.gb_calc_forces(s, x, rs, ent, p)
    return e, f


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_zero_width_switch_matches_oracle(precision):
    # InnerCutoff == Cutoff (and the same for vdW): no switching region, the
    # case PD's own defaults give when only Cutoff is set to 12
    m = synth.peptide(seed=77, residues=18)
    kw = dict(inner_cutoff=12.0, vdw_inner_cutoff=12.0)
    eo, fo = oracle_forces(m, **kw)
    with pdgpu.Context(opts(precision, **kw)) as ctx:
        i = ctx.add_system(m)
        e, f = ctx.calc_forces(i)
        assert e["status"] == 0
        assert rel(e["epot"], eo["epot"]) <= (1e-10 if precision == pdgpu.FP64 else 1e-5)
        assert rms_rel(f, fo) <= (1e-9 if precision == pdgpu.FP64 else 1e-4)


ALLPAIRS = dict(cutoff=1.0e4, inner_cutoff=1.0e4, vdw_cutoff=12.0, vdw_inner_cutoff=9.0)


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_all_pairs_gb_matches_oracle(precision):
    # all-pairs GB: Cutoff (Coulomb + GB) and the Still cutoff beyond the
    # molecule, LJ at 12 A; full tiles are evaluated densely, partial ones
    # (the last block) from lists
    m = synth.peptide(seed=88, residues=40)
    kw = dict(ALLPAIRS)
    p = oparams(still_cutoff=1.0e4, **kw)
    s = osys(m)
    rs, ent = O.nlist_build(s, m["xyz"], max(p.cutoff, p.vdw_cutoff), p.padding, True)
    eo, fo, born_o, _ = O.gb_calc_forces(s, m["xyz"], rs, ent, p)
    en, fn, _ = O.gb_calc_energies(s, m["xyz"], rs, ent, p)
    n = m["n"]
    tol_e = 1e-10 if precision == pdgpu.FP64 else 1e-5
    tol_f = 1e-9 if precision == pdgpu.FP64 else 1e-4
    with pdgpu.Context(opts(precision, gbsa_still_cutoff=1.0e4, **kw)) as ctx:
        i = ctx.add_system(m)
        e, f = ctx.calc_forces(i)
        assert e["status"] == 0
        assert e["pairs"] == n * (n - 1) // 2
        assert rel(e["epot"], eo["epot"]) <= tol_e
        assert rms_rel(f, fo) <= tol_f
        e, f = ctx.calc_energies(i)
        assert rel(e["epot"], en["epot"]) <= tol_e
        assert rms_rel(f, fn) <= tol_f


@pytest.fixture(scope="module")
def conformers():
    rng = np.random.default_rng(5)
    return [synth.peptide(seed=500 + k, residues=int(L)) for k, L in enumerate(rng.integers(10, 41, size=12))]


def test_batched_synthetic_matches_oracle(conformers):
    with pdgpu.Context(opts(pdgpu.FP64)) as c64, pdgpu.Context(opts(pdgpu.MIXED)) as cmx:
        ids64 = [c64.add_system(m) for m in conformers]
        idsmx = [cmx.add_system(m) for m in conformers]
        r64 = c64.calc_forces(pdgpu.PDGPU_ALL)
        rmx = cmx.calc_forces(pdgpu.PDGPU_ALL)
        for m, i, j in zip(conformers, ids64, idsmx):
            eo, fo = oracle_forces(m)
            assert rel(r64[i]["epot"], eo["epot"]) <= 1e-10
            assert rms_rel(c64.forces(i), fo) <= 1e-9
            assert rel(rmx[j]["epot"], eo["epot"]) <= 1e-5
            assert rms_rel(cmx.forces(j), fo) <= 1e-4


def test_large_basins_chain_matches_oracle():
    m = synth.peptide(seed=2, residues=120, mode="basins")
    eo, fo = oracle_forces(m)
    for prec, te, tf in ((pdgpu.FP64, 1e-10, 1e-9), (pdgpu.MIXED, 1e-5, 1e-4)):
        with pdgpu.Context(opts(prec)) as ctx:
            i = ctx.add_system(m)
            e, f = ctx.calc_forces(i)
            assert rel(e["epot"], eo["epot"]) <= te
            assert rms_rel(f, fo) <= tf


def test_velocity_verlet_matches_oracle():
    m = synth.peptide(seed=9, residues=8)
    rng = np.random.default_rng(1)
    v0 = rng.normal(size=(m["n"], 3)) * np.sqrt(1.38065e-23 * 300 / m["mass"])[:, None]
    steps, dt = 40, 0.5e-15
    rc, xo, vo, te, tk = O.md_vv(osys(m), m["xyz"], v0, oparams(), steps, dt, 10)
    with pdgpu.Context(opts(pdgpu.FP64)) as ctx:
        i = ctx.add_system(m)
        ctx.set_velocities(i, v0)
        e, ek = ctx.md_vv(i, steps, dt, 10)
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
    assert np.abs(x - xo).max() <= 1e-7
    assert rms_rel(v, vo) <= 1e-7
    assert rel(e["epot"], te[-1]) <= 1e-8
    assert rel(ek, tk[-1]) <= 1e-8


def test_mixed_md_tracks_fp64():
    m = synth.peptide(seed=9, residues=12)
    rng = np.random.default_rng(2)
    v0 = rng.normal(size=(m["n"], 3)) * np.sqrt(1.38065e-23 * 300 / m["mass"])[:, None]
    out = {}
    for prec in (pdgpu.FP64, pdgpu.MIXED):
        with pdgpu.Context(opts(prec)) as ctx:
            i = ctx.add_system(m)
            ctx.set_velocities(i, v0)
            e, ek = ctx.md_vv(i, 100, 1e-15, 10)
            out[prec] = (e["epot"] + ek, ctx.get_positions(i))
    assert rel(out[pdgpu.MIXED][0], out[pdgpu.FP64][0]) <= 1e-4
    assert np.abs(out[pdgpu.MIXED][1] - out[pdgpu.FP64][1]).max() <= 1e-3


def _alone(m, steps, prec):
    with pdgpu.Context(opts(prec)) as ctx:
        i = ctx.add_system(m)
        done = ctx.minimise_sd(i, steps, 2e5, 1)
        e, f = ctx.calc_forces(i)
        return done, e["epot"], ctx.get_positions(i)


def test_sd_step_api_equals_minimise():
    m = synth.peptide(seed=21, residues=10)
    done, e_alone, x_alone = _alone(m, 30, pdgpu.MIXED)
    with pdgpu.Context(opts(pdgpu.MIXED)) as ctx:
        i = ctx.add_system(m)
        ctx.sd_start(i, 30, 2e5, 1, 0.0)
        running = 1
        while running:
            d, st, running = ctx.sd_advance(7)
        e, st = ctx.sd_finish(i)
        assert st == done == 30
        assert e["epot"] == e_alone
        assert np.array_equal(ctx.get_positions(i), x_alone)


def test_continuous_batching_same_alone_and_in_mix():
    mols = [synth.peptide(seed=31 + k, residues=r) for k, r in enumerate((8, 14, 20, 11))]
    steps = 40
    alone = [_alone(m, steps, pdgpu.MIXED) for m in mols]
    with pdgpu.Context(opts(pdgpu.MIXED)) as ctx:
        ids = {}
        results = {}
        # staggered admissions while others are mid-run
        for k, m in enumerate(mols):
            ids[k] = ctx.add_system(m)
            ctx.sd_start(ids[k], steps, 2e5, 1, 0.0)
            done, st, running = ctx.sd_advance(5 + 3 * k)
            for kk, iid in list(ids.items()):
                if done[iid] and kk not in results:
                    results[kk] = ctx.sd_finish(iid) + (ctx.get_positions(iid),)
        while len(results) < len(mols):
            done, st, running = ctx.sd_advance(4)
            for kk, iid in ids.items():
                if done[iid] and kk not in results:
                    results[kk] = ctx.sd_finish(iid) + (ctx.get_positions(iid),)
        for k in range(len(mols)):
            e, st, x = results[k]
            assert st == alone[k][0]
            # 64-bit fixed-point accumulation: identical bits alone and in the mix
            assert e["epot"] == alone[k][1]
            assert np.array_equal(x, alone[k][2])


def test_server_poisson_requests_complete_and_match():
    mols = [synth.peptide(seed=60 + k, residues=int(r), bonded=True) for k, r in
            enumerate(np.random.default_rng(3).integers(6, 25, size=10))]
    reqs = [serve.Request(k, m, float(t)) for k, (m, t) in
            enumerate(zip(mols, serve.poisson_arrivals(400.0, len(mols), seed=4)))]
    srv = serve.MinimisationServer(opts(pdgpu.MIXED), memory_budget=64 << 20, max_steps=60, grms_tol=0.1,
                                   steps_per_tick=4)
    try:
        out = srv.run(reqs)
    finally:
        srv.close()
    assert len(out) == len(reqs)
    assert not any(r.failed for r in out)
    for r in out[:3]:
        with pdgpu.Context(opts(pdgpu.MIXED)) as ctx:
            i = ctx.add_system(r.mol, bonded=True)
            st = ctx.minimise_sd(i, 60, 2e5, -1, 0.1)
            e, _ = ctx.calc_forces(i)
            x = ctx.get_positions(i)
        assert st == r.steps
        # same bits alone and in the mix
        assert e["epot"] == r.energies["epot"]
        assert np.array_equal(x, r.xyz)


def test_admission_respects_budget():
    m = synth.peptide(seed=70, residues=30)
    srv = serve.MinimisationServer(opts(pdgpu.MIXED), memory_budget=16 << 10, max_steps=5)
    try:
        with pytest.raises(pdgpu.PdgpuError):
            srv.try_admit(serve.Request(0, m, 0.0), 0.0)
    finally:
        srv.close()


def test_skin_rebuild_matches_every_step_rebuild():
    """update_nlist = -1 (rebuild when an atom moved more than half the
    padding) evaluates the same pairs as PD's every-step rebuild, so the
    minimisations agree to FP64 summation order."""
    m = synth.peptide(seed=44, residues=16, bonded=True)
    out = {}
    for un in (1, -1):
        with pdgpu.Context(opts(pdgpu.FP64)) as ctx:
            i = ctx.add_system(m, bonded=True)
            st = ctx.minimise_sd(i, 40, 2e5, un)
            e, _ = ctx.calc_forces(i)
            out[un] = (st, e["epot"], ctx.get_positions(i))
    assert out[1][0] == out[-1][0] == 40
    assert rel(out[1][1], out[-1][1]) <= 1e-9
    assert np.abs(out[1][2] - out[-1][2]).max() <= 1e-8


def test_graph_replay_equals_eager():
    """Step loops replay captured HIP graphs; every replay must give the bits
    the eager launches give (mixed path: order-independent sums). Eager work
    and synchronisations (state read-back, position reads) sit between the
    replays."""
    mols = [synth.peptide(seed=90 + k, residues=r, bonded=True) for k, r in enumerate((9, 15))]
    res = {}
    for g in (0, 1):
        with pdgpu.Context(opts(pdgpu.MIXED, use_graphs=g)) as ctx:
            ids = [ctx.add_system(m, bonded=True) for m in mols]
            for i in ids:
                ctx.sd_start(i, 40, 2e5, -1, 0.0)
            snaps = []
            for _ in range(4):
                ctx.sd_advance(7)
                snaps.append([ctx.get_positions(i).copy() for i in ids])
            for i in ids:
                ctx.sd_finish(i)
            rng = np.random.default_rng(5)
            for i, m in zip(ids, mols):
                ctx.set_velocities(i, rng.normal(size=(m["n"], 3)) * 300.0)
            for _ in range(3):
                e, ek = ctx.md_vv(pdgpu.PDGPU_ALL, 12, 1e-15, 5)
                snaps.append([ctx.get_positions(i).copy() for i in ids])
            res[g] = snaps
    for a, b in zip(res[0], res[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_evaluation_graph_replay_equals_eager():
    """calc_forces / calc_energies replay a captured graph of the evaluation
    kernels; each replay (positions moved eagerly in between, energies and
    forces read back after each) gives the bits of eager launches."""
    mols = [synth.peptide(seed=120 + k, residues=r, bonded=True) for k, r in enumerate((7, 12, 18))]
    rng = np.random.default_rng(9)
    moves = [[rng.normal(scale=0.05, size=(m["n"], 3)) for m in mols] for _ in range(4)]
    res = {}
    for g in (0, 1):
        out = []
        with pdgpu.Context(opts(pdgpu.MIXED, use_graphs=g, max_instances=8)) as ctx:
            ids = [ctx.add_system(m, bonded=True) for m in mols]
            for mv in moves:
                for i, m, d in zip(ids, mols, mv):
                    ctx.set_positions(i, m["xyz"] + d)
                r = ctx.calc_forces(pdgpu.PDGPU_ALL)
                out.append([(r[i]["epot"], ctx.forces(i).copy()) for i in ids])
                e, f = ctx.calc_energies(ids[1])
                out.append([(e["epot"], f.copy())])
        res[g] = out
    for a, b in zip(res[0], res[1]):
        for (ea, fa), (eb, fb) in zip(a, b):
            assert ea == eb
            assert np.array_equal(fa, fb)
