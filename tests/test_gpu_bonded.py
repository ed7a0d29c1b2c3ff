"""GPU parity of FF_Bonded (bonds, angles, torsions, impropers) and of the
full forcefield FF_Bonded + FF_GeneralizedBorn_Still, through the C ABI,
against the compiled reference's golden vectors.

Tolerances: FP64 path -- energies relative error <= 1e-10 per component,
force RMS relative error <= 1e-9. Mixed path -- the bonded terms are FP64
maths with 64-bit fixed-point accumulation: energies <= 1e-8, force RMS
<= 1e-8; full forcefield total energy <= 1e-5 (the GB part's FP32 bound).
"""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O
import pdgpu
from gpu_helpers import context_for, options_for, rel, rms_rel
from pdgpu import synth

pytestmark = pytest.mark.gpu
CASES = G.case_names()
SD_CASES = [c for c in CASES if "sdf_xyz" in G.load(c)]
TOL = {pdgpu.FP64: (1e-10, 1e-9), pdgpu.MIXED: (1e-8, 1e-8)}


def _check_bonded(e, f, g, tag, te, tf):
    assert e["status"] == 0
    assert rel(e["epot_bond"], g[tag + "_bond"]) <= te
    assert rel(e["epot_angle"], g[tag + "_angle"]) <= te
    assert rel(e["epot_torsion"] - e["epot_improper"], g[tag + "_torsion"]) <= te
    assert rel(e["epot_improper"], g[tag + "_improper"]) <= te
    assert rel(e["epot"], g[tag + "_epot"]) <= te
    assert rms_rel(f, g[tag + "_force"]) <= tf


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_bonded_alone(name, precision):
    g = G.load(name)
    te, tf = TOL[precision]
    with context_for(g, precision, components=pdgpu.FF_BONDED) as ctx:
        i = ctx.add_system(g, bonded=True)
        e, f = ctx.calc_forces(i)
        assert e["epot_vdw"] == 0 and e["epot_pol"] == 0
        _check_bonded(e, f, g, "bf", te, tf)
        ctx.set_positions(i, g["bfp_xyz"])
        e, f = ctx.calc_forces(i)
        _check_bonded(e, f, g, "bfp", te, tf)
        # FF_Bonded::calcEnergies is its calcForces
        e2, f2 = ctx.calc_energies(i)
        _check_bonded(e2, f2, g, "bfp", te, tf)


@pytest.mark.parametrize("bit,key", [(pdgpu.FF_BONDS, "epot_bond"), (pdgpu.FF_ANGLES, "epot_angle"),
                                     (pdgpu.FF_TORSIONS, "tors"), (pdgpu.FF_IMPROPERS, "epot_improper")])
def test_component_switches(bit, key):
    g = G.load("c1_seed2")
    eo, _ = O.bonded_forces(O.OBonded(g), g["bfp_xyz"])
    want = eo["epot_torsion"] - eo["epot_improper"] if key == "tors" else eo[key]
    with context_for(g, pdgpu.FP64, components=bit) as ctx:
        i = ctx.add_system(g, xyz=g["bfp_xyz"], bonded=True)
        e, f = ctx.calc_forces(i)
        assert rel(e["epot"], want) <= 1e-10


@pytest.mark.parametrize("name", SD_CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_full_forcefield_single_point(name, precision):
    g = G.load(name)
    with context_for(g, precision) as ctx:
        i = ctx.add_system(g, bonded=True)
        e, f = ctx.calc_forces(i)
        if precision == pdgpu.FP64:
            assert rel(e["epot"], g["ff_epot"]) <= 1e-10
            for k in ("bond", "angle", "torsion", "vdw", "elec", "pol"):
                assert rel(e["epot_" + k], g["ff_" + k]) <= 1e-10, k
            assert rms_rel(f, g["ff_force"]) <= 1e-9
        else:
            assert rel(e["epot"], g["ff_epot"]) <= 1e-5
            assert rms_rel(f, g["ff_force"]) <= 1e-4


@pytest.mark.parametrize("name", SD_CASES)
def test_full_forcefield_steepest_descent(name):
    """Minimisation::runcore (SteepestDescent) over FF_Bonded + GB on the
    device against the reference's own minimisation."""
    g = G.load(name)
    steps = int(g["sd_steps"])
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, bonded=True)
        done = ctx.minimise_sd(i, steps, g["sd_stepsize"], 1)
        e, f = ctx.calc_forces(i)
        x = ctx.get_positions(i)
    assert done == steps == int(g["sdf_done"])
    assert rel(e["epot"], g["sdf_epot"]) <= 1e-9
    assert np.abs(x - g["sdf_xyz"]).max() <= 1e-7
    assert rms_rel(f, g["sdf_force"]) <= 1e-7


def test_batch_with_bonded_equals_alone_mixed():
    mols = [synth.peptide(seed=80 + k, residues=r, bonded=True) for k, r in enumerate((7, 19, 12))]
    opts = pdgpu.default_options(precision=pdgpu.MIXED, cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0,
                                 vdw_inner_cutoff=9.0, nlist_padding=2.0, dielectric_solvent=78.5)
    alone = []
    for m in mols:
        with pdgpu.Context(opts) as ctx:
            i = ctx.add_system(m, bonded=True)
            alone.append(ctx.calc_forces(i))
    with pdgpu.Context(opts) as ctx:
        filler = ctx.add_system(mols[1], bonded=True)
        ids = [ctx.add_system(m, bonded=True) for m in mols]
        ctx.remove(filler)  # freed records must not be evaluated
        r = ctx.calc_forces(pdgpu.PDGPU_ALL)
        for (e0, f0), i in zip(alone, ids):
            for k in ("epot", "epot_bond", "epot_angle", "epot_torsion", "epot_improper", "epot_pol"):
                assert r[i][k] == e0[k], k
            assert np.array_equal(ctx.forces(i), f0)


def test_bonded_set_validates_indices():
    m = synth.peptide(seed=5, residues=4, bonded=True)
    bad = dict(m)
    bad["bd_idx"] = m["bd_idx"].copy()
    bad["bd_idx"][0] = m["n"]
    with pdgpu.Context(pdgpu.default_options()) as ctx:
        i = ctx.add_system(m)
        with pytest.raises(pdgpu.PdgpuError):
            ctx.set_bonded(i, bad)
        ctx.set_bonded(i, m)
        ctx.set_bonded(i, m)  # replacing the lists keeps one copy
        e, _ = ctx.calc_forces(i)
        eo, _ = O.bonded_forces(O.OBonded(m), m["xyz"])
        # at the built geometry every torsion can sit exactly at its minimum
        # (energy 0), so the bound is relative to the amplitude sum as well
        scale = max(abs(eo["epot_torsion"]), float(np.abs(m["to_vn"]).sum()))
        assert abs(e["epot_torsion"] - eo["epot_torsion"]) <= 1e-10 * scale


MD_CASES = [c for c in CASES if "md_xyz" in G.load(c)]


@pytest.mark.parametrize("name", MD_CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_pd_velocity_verlet_matches_reference(name, precision):
    """pdgpu_md_vv_pd against PD's own MolecularDynamics run (VelocityVerlet
    integrator, FF_Bonded + GB, list refreshed every 10 steps) from the same
    positions and velocities. FP64: positions within 1e-8 A after the run;
    mixed: within 1e-3 A (FP32 pair maths through 40-60 steps)."""
    g = G.load(name)
    steps = int(g["md_steps"])
    tx, te = (1e-8, 1e-9) if precision == pdgpu.FP64 else (1e-3, 1e-4)
    with context_for(g, precision) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        e, ek = ctx.md_vv_pd(i, steps, float(g["md_dt"]), int(g["md_update_nlist"]))
        assert e["status"] == 0
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
    assert np.abs(x - g["md_xyz"].reshape(-1, 3)).max() <= tx
    vref = g["md_v"].reshape(-1, 3)
    assert np.abs(v - vref).max() <= (1e-7 if precision == pdgpu.FP64 else 1e-2) * np.abs(vref).max()
    assert rel(e["epot"], g["md_epot"]) <= te
    assert rel(ek, g["md_ekin"]) <= te


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_softvdw_alone(name, precision):
    """FF_SoftVDW on the device against the reference at the built geometry
    (PD's defaults: Hardness 12 kcal/mol, OlapOffset 0.6, VdwCutoff 15)."""
    g = G.load(name)
    te, tf = (1e-10, 1e-9) if precision == pdgpu.FP64 else (1e-5, 1e-5)
    with context_for(g, precision, components=pdgpu.FF_SOFTVDW) as ctx:
        i = ctx.add_system(g)
        e, f = ctx.calc_forces(i)
        assert e["status"] == 0
        assert rel(e["epot"], g["sv_epot"]) <= te
        assert rel(e["epot_vdw"], g["sv_vdw"]) <= te
        assert e["epot_pol"] == 0 and e["epot_elec"] == 0
        assert rms_rel(f, g["sv_force"]) <= tf


SVSD_CASES = [c for c in CASES if "svsd_xyz" in G.load(c)]


@pytest.mark.parametrize("name", SVSD_CASES)
def test_soft_clash_relief_steepest_descent(name):
    """PD's soft-core clash relief, Minimisation (SteepestDescent) over
    FF_Bonded + FF_SoftVDW, on the device against the reference's run."""
    g = G.load(name)
    steps = int(g["svsd_steps"])
    with context_for(g, pdgpu.FP64, components=pdgpu.FF_BONDED | pdgpu.FF_SOFTVDW) as ctx:
        i = ctx.add_system(g, bonded=True)
        done = ctx.minimise_sd(i, steps, g["sd_stepsize"], 1)
        e, f = ctx.calc_forces(i)
        x = ctx.get_positions(i)
    assert done == steps == int(g["svsd_done"])
    assert rel(e["epot"], g["svsd_epot"]) <= 1e-9
    assert np.abs(x - g["svsd_xyz"].reshape(-1, 3)).max() <= 1e-7
    assert rms_rel(f, g["svsd_force"]) <= 1e-7


def test_softvdw_with_gb_is_unsupported():
    g = G.load("tri_ala")
    with pytest.raises(pdgpu.PdgpuError):
        context_for(g, pdgpu.FP64, components=pdgpu.FF_GB | pdgpu.FF_SOFTVDW)
