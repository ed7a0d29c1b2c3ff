"""FF_NonBonded as a component of its own (PDGPU_FF_NONBONDED): vacuum LJ +
Coulomb with PD's default force switching, against the compiled
reference's FF_NonBonded runs in the golden fixtures (nb_*: Cutoff 15,
InnerCutoff 12, amber99 1-4 scalings, every case). FP64: energies <=
1e-10, virial <= 1e-10, forces <= 1e-9 RMS; mixed: the per-instance
contract of gpu_helpers, forces <= 1e-4 RMS. PD's FF_NonBonded evaluates
calcEnergies and calcForces the same way (UsePartialRecalc off)."""
import numpy as np
import pytest

import golden_io as G
import pdgpu
from gpu_helpers import MIXED_FORCE_RMS, mixed_energy_bound, rel, rms_rel

pytestmark = pytest.mark.gpu
CASES = G.case_names()


def nb_options(g, precision, **extra):
    cut, inner, vcut, vinner, s14v, s14e, diel = (float(x) for x in g["nb_params"])
    return pdgpu.default_options_nonbonded(precision=precision, cutoff=cut, inner_cutoff=inner, vdw_cutoff=vcut,
                                           vdw_inner_cutoff=vinner, vdw14_scaling=s14v, elec14_scaling=s14e,
                                           dielectric_solute=diel, nlist_padding=float(g["params"][17]),
                                           components=pdgpu.FF_NONBONDED, **extra)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_nonbonded_matches_reference(name, precision):
    g = G.load(name)
    with pdgpu.Context(nb_options(g, precision)) as ctx:
        i = ctx.add_system(g)
        for call in (ctx.calc_forces, ctx.calc_energies):
            e, f = call(i)
            assert e["status"] == 0
            assert e["epot_pol"] == 0.0
            if precision == pdgpu.FP64:
                assert rel(e["epot"], g["nb_epot"]) <= 1e-10
                assert rel(e["epot_vdw"], g["nb_epot_vdw"]) <= 1e-10
                assert rel(e["epot_elec"], g["nb_epot_elec"]) <= 1e-10
                assert rel(e["virial"], g["nb_virial"]) <= 1e-10
                assert rms_rel(f, g["nb_force"].reshape(-1, 3)) <= 1e-9
            else:
                assert abs(e["epot"] - g["nb_epot"]) <= mixed_energy_bound(float(g["nb_epot"]), e["pairs"])
                assert rms_rel(f, g["nb_force"].reshape(-1, 3)) <= MIXED_FORCE_RMS


def test_nonbonded_with_bonded_is_the_sum():
    """FF_Bonded + FF_NonBonded in one context: energies and forces are the
    sums of the two components evaluated alone."""
    g = G.load("c1_seed1")
    res = {}
    for comp in (pdgpu.FF_NONBONDED, pdgpu.FF_BONDED, pdgpu.FF_NONBONDED | pdgpu.FF_BONDED):
        opts = nb_options(g, pdgpu.FP64)
        opts.components = comp
        with pdgpu.Context(opts) as ctx:
            i = ctx.add_system(g, bonded=True)
            res[comp] = ctx.calc_forces(i)
    e_nb, f_nb = res[pdgpu.FF_NONBONDED]
    e_bd, f_bd = res[pdgpu.FF_BONDED]
    e_all, f_all = res[pdgpu.FF_NONBONDED | pdgpu.FF_BONDED]
    assert rel(e_all["epot"], e_nb["epot"] + e_bd["epot"]) <= 1e-12
    assert rms_rel(f_all, f_nb + f_bd) <= 1e-12


def test_nonbonded_refuses_gb_in_the_same_context():
    o = pdgpu.default_options_nonbonded(components=pdgpu.FF_NONBONDED | pdgpu.FF_GB)
    with pytest.raises(pdgpu.PdgpuError):
        pdgpu.Context(o)
