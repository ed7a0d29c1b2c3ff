"""The CPU oracle against golden vectors of the compiled reference.

Fixtures (tests/golden/*.npz) were written by oracle/ref_dump.cpp driving
the reference's own FF_GeneralizedBorn_Still / FF_NonBonded / NeighbourList /
BondOrder / Minimisation on peptides built by its NewProtein builder.
"""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O

CASES = G.case_names()


def osys(g):
    return O.OSystem(g["charge"], g["radius"], g["epsilon"], g["gb_radius"], g["vj"], g["terms123"], g["mass"],
                     g["mid"], g["band"], g["sp_i"], g["sp_j"], g["sp_o"])


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def rms_rel(f, g):
    return np.sqrt(np.mean((f - g) ** 2)) / np.sqrt(np.mean(g ** 2))


def test_fixture_set_complete():
    assert {"c1_seed1", "c1_seed2", "c1_seed3", "tri_ala", "gly_ala", "res60_seed4"} <= set(CASES)


@pytest.mark.parametrize("name", CASES)
def test_bondorder_band_exact(name):
    g = G.load(name)
    mid, band, spi, spj, spo = O.bondorder(g["resid"], g["a12"], g["b12"], g["a13"], g["b13"], g["a14"], g["b14"])
    assert mid == g["mid"]
    assert np.array_equal(band, g["band"])
    assert len(spi) == len(g["sp_i"])


@pytest.mark.parametrize("name", CASES)
def test_nlist_exact(name):
    g = G.load(name)
    s = osys(g)
    p = g["param_dict"]
    lc = max(p["cutoff"], p["vdw_cutoff"])
    rs, ent = O.nlist_build(s, g["xyz"], lc, p["padding"], True)
    assert np.array_equal(rs, g["nl_full_row_start"]) and np.array_equal(ent, g["nl_full_entries"])
    rs, ent = O.nlist_build(s, g["xyz"], lc, p["padding"], False)
    assert np.array_equal(rs, g["nl_half_row_start"]) and np.array_equal(ent, g["nl_half_entries"])


@pytest.mark.parametrize("name", CASES)
def test_fixed_terms(name):
    g = G.load(name)
    p = O.default_params(**g["param_dict"])
    vj, t = O.fixed_terms(g["gb_radius"], g["a12"], g["b12"], g["d12"], g["a13"], g["b13"], g["d13"], p)
    assert np.allclose(vj, g["vj"], rtol=1e-13, atol=1e-13)
    assert np.allclose(t, g["terms123"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", CASES)
def test_force_path(name):
    g = G.load(name)
    p = O.default_params(**g["param_dict"])
    e, f, born, deda = O.gb_calc_forces(osys(g), g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    assert np.max(np.abs(born - g["born"]) / np.abs(g["born"])) < 1e-13
    for k in ("epot", "epot_vdw", "epot_elec", "epot_pol", "epot_pol_self", "epot_pol_cross"):
        assert rel(e[k], g[k]) < 1e-12, k
    assert rms_rel(f, g["force"]) < 1e-13
    assert np.max(np.abs(deda - g["deda"])) / np.max(np.abs(g["deda"])) < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_energy_path(name):
    g = G.load(name)
    p = O.default_params(**g["param_dict"])
    e, f, _ = O.gb_calc_energies(osys(g), g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    assert rel(e["epot"], g["en_epot"]) < 1e-12
    assert rel(e["epot_vdw"], g["en_epot_vdw"]) < 1e-12
    assert rel(e["epot_elec"], g["en_epot_elec"]) < 1e-12
    assert rel(e["epot_pol"], g["en_pol"]) < 1e-12
    assert rel(e["virial"], g["en_virial"]) < 1e-12
    assert rms_rel(f, g["en_force"]) < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_nonbonded_alone_force_switch(name):
    g = G.load(name)
    npar = g["nb_params"]
    p = O.default_params(cutoff=npar[0], inner_cutoff=npar[1], vdw_cutoff=npar[2], vdw_inner_cutoff=npar[3],
                         vdw14_scaling=npar[4], elec14_scaling=npar[5], dielectric_solute=npar[6],
                         elec_mode=4, vdw_mode=2, padding=g["param_dict"]["padding"])
    s = osys(g)
    rs, ent = O.nlist_build(s, g["xyz"], max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e, f = O.nb_calc_forces(s, g["xyz"], rs, ent, p)
    assert rel(e["epot"], g["nb_epot"]) < 1e-12
    assert rel(e["virial"], g["nb_virial"]) < 1e-12
    assert rms_rel(f, g["nb_force"]) < 1e-13


@pytest.mark.parametrize("name", [c for c in CASES if "sd_xyz" in np.load(G.GOLDEN_DIR + "/" + c + ".npz").files])
def test_sd_minimiser(name):
    g = G.load(name)
    p = O.default_params(**g["param_dict"])
    rc, xyz, trace, ef, ff = O.minimise_sd(osys(g), g["xyz"], p, int(g["sd_steps"]), g["sd_stepsize"], 1)
    assert rc == int(g["sd_done"])
    assert np.max(np.abs(xyz - g["sd_xyz"])) < 1e-9
    assert rel(ef["epot"], g["sd_epot"]) < 1e-10
