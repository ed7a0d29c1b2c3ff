"""Pins of the oracle that do not depend on the compiled reference: closed
forms of the Still GB / LJ / Coulomb terms for two atoms, special cases,
invariances and finite differences."""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O

HALF_J = O.ECONV_J * 0.5


def two_atoms(r, q=(0.4, -0.3), R=(1.9, 1.7), eps_kcal=(0.1, 0.2), gbr=(1.7, 1.5), vj=(20.0, 15.0),
              t123=(-90.0, -100.0)):
    eps = np.sqrt(np.array(eps_kcal) * O.J_PER_KCALMOL)
    s = O.OSystem(q, R, eps, gbr, vj, t123, [1e-26, 1e-26], mid=1, band=np.array([7], np.int8))
    xyz = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    return s, xyz


def params(**kw):
    base = dict(cutoff=12.0, inner_cutoff=9.0, vdw_cutoff=12.0, vdw_inner_cutoff=9.0, dielectric_solvent=78.5,
                padding=2.0)
    base.update(kw)
    return O.default_params(**base)


def still_born(t_self, P4, vj_other, ccf, r):
    return -O.HALFECONV / (t_self + P4 * vj_other * ccf / r ** 4)


def test_two_atom_closed_form_near_branch():
    r = 4.0
    s, xyz = two_atoms(r)
    p = params()
    rs, ent = O.nlist_build(s, xyz, 12.0, 2.0, True)
    e, f, born, deda = O.gb_calc_forces(s, xyz, rs, ent, p)
    # Born radii: x = r^2/(R1+R2)^2 = 16/10.24 > 1/P5 so CCF = 1
    a1 = still_born(-90.0, 15.236, 15.0, 1.0, r)
    a2 = still_born(-100.0, 15.236, 20.0, 1.0, r)
    assert born == pytest.approx([a1, a2], rel=1e-14)
    premul = -HALF_J * (1.0 - 1.0 / 78.5)
    fgb = 1.0 / np.sqrt(r * r + a1 * a2 * np.exp(-r * r / (4 * a1 * a2)))
    e_pol = premul * (0.4 ** 2 / a1 + 0.3 ** 2 / a2 + 2 * 0.4 * -0.3 * fgb)
    e_elec = O.ECONV_J * 0.4 * -0.3 / r
    eps = np.sqrt(0.1 * 0.2) * O.J_PER_KCALMOL
    rmin = 1.9 + 1.7
    e_vdw = eps * ((rmin / r) ** 12 - 2 * (rmin / r) ** 6)
    assert e["epot_pol"] == pytest.approx(e_pol, rel=1e-13)
    assert e["epot_elec"] == pytest.approx(e_elec, rel=1e-13)
    assert e["epot_vdw"] == pytest.approx(e_vdw, rel=1e-13)
    assert e["epot"] == pytest.approx(e_pol + e_elec + e_vdw, rel=1e-13)
    assert np.allclose(f[0], -f[1], rtol=1e-13, atol=0)


def test_far_branch_uses_solvent_dielectric():
    # tiny Born radii push u = -r^2/(4 a1 a2) below -4: no GB pair term,
    # Coulomb with 1/eps_out instead (forcefields/gbff.cpp:1575-1661)
    r = 6.0
    s, xyz = two_atoms(r, t123=(-900.0, -900.0), vj=(0.0, 0.0))
    p = params()
    rs, ent = O.nlist_build(s, xyz, 12.0, 2.0, True)
    e, f, born, _ = O.gb_calc_forces(s, xyz, rs, ent, p)
    assert -r * r / (4 * born[0] * born[1]) < -4
    premul = -HALF_J * (1.0 - 1.0 / 78.5)
    assert e["epot_pol"] == pytest.approx(premul * (0.4 ** 2 / born[0] + 0.3 ** 2 / born[1]), rel=1e-13)
    assert e["epot_elec"] == pytest.approx(O.ECONV_J * 0.4 * -0.3 / (78.5 * r), rel=1e-13)


def test_switch_polynomial():
    # between InnerCutoff and Cutoff the Coulomb term carries S = (1 - ((r-ri)/(rc-ri))^2)^2
    r = 10.5
    s, xyz = two_atoms(r, t123=(-900.0, -900.0), vj=(0.0, 0.0))
    p = params()
    rs, ent = O.nlist_build(s, xyz, 12.0, 2.0, True)
    e, _, _, _ = O.gb_calc_forces(s, xyz, rs, ent, p)
    S = (1 - ((r - 9.0) / 3.0) ** 2) ** 2
    assert e["epot_elec"] == pytest.approx(S * O.ECONV_J * 0.4 * -0.3 / (78.5 * r), rel=1e-13)


def test_bonded_pair_excluded_from_lj_coulomb_born():
    # a 1-2 pair keeps only the GB pair term (bond order gates, Q4)
    r = 1.5
    s, xyz = two_atoms(r)
    s = O.OSystem(s.charge, s.radius, s.epsilon, s.gb_radius, s.vj, s.terms123, s.mass, mid=2,
                  band=np.array([1, 7], np.int8))
    p = params()
    rs, ent = O.nlist_build(s, xyz, 12.0, 2.0, True)
    assert (ent >> 24).tolist() == [1, 1]
    e, f, born, _ = O.gb_calc_forces(s, xyz, rs, ent, p)
    assert e["epot_vdw"] == 0.0 and e["epot_elec"] == 0.0
    assert born == pytest.approx([-O.HALFECONV / -90.0, -O.HALFECONV / -100.0], rel=1e-15)
    assert e["epot_pol_cross"] != 0.0


def test_p4_zero_gives_fixed_radii():
    g = G.load("c1_seed2")
    p = O.default_params(**dict(g["param_dict"], P4=0.0))
    s = O.OSystem(g["charge"], g["radius"], g["epsilon"], g["gb_radius"], g["vj"], g["terms123"], g["mass"],
                  g["mid"], g["band"])
    born = O.born_radii(s, g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    assert np.allclose(born, -O.HALFECONV / g["terms123"], rtol=1e-15)


def _sys(g):
    return O.OSystem(g["charge"], g["radius"], g["epsilon"], g["gb_radius"], g["vj"], g["terms123"], g["mass"],
                     g["mid"], g["band"], g["sp_i"], g["sp_j"], g["sp_o"])


def test_newton_third_law():
    g = G.load("res60_seed4")
    p = O.default_params(**g["param_dict"])
    e, f, _, _ = O.gb_calc_forces(_sys(g), g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).sum()


def test_rigid_motion_invariance():
    g = G.load("c1_seed3")
    p = O.default_params(**g["param_dict"])
    s = _sys(g)
    rs, ent = O.nlist_build(s, g["xyz"], max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e0, f0, _, _ = O.gb_calc_forces(s, g["xyz"], rs, ent, p)
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    x1 = g["xyz"] @ Rz.T + np.array([13.0, -7.0, 2.5])
    rs1, ent1 = O.nlist_build(s, x1, max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e1, f1, _, _ = O.gb_calc_forces(s, x1, rs1, ent1, p)
    assert e1["epot"] == pytest.approx(e0["epot"], rel=1e-11)
    assert np.allclose(f1, f0 @ Rz.T, rtol=0, atol=1e-10 * np.abs(f0).max())


def _fd_check(g, p, atoms, h=1e-5):
    s = _sys(g)
    rs, ent = g["nl_full_row_start"], g["nl_full_entries"]
    _, f, _, _ = O.gb_calc_forces(s, g["xyz"], rs, ent, p)
    errs = []
    for a in atoms:
        for c in range(3):
            xp = g["xyz"].copy(); xp[a, c] += h
            xm = g["xyz"].copy(); xm[a, c] -= h
            ep = O.gb_calc_forces(s, xp, rs, ent, p)[0]["epot"]
            em = O.gb_calc_forces(s, xm, rs, ent, p)[0]["epot"]
            grad = (ep - em) / (2 * h) * 1e10  # J/A -> N
            errs.append(abs(-grad - f[a, c]) / (abs(f[a, c]) + 1e-3 * np.abs(f).max()))
    return max(errs)


def test_forces_are_energy_gradient_with_fixed_radii():
    # P4 = 0 freezes the Born radii: the pair pass alone must be the exact gradient
    g = G.load("tri_ala")
    p = O.default_params(**dict(g["param_dict"], P4=0.0))
    assert _fd_check(g, p, range(0, 33, 4)) < 1e-6


def test_chain_rule_constant_quirk():
    # with Born radii moving, PD's chain rule uses 166.0 where the radii use
    # halfeconv (forcefields/gbff.cpp:1755 vs :463): forces are the gradient
    # only to ~2e-4 of the chain-rule term
    g = G.load("tri_ala")
    p = O.default_params(**g["param_dict"])
    err = _fd_check(g, p, range(0, 33, 4))
    assert 1e-7 < err < 5e-4


def test_force_path_equals_energy_path_without_far_branch():
    # with the u > -4 branch always taken, calcForces' energy and
    # calcEnergies' exact GB + FF_NonBonded energy are the same function
    g = G.load("c1_seed1")
    p = O.default_params(**dict(g["param_dict"], exp_approx_threshold=-1e300))
    s = _sys(g)
    ef, _, _, _ = O.gb_calc_forces(s, g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    ee, _, _ = O.gb_calc_energies(s, g["xyz"], g["nl_full_row_start"], g["nl_full_entries"], p)
    assert ef["epot"] == pytest.approx(ee["epot"], rel=1e-12)
    assert ef["epot_pol"] == pytest.approx(ee["epot_pol"], rel=1e-12)


def test_nlist_brute_force_small():
    rng = np.random.default_rng(3)
    n = 40
    xyz = rng.uniform(0, 15, size=(n, 3))
    band = np.full(3 * (n - 1), 7, np.int8)
    band[0::3] = 1
    s = O.OSystem(np.zeros(n), np.ones(n), np.ones(n), np.ones(n), np.ones(n), -np.ones(n), np.ones(n), 3, band)
    rs, ent = O.nlist_build(s, xyz, 6.0, 0.5, True)
    for i in range(n):
        row = ent[rs[i]:rs[i + 1]]
        js = [j for j in range(n) if j != i and ((xyz[j] - xyz[i]) ** 2).sum() <= 6.5 ** 2]
        assert (row & 0xFFFFFF).tolist() == js
        for e in row:
            j = int(e & 0xFFFFFF)
            expect = 1 if abs(i - j) == 1 else 7
            assert int(e >> 24) == expect
