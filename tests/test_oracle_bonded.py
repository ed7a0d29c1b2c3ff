"""The oracle's FF_Bonded restatement (orc_bonded_forces) pinned against
the compiled reference's FF_Bonded (golden term lists, energies and forces
at the built and at a perturbed geometry), finite differences of its own
energy, and closed forms for single terms, including the sign convention
of the dihedral (IUPAC) and the wrap of the harmonic form."""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O

CASES = G.case_names()
KJ = O.KCAL2J / O.NA  # J/molecule per kcal/mol


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("tag", ["bf", "bfp"])
def test_bonded_matches_reference(name, tag):
    d = G.load(name)
    x = d["xyz"] if tag == "bf" else d["bfp_xyz"]
    e, f = O.bonded_forces(O.OBonded(d), x)
    ref = d[tag + "_force"]
    assert np.sqrt(((f - ref) ** 2).mean() / (ref ** 2).mean()) <= 1e-12
    assert rel(e["epot_bond"], d[tag + "_bond"]) <= 1e-12
    assert rel(e["epot_angle"], d[tag + "_angle"]) <= 1e-12
    # getETorsion holds torsions only; the Hamiltonian's epot_torsion adds impropers
    assert rel(e["epot_torsion"] - e["epot_improper"], d[tag + "_torsion"]) <= 1e-10
    assert rel(e["epot_improper"], d[tag + "_improper"]) <= 1e-10
    assert rel(e["epot"], d[tag + "_epot"]) <= 1e-12


def test_term_counts_are_plausible():
    d = G.load("tri_ala")
    n = d["n"]
    nb, na = len(d["bd_lk"]) // 2, len(d["an_tk"]) // 2
    # tree-like peptide: bonds = atoms - 1; every bonded atom pair is a 1-2 link
    assert nb == n - 1 == len(d["a12"])
    assert na == len(d["a13"])
    assert len(d["to_terms"]) > 0 and len(d["im_terms"]) > 0
    assert d["to_terms"].min() >= 1 and d["to_terms"].max() <= 4


def _energy(b, x):
    return O.bonded_forces(b, x)[0]["epot"]


@pytest.mark.parametrize("tag", ["bf", "bfp"])
def test_forces_are_minus_gradient(tag):
    d = G.load("tri_ala")
    b = O.OBonded(d)
    x0 = (d["xyz"] if tag == "bf" else d["bfp_xyz"]).copy()
    _, f = O.bonded_forces(b, x0)
    h = 1e-6
    rng = np.random.default_rng(0)
    for a in rng.choice(d["n"], size=8, replace=False):
        for c in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[a, c] += h
            xm[a, c] -= h
            g = (_energy(b, xp) - _energy(b, xm)) / (2 * h * 1e-10)  # J/m
            assert abs(-g - f[a, c]) <= 1e-5 * np.abs(f).max() + 1e-6 * abs(g)


def _one(**terms):
    base = {k: np.zeros(0) for k in ("bd_idx", "bd_lk", "an_idx", "an_tk", "to_idx", "to_terms", "to_vn", "to_n",
                                     "to_gamma", "im_idx", "im_terms", "im_vn", "im_n", "im_gamma")}
    base.update(terms)
    return O.OBonded(base)


def test_bond_closed_form():
    k, l, d = 300 * KJ, 1.5, 1.62
    b = _one(bd_idx=[0, 1], bd_lk=[l, k])
    x = np.array([[0.0, 0, 0], [d, 0, 0]])
    e, f = O.bonded_forces(b, x)
    assert rel(e["epot_bond"], k * (d - l) ** 2) <= 1e-14
    assert rel(f[1, 0], -2 * k * (d - l) / 1e-10) <= 1e-14
    assert rel(e["virial"], 2 * k * (d - l) * d) <= 1e-14
    assert np.allclose(f.sum(0), 0, atol=1e-20)


def test_angle_closed_form():
    k, t0, th = 50 * KJ, np.deg2rad(109.5), np.deg2rad(121.0)
    b = _one(an_idx=[0, 1, 2], an_tk=[t0, k])
    x = np.array([[1.3, 0, 0], [0.0, 0, 0], [1.1 * np.cos(th), 1.1 * np.sin(th), 0]])
    e, f = O.bonded_forces(b, x)
    assert rel(e["epot_angle"], k * (th - t0) ** 2) <= 1e-13
    # torque-free and force-free
    assert np.abs(f.sum(0)).max() <= 1e-12 * np.abs(f).max()
    assert np.abs(np.cross(x, f).sum(0)).max() <= 1e-12 * np.abs(f).max()
    # |F_i| = |dE/dtheta| / r_i
    assert rel(np.linalg.norm(f[0]), 2 * k * (th - t0) / 1.3 / 1e-10) <= 1e-12


def _quartet(phi):
    """i-a-b-j with IUPAC dihedral phi: a at the origin, b on +z."""
    return np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 1.5], [np.cos(phi), np.sin(phi), 1.5]])


@pytest.mark.parametrize("phi_deg", [-150.0, -60.0, -3.0, 37.0, 95.0, 178.0])
def test_dihedral_sign_and_fourier_terms(phi_deg):
    phi = np.deg2rad(phi_deg)
    vn = np.array([1.0, 0.4, 0.25, 0.0]) * KJ
    nn = np.array([1.0, 2.0, 3.0, 0.0])
    gm = np.array([np.pi / 2, 0.3, np.pi, 0.0])
    b = _one(to_idx=[0, 1, 2, 3], to_terms=[3], to_vn=vn, to_n=nn, to_gamma=gm)
    e, f = O.bonded_forces(b, _quartet(phi))
    want = sum(vn[t] * (1 + np.cos(nn[t] * phi + gm[t])) for t in range(3))
    assert rel(e["epot_torsion"], want) <= 1e-12
    assert np.abs(f.sum(0)).max() <= 1e-10 * np.abs(f).max()
    # the same term as an improper lands in both epot_torsion and epot_improper
    bi = _one(im_idx=[0, 1, 2, 3], im_terms=[3], im_vn=vn, im_n=nn, im_gamma=gm)
    ei, fi = O.bonded_forces(bi, _quartet(phi))
    assert rel(ei["epot_improper"], want) <= 1e-12 and rel(ei["epot_torsion"], want) <= 1e-12
    assert np.array_equal(fi, f)


@pytest.mark.parametrize("phi_deg,gamma_deg", [(-175.0, 170.0), (175.0, -170.0), (10.0, 20.0)])
def test_harmonic_dihedral_wraps(phi_deg, gamma_deg):
    phi, g = np.deg2rad(phi_deg), np.deg2rad(gamma_deg)
    k = 2.0 * KJ
    b = _one(to_idx=[0, 1, 2, 3], to_terms=[1], to_vn=[k, 0, 0, 0], to_n=[0.0, 0, 0, 0], to_gamma=[g, 0, 0, 0])
    e, _ = O.bonded_forces(b, _quartet(phi))
    dd = (phi - g + np.pi) % (2 * np.pi) - np.pi
    assert rel(e["epot_torsion"], k * dd * dd) <= 1e-10


def test_dihedral_both_gradient_branches():
    # near-planar (|sin phi| <= 0.1: sine branch) and generic (cosine branch)
    for phi_deg in (179.0, -2.0, 60.0):
        x0 = _quartet(np.deg2rad(phi_deg))
        x0[0, 1] += 0.05
        b = _one(to_idx=[0, 1, 2, 3], to_terms=[2], to_vn=[KJ, 0.5 * KJ, 0, 0], to_n=[2.0, 3.0, 0, 0],
                 to_gamma=[np.pi, 0.2, 0, 0])
        _, f = O.bonded_forces(b, x0)
        h = 1e-7
        for a in range(4):
            for c in range(3):
                xp, xm = x0.copy(), x0.copy()
                xp[a, c] += h
                xm[a, c] -= h
                g = (_energy(b, xp) - _energy(b, xm)) / (2 * h * 1e-10)
                assert abs(-g - f[a, c]) <= 1e-6 * np.abs(f).max()


def _system(d, bonded):
    return O.OSystem(d["charge"], d["radius"], d["epsilon"], d["gb_radius"], d["vj"], d["terms123"], d["mass"],
                     d["mid"], d["band"], d["sp_i"], d["sp_j"], d["sp_o"], bonded=bonded)


SD_CASES = [c for c in CASES if "sdf_xyz" in np.load(G.os.path.join(G.GOLDEN_DIR, c + ".npz")).files]


@pytest.mark.parametrize("name", SD_CASES)
def test_full_forcefield_single_point(name):
    """FF_Bonded + FF_GeneralizedBorn_Still in one Forcefield."""
    d = G.load(name)
    b = O.OBonded(d)
    s = _system(d, b)
    p = O.default_params(**d["param_dict"])
    rs, ent = O.nlist_build(s, d["xyz"], max(p.cutoff, p.vdw_cutoff), p.padding, True)
    e, f, _, _ = O.gb_calc_forces(s, d["xyz"], rs, ent, p)
    eb, fb = O.bonded_forces(b, d["xyz"])
    assert rel(e["epot"] + eb["epot"], d["ff_epot"]) <= 1e-12
    assert rel(eb["epot_torsion"], d["ff_torsion"]) <= 1e-12
    ft = f + fb
    assert np.sqrt(((ft - d["ff_force"]) ** 2).mean() / (d["ff_force"] ** 2).mean()) <= 1e-12


@pytest.mark.parametrize("name", SD_CASES)
def test_full_forcefield_steepest_descent(name):
    d = G.load(name)
    s = _system(d, O.OBonded(d))
    p = O.default_params(**d["param_dict"])
    steps = int(d["sd_steps"])
    rc, x, _, ef, _ = O.minimise_sd(s, d["xyz"], p, steps, d["sd_stepsize"], 1)
    assert rc == steps == int(d["sdf_done"])
    assert rel(ef["epot"], d["sdf_epot"]) <= 1e-10
    assert np.abs(x - d["sdf_xyz"]).max() <= 1e-9


MD_CASES = [n for n in CASES if "md_xyz" in G.load(n)]


@pytest.mark.parametrize("name", MD_CASES)
def test_pd_velocity_verlet_matches_reference(name):
    """PD's MolecularDynamics with its VelocityVerlet update as written
    (protocols/md.cpp:788-838), FF_Bonded + GB, against the compiled
    reference's run from the same positions and velocities."""
    d = G.load(name)
    s = _system(d, O.OBonded(d))
    p = O.default_params(**d["param_dict"])
    steps = int(d["md_steps"])
    rc, x, v, te, tk = O.md_vv_pd(s, d["md_xyz0"], d["md_v0"], p, steps, float(d["md_dt"]),
                                   int(d["md_update_nlist"]))
    assert rc == steps
    assert np.abs(x - d["md_xyz"].reshape(-1, 3)).max() <= 1e-9
    assert np.abs(v - d["md_v"].reshape(-1, 3)).max() <= 1e-9 * np.abs(d["md_v"]).max()
    assert rel(te[-1], d["md_epot"]) <= 1e-10
    assert rel(tk[-1], d["md_ekin"]) <= 1e-10
