"""Pins of the oracle's textbook velocity Verlet (orc_md_vv: half kick,
drift, force, half kick) against what the mathematics fixes. PD has no
such integrator (its VelocityVerlet, protocols/md.cpp:788-838, is pinned by
the reference's own runs through orc_md_vv_pd), so these closed forms are
its only pins:

- free particles: x_n = x_0 + n v dt exactly, v unchanged;
- a harmonic dimer (one FF_Bonded bond, E = k (d - l)^2, no charges, no
  LJ): velocity Verlet is the linear map M = [[c, h], [-h w^2 (1 - h^2 w^2
  / 4), c]] on (d - l, v_rel) with c = 1 - h^2 w^2 / 2, w^2 = 2 k / mu, so
  the stretch follows d_n - l = cos(n t) (d_0 - l) + sin(n t) / sin(t) h
  v_0 with cos t = c, and v^2 + w^2 (1 - h^2 w^2 / 4) (d - l)^2 is
  conserved exactly;
- the energy error of that dimer over a fixed time scales as dt^2.

A drift with the full step's kick (PD's variant) or a missing half kick
fails the closed form by a relative O(h w) or more."""
import numpy as np
import pytest

import oracle_lib as O

AMU = 1.66054e-27
ANG = 1e-10
KCAL_J = 4184.0 / 6.02214076e23  # J/molecule per kcal/mol


def dimer(m1=12.011, m2=15.999, k_kcal=300.0, l=1.229, bonded=True):
    """Two uncharged, LJ-free atoms with one bond of k kcal/mol/A^2."""
    n = 2
    z = np.zeros(n)
    d = {"bd_idx": np.array([0, 1]) if bonded else np.zeros(0), "bd_lk": np.array([l, k_kcal * KCAL_J])
         if bonded else np.zeros(0)}
    for key in ("an_idx", "an_tk", "to_idx", "to_terms", "to_vn", "to_n", "to_gamma", "im_idx", "im_terms",
                "im_vn", "im_n", "im_gamma"):
        d[key] = np.zeros(0)
    b = O.OBonded(d)
    mass = np.array([m1, m2]) * AMU
    s = O.OSystem(z, np.full(n, 1.7), z, np.full(n, 1.6), z, np.full(n, -0.5), mass, 1,
                  np.array([1 if bonded else 7], np.int8), bonded=b)
    return s, mass


def run(s, xyz, vel, steps, dt):
    p = O.default_params()
    rc, x, v, te, tk = O.md_vv(s, xyz, vel, p, steps, dt, update_nlist=10)
    assert rc == steps
    return x, v, te, tk


def test_free_particles_move_in_straight_lines():
    s, _ = dimer(bonded=False)
    xyz = np.array([[0.0, 0.0, 0.0], [40.0, 0.0, 0.0]])  # beyond every cutoff
    vel = np.array([[310.0, -120.0, 55.0], [-40.0, 250.0, -180.0]])
    dt, steps = 2e-15, 37
    x, v, te, _ = run(s, xyz, vel, steps, dt)
    np.testing.assert_allclose(x, xyz + steps * vel * dt / ANG, rtol=0, atol=1e-12)
    assert np.array_equal(v, vel)
    assert np.all(te == 0.0)


@pytest.mark.parametrize("v0", [0.0, 900.0])
@pytest.mark.parametrize("dt", [0.5e-15, 1e-15, 2e-15])
def test_harmonic_dimer_follows_the_verlet_recurrence(dt, v0):
    k_kcal, l, stretch = 300.0, 1.229, 0.08
    s, mass = dimer(k_kcal=k_kcal, l=l)
    mu = mass[0] * mass[1] / mass.sum()
    w2 = 2.0 * k_kcal * KCAL_J / ANG ** 2 / mu  # 1/s^2
    xyz = np.array([[0.0, 0.0, 0.0], [l + stretch, 0.0, 0.0]])
    # relative velocity v0 along the bond, total momentum zero
    vel = np.array([[-v0 * mass[1] / mass.sum(), 0, 0], [v0 * mass[0] / mass.sum(), 0, 0]])
    steps = 150
    h = dt
    c = 1.0 - 0.5 * h * h * w2
    t = np.arccos(c)
    n = np.arange(1, steps + 1)
    d_exact = l + np.cos(n * t) * stretch + np.sin(n * t) / np.sin(t) * h * v0 / ANG
    d_got = np.empty(steps)
    inv_got = np.empty(steps)
    x, v = xyz, vel
    for i in range(steps):  # one step per call: the trajectory at every step
        x, v, _, _ = run(s, x, v, 1, dt)
        d_got[i] = x[1, 0] - x[0, 0]
        vr = v[1, 0] - v[0, 0]
        inv_got[i] = vr ** 2 + w2 * (1 - 0.25 * h * h * w2) * ((d_got[i] - l) * ANG) ** 2
        assert abs(x[:, 1:]).max() == 0.0 and abs(v[:, 1:]).max() == 0.0
    assert np.abs(d_got - d_exact).max() <= 1e-11 * stretch + 1e-13
    inv0 = v0 ** 2 + w2 * (1 - 0.25 * h * h * w2) * (stretch * ANG) ** 2
    assert np.abs(inv_got / inv0 - 1).max() <= 1e-10
    # the centre of mass stays where it was
    com = (mass[:, None] * x).sum(0) / mass.sum()
    com0 = (mass[:, None] * xyz).sum(0) / mass.sum()
    assert np.abs(com - com0).max() <= 1e-12


def test_harmonic_dimer_energy_error_scales_as_dt_squared():
    """Over the same 150 fs, halving dt divides the largest deviation of the
    total energy from its start by 4 (to leading order in h w)."""
    k_kcal, l, stretch = 300.0, 1.229, 0.08
    s, mass = dimer(k_kcal=k_kcal, l=l)
    xyz = np.array([[0.0, 0.0, 0.0], [l + stretch, 0.0, 0.0]])
    vel = np.zeros((2, 3))
    e0 = k_kcal * KCAL_J * stretch ** 2
    dev = []
    for dt, steps in ((1e-15, 150), (0.5e-15, 300)):
        _, _, te, tk = run(s, xyz, vel, steps, dt)
        dev.append(np.abs(te + tk - e0).max() / e0)
    assert 3.8 < dev[0] / dev[1] < 4.2
    # and the size itself is (h w)^2 / 8 of the energy at its peak
    mu = mass[0] * mass[1] / mass.sum()
    hw2 = (1e-15) ** 2 * 2.0 * k_kcal * KCAL_J / ANG ** 2 / mu
    assert 0.9 < dev[0] / (hw2 / 4) < 1.1
