"""FF_SoftVDW (calcForces_Soft, forcefields/ffsoftvdw.cpp:71-145) in the
oracle: against the compiled reference's dumps (PD's defaults, built
geometries of every fixture case) and against closed forms."""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O
from test_oracle_bonded import _system

CASES = G.case_names()
J = 4184.0 / 6.0221418e23  # J per molecule per kcal/mol


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_softvdw_matches_reference(name):
    d = G.load(name)
    s = _system(d, None)
    rs, ent = O.nlist_build(s, d["xyz"], 15.0, float(d["param_dict"]["padding"]), True)
    e, f = O.softvdw_forces(s, d["xyz"], rs, ent)
    assert rel(e["epot"], d["sv_epot"]) <= 1e-12
    assert rel(e["epot_vdw"], d["sv_vdw"]) <= 1e-12
    fr = d["sv_force"].reshape(-1, 3)
    assert np.sqrt(np.mean((f - fr) ** 2)) <= 1e-12 * np.sqrt(np.mean(fr ** 2))


def _pair_system(r, R=(1.9, 1.7), order=7):
    """Two atoms a distance r apart with LJ radii R; order is their bond
    order (7: unbonded)."""
    n = 2
    m = {"charge": np.zeros(n), "radius": np.array(R), "epsilon": np.zeros(n), "gb_radius": np.full(n, 1.5),
         "vj": np.zeros(n), "terms123": np.full(n, -1.0), "mass": np.full(n, 2e-26)}
    if order < 7:
        mid, band = 2, np.array([order, 0], np.int8)  # mid * (n - 1) band entries
    else:
        mid, band = 1, np.zeros(0, np.int8)
    s = O.OSystem(m["charge"], m["radius"], m["epsilon"], m["gb_radius"], m["vj"], m["terms123"], m["mass"],
                  mid, band, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    rs, ent = O.nlist_build(s, x, 15.0, 0.5, True)
    return s, x, rs, ent


@pytest.mark.parametrize("r", [0.9, 2.0, 2.9])
def test_two_atom_closed_form_and_gradient(r):
    s, x, rs, ent = _pair_system(r)
    e, f = O.softvdw_forces(s, x, rs, ent)
    sig = 1.9 + 1.7 - 0.6
    want = 12.0 * (1.0 - (r / sig) ** 2) ** 2
    assert rel(e["epot"] / J, want) <= 1e-13
    # the force on atom 1 is -dE/dr (N; E in J, r in A)
    h = 1e-6
    ep, _ = O.softvdw_forces(s, x + [[0, 0, 0], [h, 0, 0]], rs, ent)
    em, _ = O.softvdw_forces(s, x - [[0, 0, 0], [h, 0, 0]], rs, ent)
    dEdr = (ep["epot"] - em["epot"]) / (2 * h * 1e-10)
    assert rel(f[1, 0], -dEdr) <= 1e-6
    assert f[0, 0] == -f[1, 0] and f[0, 1] == 0 and f[1, 2] == 0


def test_range_and_bond_order_gates():
    # beyond s = R_i + R_j - OlapOffset: nothing
    s, x, rs, ent = _pair_system(3.01)
    e, f = O.softvdw_forces(s, x, rs, ent)
    assert e["epot"] == 0 and np.all(f == 0)
    # 1-4 pairs (bond order 3) and closer are excluded; 1-5 (4) counts
    for order, on in ((1, False), (2, False), (3, False), (4, True)):
        s, x, rs, ent = _pair_system(1.5, order=order)
        e, _ = O.softvdw_forces(s, x, rs, ent)
        assert (e["epot"] > 0) == on, order
