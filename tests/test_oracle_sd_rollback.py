"""The oracle's steepest descent on an uphill step (Minimisation::
doSteepestDescentStep, protocols/minimise.cpp:181-235), pinned by what the
algorithm fixes in closed form. A run returns the last accepted positions
(the final move is not kept). Step 0 moves every atom by StepSize * f0;
when the energy there is higher, step 1 restores the saved positions and
returns with the multiplier scaled by 0.5 / 1.2; step 2 is then always
accepted (the reference energy was raised by 1 J) from the restored
positions, scales the multiplier by 1.2 and moves by
StepSize * (0.5 / 1.2 * 1.2) * f0; a downhill step 3 accepts that point.
A downhill first step instead keeps moving. A wrong multiplier, a missing
restore or a rollback that moves the atoms fails one of these."""
import numpy as np

import golden_io as G
import oracle_lib as O


def _setup(name):
    g = G.load(name)
    p = O.default_params(**g["param_dict"])
    s = O.OSystem(g["charge"], g["radius"], g["epsilon"], g["gb_radius"], g["vj"], g["terms123"], g["mass"],
                  g["mid"], g["band"], g["sp_i"], g["sp_j"], g["sp_o"])
    return g, p, s


def test_uphill_step_is_rolled_back():
    g, p, s = _setup("c1_seed1")
    x0 = g["xyz"]
    S = 2e5 * 250  # step 0 goes uphill, the 0.5 x step downhill
    rc, x2, trace, _, _ = O.minimise_sd(s, x0, p, 4, S, 1)
    assert trace[1] > trace[0]  # uphill: rolled back
    assert trace[2] == trace[0]  # step 2 evaluates the restored positions
    assert trace[3] <= trace[2]  # the shorter step is accepted
    rs, ent = O.nlist_build(s, x0, p.cutoff, p.padding, True)
    _, f0, _, _ = O.gb_calc_forces(s, x0, rs, ent, p)
    m = 1.0
    m *= 0.5 / 1.2
    m *= 1.2
    assert np.max(np.abs(x2 - (x0 + f0 * (S * m)))) <= 1e-12
    # three steps end on the restored start: the last accepted positions
    rc, x3, _, _, _ = O.minimise_sd(s, x0, p, 3, S, 1)
    assert np.array_equal(x3, x0)


def test_downhill_steps_keep_moving():
    g, p, s = _setup("c1_seed1")
    x0 = g["xyz"]
    rc, x3, trace, _, _ = O.minimise_sd(s, x0, p, 3, 2e5, 1)
    assert trace[1] <= trace[0] and trace[2] <= trace[1]
    rs, ent = O.nlist_build(s, x0, p.cutoff, p.padding, True)
    _, f0, _, _ = O.gb_calc_forces(s, x0, rs, ent, p)
    x1 = x0 + f0 * 2e5
    rs1, ent1 = O.nlist_build(s, x1, p.cutoff, p.padding, True)
    _, f1, _, _ = O.gb_calc_forces(s, x1, rs1, ent1, p)
    # three downhill steps return the point of step 2: x1 + 1.2 StepSize f1
    assert np.max(np.abs(x3 - (x1 + f1 * (2e5 * 1.2)))) <= 1e-12
