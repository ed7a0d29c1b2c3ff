"""PD's Langevin and Beeman MolecularDynamics on the device
(pdgpu_md_langevin_pd, pdgpu_md_beeman_pd) through the C ABI.

Parity: against the compiled reference's own runs (golden lg_* / bm_*,
FF_Bonded + GB, list refreshed every 10 steps) from the same positions and
velocities; the Langevin run is fed the reference's libc rand() deviates
(the oracle reproduces that stream). FP64: positions within 1e-8 A;
mixed: within 1e-3 A (FP32 pair maths through 40-60 steps).

The device noise (Philox4x32-10, Box-Muller) has no reference
counterpart: it is checked for what the mathematics fixes -- an ideal gas
thermalises to equipartition at the target temperature, the same seed and
context layout reproduce a run exactly, and the stream advances between
calls."""
import numpy as np
import pytest

import golden_io as G
import oracle_lib as O
import pdgpu
from gpu_helpers import context_for, rel
from pdgpu import synth

pytestmark = pytest.mark.gpu
CASES = G.case_names()
LG_CASES = [c for c in CASES if "lg_xyz" in G.load(c)]
BM_CASES = [c for c in CASES if "bm_xyz" in G.load(c)]
TOL = {pdgpu.FP64: (1e-8, 1e-7, 1e-9), pdgpu.MIXED: (1e-3, 1e-2, 1e-4)}


@pytest.mark.parametrize("tag", ["lg", "lx", "lh"])
@pytest.mark.parametrize("name", LG_CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_pd_langevin_matches_reference(name, precision, tag):
    """lg: PD's default friction (Taylor coefficients); lx: 100/ps
    (exponential coefficients); lh: LangevinOnHydrogens = false, hydrogens
    flagged as PD flags them."""
    g = G.load(name)
    steps = int(g["md_steps"])
    tx, tv, te = TOL[precision]
    hyd = g["is_h"] if tag == "lh" else None
    noise = O.langevin_noise(int(g[tag + "_seed"]), len(g["mass"]), steps, skip=hyd)
    with context_for(g, precision, langevin_on_hydrogens=0 if tag == "lh" else 1) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        if hyd is not None:
            ctx.set_hydrogens(i, hyd)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        e, ek = ctx.md_langevin_pd(i, steps, float(g["md_dt"]), float(g[tag + "_temp"]), float(g[tag + "_fric"]),
                                   noise=noise, update_nlist=int(g["md_update_nlist"]))
        assert e["status"] == 0
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
    assert np.abs(x - g[tag + "_xyz"]).max() <= tx
    assert np.abs(v - g[tag + "_v"]).max() <= tv * np.abs(g[tag + "_v"]).max()
    assert rel(e["epot"], g[tag + "_epot"]) <= te
    assert rel(ek, g[tag + "_ekin"]) <= te


@pytest.mark.parametrize("name", BM_CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_pd_beeman_matches_reference(name, precision):
    g = G.load(name)
    steps = int(g["md_steps"])
    tx, tv, te = TOL[precision]
    with context_for(g, precision) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        e, ek = ctx.md_beeman_pd(i, steps, float(g["md_dt"]), int(g["md_update_nlist"]))
        assert e["status"] == 0
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
    assert np.abs(x - g["bm_xyz"]).max() <= tx
    assert np.abs(v - g["bm_v"]).max() <= tv * np.abs(g["bm_v"]).max()
    assert rel(e["epot"], g["bm_epot"]) <= te
    assert rel(ek, g["bm_ekin"]) <= te


def test_beeman_keeps_old_forces_between_calls():
    """Two calls of 12 and 9 steps against the oracle with the workspace's
    old forces carried over (each call restarts PD's step count)."""
    g = G.load("tri_ala")
    s = O.OSystem(g["charge"], g["radius"], g["epsilon"], g["gb_radius"], g["vj"], g["terms123"], g["mass"],
                  g["mid"], g["band"], g["sp_i"], g["sp_j"], g["sp_o"], bonded=O.OBonded(g))
    p = O.default_params(**g["param_dict"])
    dt = float(g["md_dt"])
    _, x1, v1, fo, _, _ = O.md_beeman_pd(s, g["md_xyz0"], g["md_v0"], p, 12, dt)
    _, x2, v2, fo, _, _ = O.md_beeman_pd(s, x1, v1, p, 9, dt, f_old=fo)
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        ctx.md_beeman_pd(i, 12, dt)
        ctx.md_beeman_pd(i, 9, dt)
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
    assert np.abs(x - x2).max() <= 1e-9
    assert np.abs(v - v2).max() <= 1e-8 * np.abs(v2).max()


def test_langevin_small_friction_is_pd_velocity_verlet():
    """Below gamma dt = 1e-5 PD runs its VelocityVerlet update. Mixed
    precision: its fixed-point sums make runs bit-reproducible (the FP64
    path's floating-point atomics do not)."""
    g = G.load("tri_ala")
    dt = float(g["md_dt"])
    out = []
    for fric in (None, 1e9):
        with context_for(g, pdgpu.MIXED) as ctx:
            i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
            ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
            if fric is None:
                ctx.md_vv_pd(i, 15, dt)
            else:
                ctx.md_langevin_pd(i, 15, dt, 300.0, fric)  # gamma dt = 1e-6
            out.append(ctx.get_positions(i))
    assert np.array_equal(out[0], out[1])


def gas(n, spacing=16.0, seed=3):
    """n uncharged carbon atoms on a cubic grid, no bonds: every pair is
    beyond the cutoff, so the forces vanish and Langevin dynamics is an
    Ornstein-Uhlenbeck process per velocity component."""
    k = int(np.ceil(n ** (1 / 3)))
    grid = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n] * spacing
    return synth._finish(grid.astype(np.float64), ["C"] * n, np.zeros(n), np.arange(n, dtype=np.int32),
                         np.zeros((0, 2), np.int32))


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_device_noise_thermalises_an_ideal_gas(precision):
    """From rest, 30 steps at gamma dt = 0.5 (exponential coefficients, c0^60
    ~ 1e-13): each velocity component must be N(0, kB T / m). 3,000 atoms
    give 9,000 samples: the mean kinetic energy per atom is 3/2 kB T within
    a few percent, the components have zero mean and Gaussian kurtosis."""
    n, T = 3000, 300.0
    m = gas(n)
    opts = pdgpu.default_options(precision=precision, cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0,
                                 vdw_inner_cutoff=7.0, nlist_padding=1.0)
    with pdgpu.Context(opts) as ctx:
        i = ctx.add_system(m)
        e, ek = ctx.md_langevin_pd(i, 31, 1e-15, T, 5e14, seed=42)
        v = ctx.get_velocities(i)
        x = ctx.get_positions(i)
    assert e["pairs"] == 0
    kB = 1.38065e-23
    s = v / np.sqrt(kB * T / m["mass"])[:, None]  # standard normal if thermalised
    assert abs(s.mean()) < 4 / np.sqrt(s.size)
    assert abs((s ** 2).mean() - 1) < 0.05
    k4 = (s ** 4).mean() / (s ** 2).mean() ** 2
    assert abs(k4 - 3) < 0.25
    assert abs(ek / (1.5 * n * kB * T) - 1) < 0.05
    # the positions moved by a random walk, not a drift
    d = x - m["xyz"]
    assert np.abs(d.mean(0)).max() < 5 * d.std() / np.sqrt(n)


def test_device_noise_reproducible_and_advancing():
    m = gas(200)
    opts = pdgpu.default_options(precision=pdgpu.FP64, cutoff=10.0, inner_cutoff=7.0, vdw_cutoff=10.0,
                                 vdw_inner_cutoff=7.0, nlist_padding=1.0)
    runs = []
    for seed in (7, 7, 8):
        with pdgpu.Context(opts) as ctx:
            i = ctx.add_system(m)
            ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=seed)
            a = ctx.get_velocities(i)
            ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=seed)
            b = ctx.get_velocities(i)
            runs.append((a, b))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert not np.array_equal(runs[0][0], runs[2][0])
    # the second call drew new deviates: its velocity change is not the first's
    with pdgpu.Context(opts) as ctx:
        i = ctx.add_system(m)
        ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=7)
        ctx.set_velocities(i, np.zeros((200, 3)))
        ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=7)
        c = ctx.get_velocities(i)
    with pdgpu.Context(opts) as ctx:
        i = ctx.add_system(m)
        ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=7)
        d0 = ctx.get_velocities(i)
    assert not np.allclose(c, d0)


def test_langevin_rejects_bad_noise_shape():
    g = G.load("tri_ala")
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, bonded=True)
        with pytest.raises(ValueError):
            ctx.md_langevin_pd(i, 4, 1e-15, noise=np.zeros((3, len(g["mass"]), 6)))


@pytest.mark.parametrize("tag", ["vb", "bb"])
@pytest.mark.parametrize("name", [c for c in CASES if "vb_xyz" in G.load(c)])
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_pd_berendsen_matches_reference(name, precision, tag):
    """PD's VelocityVerlet (vb) and Beeman (bb) runs with its Berendsen
    thermostat (300 K, BerendsenTau 100 fs)."""
    g = G.load(name)
    steps = int(g["md_steps"])
    tx, tv, te = TOL[precision]
    with context_for(g, precision) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        ctx.set_thermostat(float(g[tag + "_tau"]), 300.0)
        run = ctx.md_vv_pd if tag == "vb" else ctx.md_beeman_pd
        e, ek = run(i, steps, float(g["md_dt"]), int(g["md_update_nlist"]))
        assert e["status"] == 0
        x = ctx.get_positions(i)
        v = ctx.get_velocities(i)
        # PD's momentum removal at step 200 is not built: longer runs refuse
        with pytest.raises(pdgpu.PdgpuError):
            run(i, 201, float(g["md_dt"]))
    assert np.abs(x - g[tag + "_xyz"]).max() <= tx
    assert np.abs(v - g[tag + "_v"]).max() <= tv * np.abs(g[tag + "_v"]).max()
    assert rel(e["epot"], g[tag + "_epot"]) <= te
    assert rel(ek, g[tag + "_ekin"]) <= te


@pytest.mark.parametrize("run_name", ["md_vv_pd", "md_beeman_pd"])
def test_pd_md_graph_replay_follows_thermostat(run_name):
    """Mixed precision (fixed-point sums, so runs are bit-reproducible):
    with use_graphs 0 and 1, the same sequence of calls -- thermostat off,
    on, then at a changed tau and temperature -- gives identical bits. A
    replayed step captured under other thermostat settings would not."""
    g = G.load("tri_ala")
    dt = float(g["md_dt"])
    out = []
    for graphs in (0, 1):
        states = []
        with context_for(g, pdgpu.MIXED, use_graphs=graphs) as ctx:
            i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
            ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
            run = getattr(ctx, run_name)
            for tau, temp in ((0.0, 300.0), (1e-13, 300.0), (5e-14, 350.0), (0.0, 300.0)):
                ctx.set_thermostat(tau, temp)
                run(i, 12, dt, 5)
                states.append((ctx.get_positions(i), ctx.get_velocities(i)))
        out.append(states)
    for (xa, va), (xb, vb) in zip(*out):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    # the thermostat changed the trajectory
    assert not np.array_equal(out[0][0][1], out[0][1][1])


def test_langevin_with_thermostat_refuses_past_momentum_removal():
    """PD zeroes the total momenta at step 200 whenever a thermostat is set,
    for every integrator (md.cpp:246-258); that is not built, so Langevin
    runs beyond 200 steps with tau > 0 are refused like the VV and Beeman
    calls."""
    g = G.load("tri_ala")
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_thermostat(1e-13, 300.0)
        with pytest.raises(pdgpu.PdgpuError):
            ctx.md_langevin_pd(i, 201, 1e-15, 300.0, 1e13, seed=1)
        ctx.md_langevin_pd(i, 5, 1e-15, 300.0, 1e13, seed=1)
        ctx.set_thermostat(0.0, 300.0)
        ctx.md_langevin_pd(i, 201, 1e-15, 300.0, 1e13, seed=1)
