"""Tiles split over several wavefronts (pdgpu_options.tile_split): every
segment count meets the golden tolerances of the unsplit path, in both
precisions, on the list path, the dense off-diagonal path and the energy
path, for GB, FF_NonBonded and FF_SoftVDW; results are deterministic and
the same alone and in a batch for a given split.

Tolerances as tests/test_gpu_parity.py: FP64 energy relative error <= 1e-10,
force RMS <= 1e-9; mixed energy <= 1e-5 relative, force RMS <= 1e-4.
"""
import numpy as np
import pytest

import golden_io as G
import pdgpu
from gpu_helpers import MIXED_FORCE_RMS, context_for, mixed_energy_bound, rel, rms_rel

pytestmark = pytest.mark.gpu
CASES = G.case_names()
SPLITS = [2, 4, 16]
E_KEYS = ("epot", "epot_vdw", "epot_elec", "epot_pol", "epot_pol_self", "epot_pol_cross")


def _tol(precision):
    return (1e-10, 1e-9) if precision == pdgpu.FP64 else (1e-5, 1e-4)


@pytest.mark.parametrize("split", SPLITS)
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_split_force_and_energy_paths(name, precision, split):
    g = G.load(name)
    et, ft = _tol(precision)
    print(f"split {split} case {name} precision {precision}", flush=True)
    with context_for(g, precision, tile_split=split) as ctx:
        i = ctx.add_system(g)
        e, f = ctx.calc_forces(i)
        assert e["status"] == 0
        if precision == pdgpu.FP64:
            for k in E_KEYS:
                assert rel(e[k], g[k]) <= et, (k, e[k], g[k])
        else:
            assert rel(e["epot"], g["epot"]) <= et
        assert rms_rel(f, g["force"]) <= ft
        assert e["pairs"] > 0
        born = ctx.born_radii(i)
        assert np.max(np.abs(born - g["born"]) / np.abs(g["born"])) <= (1e-10 if precision == pdgpu.FP64 else 1e-5)
        e, f = ctx.calc_energies(i)
        assert rel(e["epot"], g["en_epot"]) <= et
        assert rms_rel(f, g["en_force"]) <= ft


@pytest.mark.parametrize("split", [4, 16])
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_split_dense_offdiagonal(precision, split):
    g = G.load("res60_seed4")
    et, ft = _tol(precision)
    with context_for(g, precision, tile_split=split, dense_tile_pairs=1) as ctx:
        i = ctx.add_system(g)
        e, f = ctx.calc_forces(i)
        assert rel(e["epot"], g["epot"]) <= et
        assert rms_rel(f, g["force"]) <= ft


@pytest.mark.parametrize("split", [1, 8])
def test_split_pair_counts_equal_unsplit(split):
    """Segments partition each tile: the pairs the passes count are the
    unsplit path's, in both precisions."""
    g = G.load("res60_seed4")
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g)
        ref, _ = ctx.calc_forces(i)
    for prec in (pdgpu.FP64, pdgpu.MIXED):
        with context_for(g, prec, tile_split=split) as ctx:
            i = ctx.add_system(g)
            e, _ = ctx.calc_forces(i)
            assert e["pairs"] == ref["pairs"] and e["born_pairs"] == ref["born_pairs"]


def test_split_mixed_deterministic_and_batch_equals_alone():
    names = ["c1_seed1", "res60_seed4", "tri_ala"]
    gs = [G.load(n) for n in names]
    alone = []
    for g in gs:
        with context_for(g, pdgpu.MIXED, tile_split=8) as ctx:
            i = ctx.add_system(g)
            e1, f1 = ctx.calc_forces(i)
            e2, f2 = ctx.calc_forces(i)
            assert e1["epot"] == e2["epot"] and np.array_equal(f1, f2)
            alone.append((e1, f1))
    with context_for(gs[0], pdgpu.MIXED, tile_split=8) as ctx:
        ids = [ctx.add_system(g) for g in gs]
        ctx.nlist_build()
        res = ctx.calc_forces(pdgpu.PDGPU_ALL)
        for (ea, fa), i in zip(alone, ids):
            assert res[i]["epot"] == ea["epot"]
            assert np.array_equal(ctx.forces(i), fa)


@pytest.mark.parametrize("split", [4, 16])
@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_split_nonbonded_alone(precision, split):
    """FF_NonBonded alone (k_pair MODE 2) against the reference's nb_* runs."""
    from test_gpu_nonbonded import nb_options
    g = G.load("res60_seed4")
    with pdgpu.Context(nb_options(g, precision, tile_split=split)) as ctx:
        i = ctx.add_system(g)
        e, f = ctx.calc_forces(i)
    if precision == pdgpu.FP64:
        assert rel(e["epot"], g["nb_epot"]) <= 1e-10
        assert rel(e["virial"], g["nb_virial"]) <= 1e-10
        assert rms_rel(f, g["nb_force"].reshape(-1, 3)) <= 1e-9
    else:
        assert abs(e["epot"] - g["nb_epot"]) <= mixed_energy_bound(float(g["nb_epot"]), e["pairs"])
        assert rms_rel(f, g["nb_force"].reshape(-1, 3)) <= MIXED_FORCE_RMS


@pytest.mark.parametrize("split", [4])
def test_split_softvdw(split):
    """FF_SoftVDW alone (k_soft) split against the unsplit evaluation."""
    g = G.load("res60_seed4")
    res = []
    for sp in (1, split):
        with context_for(g, pdgpu.FP64, components=pdgpu.FF_SOFTVDW, tile_split=sp) as ctx:
            i = ctx.add_system(g)
            res.append(ctx.calc_forces(i))
    (ref, fref), (e, f) = res
    assert ref["epot"] != 0.0
    assert rel(e["epot"], ref["epot"]) <= 1e-12
    assert rms_rel(f, fref) <= 1e-12

