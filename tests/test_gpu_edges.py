"""GPU edge cases against the oracle: instance sizes around the 64-atom
block boundary (1, 2, 3, 63, 64, 65, 128, 129 atoms), all of them in one
batch, an empty context, and a context emptied by removals."""
import numpy as np
import pytest

import pdgpu
from gpu_helpers import rel, rms_rel
from pdgpu import synth
from test_gpu_dynamics import opts, oracle_forces

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 3, 63, 64, 65, 128, 129)


def prefix(m, n):
    """The first n atoms of a peptide as a molecule of their own (bonds,
    1-3 / 1-4 pairs, bond orders and GB fixed terms rederived)."""
    b = np.asarray(m["bonds"])
    keep = (b[:, 0] < n) & (b[:, 1] < n)
    return synth._finish(m["xyz"][:n].copy(), list(m["el"][:n]), np.asarray(m["charge"][:n]).copy(),
                         np.asarray(m["resid"][:n]).copy(), b[keep])


@pytest.fixture(scope="module")
def pieces():
    base = synth.peptide(seed=404, residues=12)
    return {n: prefix(base, n) for n in SIZES}


def check(e, f, eo, fo, precision):
    assert e["status"] == 0
    assert rel(e["epot"], eo["epot"]) <= (1e-10 if precision == pdgpu.FP64 else 1e-5)
    if np.any(fo != 0):
        assert rms_rel(f, fo) <= (1e-9 if precision == pdgpu.FP64 else 1e-4)
    else:  # one atom: the GB self term only, no force
        assert np.all(f == 0)


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
@pytest.mark.parametrize("n", SIZES)
def test_block_boundary_sizes_match_oracle(pieces, n, precision):
    m = pieces[n]
    eo, fo = oracle_forces(m)
    with pdgpu.Context(opts(precision)) as ctx:
        i = ctx.add_system(m)
        e, f = ctx.calc_forces(i)
        check(e, f, eo, fo, precision)
        x = m["xyz"]
        d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
        # unordered pairs in the cutoff; FP32 distances may differ only for
        # pairs within rounding of the cutoff itself
        edge = int(np.count_nonzero(np.triu(np.abs(d2 - 144.0) < 1e-3, 1)))
        assert abs(e["pairs"] - int(np.count_nonzero(np.triu(d2 <= 144.0, 1)))) <= edge


@pytest.mark.parametrize("precision", [pdgpu.FP64, pdgpu.MIXED])
def test_boundary_sizes_in_one_batch(pieces, precision):
    ref = {n: oracle_forces(pieces[n]) for n in SIZES}
    with pdgpu.Context(opts(precision, max_instances=16)) as ctx:
        ids = {n: ctx.add_system(pieces[n]) for n in SIZES}
        res = ctx.calc_forces(pdgpu.PDGPU_ALL)
        for n, i in ids.items():
            check(res[i], ctx.forces(i), *ref[n], precision)


def test_empty_context():
    with pdgpu.Context(opts(pdgpu.MIXED, max_instances=4)) as ctx:
        ctx.nlist_build()
        assert ctx.calc_forces(pdgpu.PDGPU_ALL) == {}
        assert ctx.calc_energies(pdgpu.PDGPU_ALL) == {}


def test_context_emptied_by_removal_then_reused(pieces):
    m = pieces[65]
    eo, fo = oracle_forces(m)
    with pdgpu.Context(opts(pdgpu.FP64, max_instances=4)) as ctx:
        a = ctx.add_system(pieces[129])
        b = ctx.add_system(pieces[3])
        ctx.calc_forces(pdgpu.PDGPU_ALL)
        ctx.remove(a)
        ctx.remove(b)
        assert ctx.calc_forces(pdgpu.PDGPU_ALL) == {}
        c = ctx.add_system(m)
        e, f = ctx.calc_forces(c)
        check(e, f, eo, fo, pdgpu.FP64)
