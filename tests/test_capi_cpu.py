"""C-ABI checks that need no GPU: the library loads, exports every entry
point the header declares, and its host-side setup helpers (PD's BondOrder
band and Still fixed terms) match the compiled reference's values."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import golden_io as G
import pdgpu

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pdgpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(pdgpu_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_binding_set():
    assert set(declared_functions()) == set(pdgpu.EXPORTS)


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    assert hasattr(pdgpu.lib, name)


def test_options_default_match_reference_constructor():
    o = pdgpu.default_options()
    assert (o.cutoff, o.inner_cutoff, o.vdw_cutoff, o.vdw_inner_cutoff) == (15.0, 12.0, 15.0, 12.0)
    assert (o.dielectric_solute, o.dielectric_solvent, o.gbsa_still_cutoff) == (1.0, 80.0, 10.0)
    assert (o.exp_approx_threshold, o.dielectric_offset, o.born_radius_offset) == (-4.0, -0.09, 0.0)
    assert (o.p1, o.p2, o.p3, o.p4, o.p5) == (0.073, 0.921, 6.211, 15.236, 1.254)
    assert o.fast_mode == 1 and o.elec_mode == 3 and o.nlist_padding == 0.5


@pytest.mark.parametrize("name", G.case_names())
def test_bondorder_build_matches_reference(name):
    g = G.load(name)
    mid, band, si, sj, so = pdgpu.bondorder_build(g["resid"], g["a12"], g["b12"], g["a13"], g["b13"],
                                                  g["a14"], g["b14"])
    assert mid == g["mid"]
    assert np.array_equal(band, g["band"])
    assert len(si) == len(g["sp_i"])


def test_bondorder_sparse_ring_and_long_bond():
    # 10-atom chain in residues of 2 atoms plus a long-range 0-9 bond (a
    # disulfide-like link): the link is outside the band and must come back
    # as a sparse pair with order 1, its 1-3 partners with order 2.
    n = 10
    resid = np.arange(n) // 2
    a12 = list(range(n - 1)) + [0]
    b12 = list(range(1, n)) + [9]
    a13 = list(range(n - 2)) + [1, 0]
    b13 = list(range(2, n)) + [9, 8]
    a14 = list(range(n - 3))
    b14 = list(range(3, n))
    mid, band, si, sj, so = pdgpu.bondorder_build(resid, a12, b12, a13, b13, a14, b14)
    assert mid == 4  # widest in-neighbour-residue gap is 3 (1-4), plus one
    sp = {(int(i), int(j)): int(o) for i, j, o in zip(si, sj, so)}
    assert sp[(0, 9)] == 1 and sp[(1, 9)] == 2 and sp[(0, 8)] == 2
    assert band[0 * mid + 0] == 1 and band[0 * mid + 1] == 2 and band[0 * mid + 2] == 3


@pytest.mark.parametrize("name", G.case_names())
def test_fixed_terms_match_reference(name):
    g = G.load(name)
    o = pdgpu.default_options(dielectric_offset=g["param_dict"]["dielectric_offset"])
    vj, t = pdgpu.fixed_terms(g["gb_radius"], g["a12"], g["b12"], g["d12"], g["a13"], g["b13"], g["d13"], o)
    assert np.allclose(vj, g["vj"], rtol=1e-13, atol=1e-13)
    assert np.allclose(t, g["terms123"], rtol=1e-13, atol=0)


def test_fixed_terms_isolated_atom_closed_form():
    o = pdgpu.default_options()
    R = 1.7
    vj, t = pdgpu.fixed_terms([R], [], [], [], [], [], [], o)
    assert vj[0] == pytest.approx(4.0 / 3.0 * np.pi * R ** 3, rel=1e-15)
    assert t[0] == pytest.approx(-166.0 / (R - 0.09 + 0.073), rel=1e-15)
    # hydrogen rule: a type radius <= 0.01 is replaced by 1.15
    vj, t = pdgpu.fixed_terms([0.0], [], [], [], [], [], [], o)
    assert vj[0] == pytest.approx(4.0 / 3.0 * np.pi * 1.15 ** 3, rel=1e-15)


def test_ctx_create_rejects_bad_options():
    for kw in ({"fast_mode": 0}, {"cutoff": -1.0}, {"inner_cutoff": 20.0}, {"precision": 7}, {"tile_split": 3},
               {"tile_split": 32}, {"tile_split": -1}):
        o = pdgpu.default_options(**kw)
        h = C.c_void_p()
        rc = pdgpu.lib.pdgpu_ctx_create(C.byref(o), C.byref(h))
        assert rc in (pdgpu.PDGPU_ERR_UNSUPPORTED, pdgpu.PDGPU_ERR_INVALID)
        assert not h.value
