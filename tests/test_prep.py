"""Input preparation (pdgpu_ff_read + pdgpu_peptide_build, pdgpu.prep)
against the compiled reference's own NewProtein runs: every golden case is
rebuilt from PD's parameter file and its sequence and must give the same
molecule. The topology (atom order, residues, 1-2 / 1-3 / 1-4 lists, bond
order band, FF_Bonded term indices) is bit-exact; charges, radii, epsilon,
masses, Still radii and the bonded parameters are parsed by the same
sscanf and must be equal; the equilibrium distances, the bond and angle
terms (some of which PD takes from the built geometry) and the fixed GB
terms within 1e-12 relative; the extended-chain coordinates (two sets of phi /
psi / omega rotations over the whole chain) within 1e-10 A.

Needs PD's parameter file (the user's own; none ships with the library),
so it runs where the reference tree is present and is skipped elsewhere.
The grammar and setter tests beside it check properties that hold for any
parameter file: equivalent spellings build identical molecules, malformed
sequences are refused, a set phi/psi reads back, and rotations keep every
bond length."""
import os

import numpy as np
import pytest

import golden_io as G
import pdgpu
from pdgpu import prep, synth

PARAM = "/root/reference/param/amber99aa.ff"
pytestmark = pytest.mark.skipif(not os.path.exists(PARAM), reason="PD's parameter file not present")

# the sequences ref_dump built each golden case from
SEQUENCES = {"tri_ala": "AAA", "gly_ala": "GA", "c1_seed1": synth.random_sequence(1, 20),
             "c1_seed2": synth.random_sequence(2, 20), "c1_seed3": synth.random_sequence(3, 20),
             "res60_seed4": synth.random_sequence(4, 60), "pd_defaults_seed5": synth.random_sequence(5, 12),
             "vdw_default_seed6": synth.random_sequence(6, 12), "cut10_seed7": synth.random_sequence(7, 30)}


@pytest.fixture(scope="module")
def ff():
    return prep.read_ff(PARAM)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def test_parameter_file_globals(ff):
    # VDW14SCALING / ELEC14SCALING lines and GB GENERALPARAMETERS
    g = G.load("c1_seed1")
    assert ff.vdw14_scaling == g["param_dict"]["vdw14_scaling"]
    assert ff.elec14_scaling == g["param_dict"]["elec14_scaling"]
    assert np.array_equal(ff.gb_p, [g["param_dict"][k] for k in ("P1", "P2", "P3", "P4", "P5")])


@pytest.mark.parametrize("name", sorted(SEQUENCES))
def test_newprotein_matches_reference(ff, name):
    g = G.load(name)
    m = ff.peptide(prep.pd_grammar(SEQUENCES[name]))
    assert m["n"] == g["n"]
    for k in ("resid", "a12", "b12", "a13", "b13", "a14", "b14", "band", "sp_i", "sp_j", "sp_o",
              "bd_idx", "an_idx", "to_idx", "to_terms", "im_idx", "im_terms"):
        assert np.array_equal(np.asarray(m[k]), np.asarray(g[k])), k
    assert m["mid"] == g["mid"]
    for k in ("charge", "radius", "epsilon", "mass", "gb_radius", "to_vn", "to_n", "to_gamma", "im_vn", "im_n",
              "im_gamma"):
        assert np.array_equal(m[k], g[k]), k
    # bond and angle terms without a parameter line take the built
    # geometry's length or angle (FF_Bonded's fallback), so they follow the
    # coordinates' last bits
    for k in ("d12", "bd_lk", "an_tk", "d13", "vj", "terms123"):
        assert rel(m[k], g[k]) <= 1e-12, k
    assert np.max(np.abs(m["xyz"] - g["xyz"])) <= 1e-10
    if "is_h" in g:
        assert np.array_equal(m["is_h"], g["is_h"].astype(np.uint8))
    if "io_names" in g:
        names = g["io_names"].view(np.uint8).reshape(g["n"], 4, 16)
        text = lambda row: bytes(row).split(b"\0")[0].decode()
        assert [tuple(text(names[i, q]) for q in range(4)) for i in range(g["n"])] == m["names"]
        assert np.array_equal(m["backbone"].ravel(), g["io_backbone"])
        assert np.array_equal(m["n_cov12"], g["io_ncov12"])
        assert np.array_equal(m["cov12"].ravel(), g["io_cov12"])


def test_btf_header_of_a_built_peptide_matches_reference(ff):
    """The BTF system definition written from the builder's own names and
    covalent lists equals the reference's (same masks as the header test)."""
    g = G.load("c1_seed2")
    m = ff.peptide(prep.pd_grammar(SEQUENCES["c1_seed2"]))
    ref = np.frombuffer(g["io_btf"].view(np.uint8).tobytes(), np.uint8)
    date = np.frombuffer(ref[24:36].tobytes(), np.int32)
    atoms = prep.btf_atoms(m)
    for a, r in zip(atoms, g["io_ref"].reshape(-1, 3)):
        a["ref"] = r  # the reference position of the run the file came from
    h = np.frombuffer(pdgpu.btf_header(atoms, m["nres"], pdgpu.BTF_ALL, date), np.uint8)
    undefined = G.btf_undefined_bytes(g, len(h))
    assert np.array_equal(h[~undefined], ref[:len(h)][~undefined])


@pytest.mark.parametrize("spelling", ["*A-(AA)-A*", "*A-A[2]-A*", "*A-<A>[2]-A*", "*ALA-ALA-ALA-ALA*",
                                      "*a-(AA)-a*", "*A-(A[2])-A*"])
def test_grammar_spellings_build_the_same_molecule(ff, spelling):
    ref = ff.peptide("*A-(AA)-A*")
    m = ff.peptide(spelling)
    assert m["names"] == ref["names"]
    assert np.array_equal(m["xyz"], ref["xyz"])


@pytest.mark.parametrize("bad", ["*A*", "A-(AA", "*A-(aa)-A*", "*A-A[0]-A*", "*A-(A*)-A*", "*A-ZZZ-A*", "*A-(A)-A*-A", "A>"])
def test_grammar_refuses_malformed_sequences(ff, bad):
    with pytest.raises(pdgpu.PdgpuError):
        ff.peptide(bad)


def test_terminal_residues_take_the_class_codes(ff):
    m = ff.peptide("*G-(AS)-W*")
    parents = [n[3] for n in m["names"]]
    assert parents[0] == "NGLY" and parents[-1] == "CTRP"
    assert "ALA" in parents and "SER" in parents


def test_phipsi_setters_read_back_and_keep_bond_lengths(ff):
    seq = prep.pd_grammar(synth.random_sequence(9, 15))
    rng = np.random.default_rng(3)
    target = np.stack([rng.uniform(-np.pi, np.pi, 15), rng.uniform(-np.pi, np.pi, 15),
                       np.full(15, np.pi)], axis=1)
    base = prep.Molecule(ff, seq)
    mol = prep.Molecule(ff, seq, phipsi=target)
    got = mol.phipsi()
    # phi of the first and psi of the last residue are undefined (0)
    ang = lambda a, b: np.abs(np.angle(np.exp(1j * (a - b))))
    assert np.all(ang(got[1:, 0], target[1:, 0]) < 1e-9)
    assert np.all(ang(got[:-1, 1], target[:-1, 1]) < 1e-9)
    assert np.all(ang(got[1:, 2], np.pi) < 1e-9)
    d0, d1 = base.as_dict(), mol.as_dict()
    bond = lambda d: np.linalg.norm(d["xyz"][d["a12"]] - d["xyz"][d["b12"]], axis=1)
    assert np.max(np.abs(bond(d0) - bond(d1))) < 1e-9
    assert not np.allclose(d0["xyz"], d1["xyz"])


def test_missing_parameter_file_is_an_error():
    with pytest.raises(pdgpu.PdgpuError):
        prep.read_ff("/nonexistent/file.ff")


def test_workloads_build_from_the_parameter_file():
    """use_parameter_file switches the workload builders to NewProtein of
    the same sequences and lengths."""
    from pdgpu import workloads as W
    W.use_parameter_file(PARAM)
    try:
        mols = W.batch_molecules(4, workers=1)
        ch = W.chain_molecule()
    finally:
        W.use_parameter_file(None)
    lengths = W.batch_lengths(4)
    for k, m in enumerate(mols):
        assert m["nres"] == lengths[k]
        assert m["sequence"] == prep.pd_grammar(synth.random_sequence(1000 + k, lengths[k]))
        for key in ("charge", "radius", "epsilon", "gb_radius", "vj", "terms123", "mid", "band", "xyz", "bd_idx"):
            assert key in m
    assert ch["nres"] == 500
    # the chain takes the basin angles
    pp = prep.Molecule(prep.read_ff(PARAM), ch["sequence"], W.basin_angles(2, 500)).phipsi()
    assert abs(np.angle(np.exp(1j * (pp[10, 0] - W.basin_angles(2, 500)[10, 0])))) < 1e-9
