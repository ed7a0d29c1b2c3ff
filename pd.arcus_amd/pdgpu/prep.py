"""Input preparation through the library's host code: a PD parameter file
read (pdgpu_ff_read, FFParamSet::readLib) and peptides built with PD's
NewProtein (pdgpu_peptide_build), returned as the per-atom mapping that
Context.add_system takes, with the FF_Bonded terms, the covalent pair lists
and the names a BTF header needs.

    ff = prep.read_ff("amber99aa.ff")       # the user's own PD parameter file
    mol = ff.peptide("*A-(CDEFGHIKLMNPQRSTVW)-Y*")
    ctx.add_system(mol, bonded=True)

No parameter file ships with this package: PD's are the user's own."""
import ctypes as C

import numpy as np

from . import (PdgpuError, bondorder_build, default_options, fixed_terms, lib, _dp, _ip, _p)

_bp8 = C.POINTER(C.c_uint8)


class _BondedOut(C.Structure):
    _fields_ = [("nbonds", C.c_int32), ("nangles", C.c_int32), ("ntorsions", C.c_int32), ("nimpropers", C.c_int32),
                ("bond_ij", _ip), ("bond_lk", _dp), ("angle_iaj", _ip), ("angle_tk", _dp),
                ("torsion_iabj", _ip), ("torsion_terms", _ip), ("torsion_vn", _dp), ("torsion_n", _dp),
                ("torsion_gamma", _dp), ("improper_iabj", _ip), ("improper_terms", _ip), ("improper_vn", _dp),
                ("improper_n", _dp), ("improper_gamma", _dp)]


_SIG = {
    "pdgpu_ff_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "pdgpu_ff_free": (None, [C.c_void_p]),
    "pdgpu_prep_error": (C.c_char_p, []),
    "pdgpu_ff_info": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pdgpu_peptide_build": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "pdgpu_molecule_free": (None, [C.c_void_p]),
    "pdgpu_molecule_counts": (C.c_int, [C.c_void_p, _ip]),
    "pdgpu_molecule_atoms": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _bp8]),
    "pdgpu_molecule_atom_names": (C.c_int, [C.c_void_p, C.c_int32] + [C.POINTER(C.c_char_p)] * 4),
    "pdgpu_molecule_pairs": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _dp]),
    "pdgpu_molecule_cov12": (C.c_int, [C.c_void_p, _ip, _ip]),
    "pdgpu_molecule_backbone": (C.c_int, [C.c_void_p, _ip]),
    "pdgpu_molecule_set_phipsi": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double]),
    "pdgpu_molecule_get_phipsi": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp]),
    "pdgpu_molecule_bonded": (C.c_int, [C.c_void_p, C.POINTER(_BondedOut)]),
}
for _name, (_res, _args) in _SIG.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _check(rc, what):
    if rc != 0:
        raise PdgpuError(rc, f"{what}: {lib.pdgpu_prep_error().decode()}")


class ForceField:
    """A PD parameter set read by pdgpu_ff_read."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib.pdgpu_ff_read(str(path).encode(), C.byref(h)), f"ff_read {path}")
        self.h = h
        self._free = lib.pdgpu_ff_free  # kept: module globals may be gone when __del__ runs at exit
        v, e = C.c_double(), C.c_double()
        p = np.zeros(5)
        lib.pdgpu_ff_info(self.h, C.byref(v), C.byref(e), _p(p, _dp))
        self.vdw14_scaling, self.elec14_scaling, self.gb_p = v.value, e.value, p

    def __del__(self):
        if getattr(self, "h", None):
            self._free(self.h)
            self.h = None

    def options(self, **kw):
        """pdgpu options with this parameter set's 1-4 scalings and GB
        P1..P5 (what FF_GeneralizedBorn_Still reads from it)."""
        d = dict(vdw14_scaling=self.vdw14_scaling, elec14_scaling=self.elec14_scaling,
                 p1=self.gb_p[0], p2=self.gb_p[1], p3=self.gb_p[2], p4=self.gb_p[3], p5=self.gb_p[4])
        d.update(kw)
        return default_options(**d)

    def peptide(self, sequence, phipsi=None):
        """NewProtein(sequence) as a mapping for Context.add_system.
        phipsi: optional (nres, 2) or (nres, 3) radians (phi, psi[, omega])
        applied after the build with PD's setters."""
        return Molecule(self, sequence, phipsi).as_dict()


class Molecule:
    """A peptide built by pdgpu_peptide_build."""

    def __init__(self, ff, sequence, phipsi=None):
        h = C.c_void_p()
        _check(lib.pdgpu_peptide_build(ff.h, sequence.encode(), C.byref(h)), f"peptide_build {sequence!r}")
        self.h, self.ff, self.sequence = h, ff, sequence
        c = np.zeros(9, np.int32)
        lib.pdgpu_molecule_counts(self.h, _p(c, _ip))
        self.counts = c
        if phipsi is not None:
            pp = np.asarray(phipsi, dtype=np.float64)
            for r in range(len(pp)):
                om = pp[r, 2] if pp.shape[1] > 2 else np.nan
                _check(lib.pdgpu_molecule_set_phipsi(self.h, r, pp[r, 0], pp[r, 1], om), "set_phipsi")

    def __del__(self):
        if getattr(self, "h", None):
            lib.pdgpu_molecule_free(self.h)
            self.h = None

    def phipsi(self):
        nres = int(self.counts[1])
        out = np.zeros((nres, 3))
        for r in range(nres):
            a, b, o = C.c_double(), C.c_double(), C.c_double()
            lib.pdgpu_molecule_get_phipsi(self.h, r, C.byref(a), C.byref(b), C.byref(o))
            out[r] = a.value, b.value, o.value
        return out

    def as_dict(self):
        n, nres, n12, n13, n14, nb, na, nt, ni = (int(x) for x in self.counts)
        xyz = np.zeros(3 * n)
        q, rad, eps, mass, gbr = (np.zeros(n) for _ in range(5))
        resid = np.zeros(n, np.int32)
        is_h = np.zeros(n, np.uint8)
        _check(lib.pdgpu_molecule_atoms(self.h, _p(xyz, _dp), _p(q, _dp), _p(rad, _dp), _p(eps, _dp),
                                        _p(mass, _dp), _p(gbr, _dp), _p(resid, _ip),
                                        is_h.ctypes.data_as(_bp8)), "molecule_atoms")
        pairs = {}
        for order, m in ((2, n12), (3, n13), (4, n14)):
            a, b, d = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m)
            _check(lib.pdgpu_molecule_pairs(self.h, order, _p(a, _ip), _p(b, _ip), _p(d, _dp)), "molecule_pairs")
            pairs[order] = (a, b, d)
        cov = np.zeros(6 * n, np.int32)
        ncov = np.zeros(n, np.int32)
        lib.pdgpu_molecule_cov12(self.h, _p(cov, _ip), _p(ncov, _ip))
        bb = np.zeros(3 * nres, np.int32)
        lib.pdgpu_molecule_backbone(self.h, _p(bb, _ip))
        names = []
        for i in range(n):
            s = [C.c_char_p() for _ in range(4)]
            lib.pdgpu_molecule_atom_names(self.h, i, *[C.byref(x) for x in s])
            names.append(tuple(x.value.decode() for x in s))
        arr = dict(bd_idx=np.zeros(2 * nb, np.int32), bd_lk=np.zeros(2 * nb), an_idx=np.zeros(3 * na, np.int32),
                   an_tk=np.zeros(2 * na), to_idx=np.zeros(4 * nt, np.int32), to_terms=np.zeros(nt, np.int32),
                   to_vn=np.zeros(4 * nt), to_n=np.zeros(4 * nt), to_gamma=np.zeros(4 * nt),
                   im_idx=np.zeros(4 * ni, np.int32), im_terms=np.zeros(ni, np.int32), im_vn=np.zeros(4 * ni),
                   im_n=np.zeros(4 * ni), im_gamma=np.zeros(4 * ni))
        o = _BondedOut()
        for f, k in (("bond_ij", "bd_idx"), ("bond_lk", "bd_lk"), ("angle_iaj", "an_idx"), ("angle_tk", "an_tk"),
                     ("torsion_iabj", "to_idx"), ("torsion_terms", "to_terms"), ("torsion_vn", "to_vn"),
                     ("torsion_n", "to_n"), ("torsion_gamma", "to_gamma"), ("improper_iabj", "im_idx"),
                     ("improper_terms", "im_terms"), ("improper_vn", "im_vn"), ("improper_n", "im_n"),
                     ("improper_gamma", "im_gamma")):
            setattr(o, f, _p(arr[k], _ip if arr[k].dtype == np.int32 else _dp))
        _check(lib.pdgpu_molecule_bonded(self.h, C.byref(o)), "molecule_bonded")
        (a12, b12, d12), (a13, b13, d13), (a14, b14, _) = pairs[2], pairs[3], pairs[4]
        mid, band, si, sj, so = bondorder_build(resid, a12, b12, a13, b13, a14, b14)
        vj, t123 = fixed_terms(gbr, a12, b12, d12, a13, b13, d13, self.ff.options())
        d = {"xyz": xyz.reshape(-1, 3), "charge": q, "radius": rad, "epsilon": eps, "gb_radius": gbr, "mass": mass,
             "resid": resid, "is_h": is_h, "a12": a12, "b12": b12, "d12": d12, "a13": a13, "b13": b13, "d13": d13,
             "a14": a14, "b14": b14, "mid": mid, "band": band, "sp_i": si, "sp_j": sj, "sp_o": so, "vj": vj,
             "terms123": t123, "n": n, "nres": nres, "names": names, "cov12": cov.reshape(n, 6), "n_cov12": ncov,
             "backbone": bb.reshape(nres, 3), "sequence": self.sequence}
        d.update(arr)
        return d


def read_ff(path):
    return ForceField(path)


def pd_grammar(seq):
    """A one-letter sequence in PD's grammar with charged termini:
    '*A-(CDE...)-Y*'."""
    if len(seq) < 2:
        raise ValueError("a polymer needs two residues or more")
    mid = seq[1:-1]
    return f"*{seq[0]}-({mid})-{seq[-1]}*" if mid else f"*{seq[0]}-{seq[-1]}*"


def btf_atoms(mol):
    """BTF system-definition atoms (pdgpu.btf_header) of a built peptide:
    not from experimental data, the built positions as reference."""
    out = []
    for i, (pdb, typ, raw, parent) in enumerate(mol["names"]):
        out.append(dict(pdbname=pdb, type_name=typ, rawname=raw, parentname=parent, ir=int(mol["resid"][i]),
                        known_structure=0, ref=mol["xyz"][i], cov12=list(mol["cov12"][i, :mol["n_cov12"][i]]),
                        charge=mol["charge"][i], radius=mol["radius"][i], epsilon=mol["epsilon"][i],
                        mass=mol["mass"][i]))
    return out
