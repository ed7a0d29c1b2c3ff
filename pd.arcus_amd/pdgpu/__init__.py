"""Python binding of libpdgpu.so, the MI355X implementation of PD's
Generalized Born / nonbonded forcefield evaluation (C ABI: include/pdgpu.h).

The binding is a thin ctypes layer: every call goes through the C ABI.
There is no fallback; if the shared library is missing, importing the
module raises.

Units follow PD (forcefields/forcefield.cpp:18-38): Angstrom, e,
sqrt(J/molecule) for epsilon, J/molecule, Newton, kg.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpdgpu.so")

PDGPU_OK = 0
PDGPU_ERR_INVALID = -1
PDGPU_ERR_NOMEM = -2
PDGPU_ERR_NONFINITE = -3
PDGPU_ERR_HIP = -4
PDGPU_ERR_UNSUPPORTED = -5
PDGPU_ERR_NODEVICE = -6
PDGPU_ALL = -1
FP64 = 0
MIXED = 1

KCAL2J = 4184.0
NA = 6.0221418e23
J_PER_KCALMOL = KCAL2J / NA
N_PER_KCALMOL_A = J_PER_KCALMOL / 1e-10

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_bp = C.POINTER(C.c_int8)

# every entry point declared in include/pdgpu.h
EXPORTS = (
    "pdgpu_options_default", "pdgpu_ctx_create", "pdgpu_ctx_destroy", "pdgpu_last_error",
    "pdgpu_instance_add", "pdgpu_instance_remove", "pdgpu_instance_natoms",
    "pdgpu_set_positions", "pdgpu_get_positions", "pdgpu_nlist_build", "pdgpu_nlist_export",
    "pdgpu_calc_forces", "pdgpu_calc_energies", "pdgpu_get_born_radii", "pdgpu_get_forces",
    "pdgpu_minimise_sd", "pdgpu_set_velocities", "pdgpu_get_velocities", "pdgpu_md_vv",
    "pdgpu_bondorder_build", "pdgpu_fixed_terms", "pdgpu_instance_bytes", "pdgpu_bytes_free",
    "pdgpu_profile", "pdgpu_profile_read", "pdgpu_sd_start", "pdgpu_sd_advance", "pdgpu_sd_finish",
    "pdgpu_bonded_set", "pdgpu_bonded_bytes", "pdgpu_autotune", "pdgpu_md_vv_pd", "pdgpu_md_langevin_pd",
    "pdgpu_md_beeman_pd", "pdgpu_mime_encode", "pdgpu_snapshot_mime", "pdgpu_dcd_header", "pdgpu_dcd_frame",
    "pdgpu_dcd_create", "pdgpu_dcd_append", "pdgpu_traj_dcd", "pdgpu_set_hydrogens",
    "pdgpu_set_thermostat", "pdgpu_nlist_export_noalloc", "pdgpu_options_default_nonbonded",
    "pdgpu_btf_header", "pdgpu_btf_frame", "pdgpu_btf_append",
    "pdgpu_ff_read", "pdgpu_ff_free", "pdgpu_prep_error", "pdgpu_ff_info", "pdgpu_peptide_build",
    "pdgpu_molecule_free", "pdgpu_molecule_counts", "pdgpu_molecule_atoms", "pdgpu_molecule_atom_names",
    "pdgpu_molecule_pairs", "pdgpu_molecule_cov12", "pdgpu_molecule_backbone", "pdgpu_molecule_set_phipsi",
    "pdgpu_molecule_get_phipsi", "pdgpu_molecule_bonded",
)
# BTF includes (PD's BFT_StandardIncludes)
BTF_PHIPSI, BTF_ENERGY, BTF_FORCE = 2, 8, 16
BTF_ALL = BTF_PHIPSI | BTF_ENERGY | BTF_FORCE
# forcefield components (pdgpu_options.components)
FF_GB, FF_BONDS, FF_ANGLES, FF_TORSIONS, FF_IMPROPERS = 1, 2, 4, 8, 16
FF_BONDED = FF_BONDS | FF_ANGLES | FF_TORSIONS | FF_IMPROPERS
FF_SOFTVDW = 32  # FF_SoftVDW soft-core repulsion (not in FF_ALL; not with FF_GB)
FF_NONBONDED = 64  # FF_NonBonded alone: vacuum LJ + Coulomb (not in FF_ALL; not with FF_GB)
FF_ALL = FF_GB | FF_BONDED
KERNEL_GROUPS = ("nlist", "atom", "born", "pair", "chain", "integrate", "enum", "bonded")


def mime_encode(xyz, vel, force, epot, name=""):
    """PD's SnapShot MIME text of the given arrays (A, m/s, N; epot in
    J/molecule), as pdgpu_mime_encode writes it."""
    x, v, f = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (xyz, vel, force))
    n = len(x) // 3
    nb = name.encode()
    need = lib.pdgpu_mime_encode(_p(x, _dp), _p(v, _dp), _p(f, _dp), n, epot, nb, None, 0)
    if need < 0:
        raise PdgpuError(need, "mime_encode")
    buf = C.create_string_buffer(need + 1)
    lib.pdgpu_mime_encode(_p(x, _dp), _p(v, _dp), _p(f, _dp), n, epot, nb, buf, need + 1)
    return buf.raw[:need].decode()


def dcd_header(natoms):
    """The DCD header PD's OutTra_NAMD writes for natoms atoms (bytes)."""
    need = lib.pdgpu_dcd_header(natoms, None, 0)
    buf = C.create_string_buffer(need)
    lib.pdgpu_dcd_header(natoms, buf, need)
    return buf.raw


class BtfAtom(C.Structure):
    _fields_ = [("pdbname", C.c_char_p), ("type_name", C.c_char_p), ("rawname", C.c_char_p),
                ("parentname", C.c_char_p), ("ir", C.c_int32), ("known_structure", C.c_int32),
                ("ref", C.c_double * 3), ("n_cov12", C.c_int32), ("cov12", C.c_int32 * 6),
                ("charge", C.c_double), ("radius", C.c_double), ("epsilon", C.c_double), ("mass", C.c_double)]


class BtfState(C.Structure):
    _fields_ = [("step", C.c_int64), ("crms", C.c_double), ("drms", C.c_double), ("ekin", C.c_double),
                ("etot", C.c_double), ("data", C.c_double * 10)]


def btf_header(atoms, nresidues, includes=BTF_ALL, date=None):
    """The start of a BTF file (bytes): header, system definition, TRASTART.
    atoms: list of dicts with keys pdbname, type_name, rawname, parentname
    (str), ir, known_structure (int), ref (3 floats), cov12 (up to 6 ints),
    charge, radius, epsilon, mass; date: (day, month, year) or None."""
    arr = (BtfAtom * max(1, len(atoms)))()
    keep = []
    for k, a in enumerate(atoms):
        names = [a[f].encode() for f in ("pdbname", "type_name", "rawname", "parentname")]
        keep += names
        r = arr[k]
        r.pdbname, r.type_name, r.rawname, r.parentname = names
        r.ir, r.known_structure = int(a["ir"]), int(a["known_structure"])
        r.ref[:] = [float(x) for x in a["ref"]]
        cov = [int(x) for x in a["cov12"]]
        r.n_cov12 = len(cov)
        for q, j in enumerate(cov[:6]):
            r.cov12[q] = j
        r.charge, r.radius, r.epsilon, r.mass = (float(a[f]) for f in ("charge", "radius", "epsilon", "mass"))
    d = None if date is None else (C.c_int32 * 3)(*[int(x) for x in date])
    need = lib.pdgpu_btf_header(arr, len(atoms), nresidues, includes, d, None, 0)
    if need < 0:
        raise PdgpuError(need, "btf_header")
    buf = C.create_string_buffer(need)
    got = lib.pdgpu_btf_header(arr, len(atoms), nresidues, includes, d, buf, need)
    if got != need:
        raise PdgpuError(got, "btf_header")
    return buf.raw


def btf_state(step=0, crms=0.0, drms=0.0, ekin=0.0, etot=0.0, data=None):
    s = BtfState()
    s.step, s.crms, s.drms, s.ekin, s.etot = int(step), crms, drms, ekin, etot
    s.data[:] = [float(x) for x in (data if data is not None else [0.0] * 10)]
    return s


class PdgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdgpu error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32), ("fast_mode", C.c_int32),
                ("elec_mode", C.c_int32), ("vdw_mode", C.c_int32), ("tiles_per_wave", C.c_int32)] + \
               [(n, C.c_double) for n in (
                   "cutoff", "inner_cutoff", "vdw_cutoff", "vdw_inner_cutoff", "vdw14_scaling",
                   "elec14_scaling", "dielectric_solute", "dielectric_solvent", "gbsa_still_cutoff",
                   "exp_approx_threshold", "dielectric_offset", "born_radius_offset",
                   "p1", "p2", "p3", "p4", "p5", "nlist_padding")] + \
               [("device_memory_budget", C.c_int64), ("max_instances", C.c_int32), ("components", C.c_int32),
                ("use_graphs", C.c_int32), ("dense_tile_pairs", C.c_int32),
                ("langevin_on_hydrogens", C.c_int32), ("softvdw_hardness", C.c_double), ("softvdw_olap_offset", C.c_double),
                ("softvdw_cutoff", C.c_double), ("tile_split", C.c_int32)]


class Bonded(C.Structure):
    _fields_ = [("nbonds", C.c_int32), ("nangles", C.c_int32), ("ntorsions", C.c_int32), ("nimpropers", C.c_int32),
                ("bond_ij", _ip), ("bond_lk", _dp), ("angle_iaj", _ip), ("angle_tk", _dp),
                ("torsion_iabj", _ip), ("torsion_terms", _ip), ("torsion_vn", _dp), ("torsion_n", _dp),
                ("torsion_gamma", _dp), ("improper_iabj", _ip), ("improper_terms", _ip), ("improper_vn", _dp),
                ("improper_n", _dp), ("improper_gamma", _dp)]


# keys of a bonded-term dict (the layout of pdgpu.synth and of the reference fixtures)
BONDED_KEYS = ("bd_idx", "bd_lk", "an_idx", "an_tk", "to_idx", "to_terms", "to_vn", "to_n", "to_gamma",
               "im_idx", "im_terms", "im_vn", "im_n", "im_gamma")


class Topology(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("max_index_delta", C.c_int32),
                ("charge", _dp), ("radius", _dp), ("epsilon", _dp), ("gb_radius", _dp),
                ("vj", _dp), ("terms123", _dp), ("mass", _dp), ("band", _bp),
                ("nsparse", C.c_int32), ("reserved", C.c_int32),
                ("sparse_i", _ip), ("sparse_j", _ip), ("sparse_order", _ip)]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "epot", "epot_vdw", "epot_elec", "epot_pol", "epot_pol_self", "epot_pol_cross", "virial",
        "epot_bond", "epot_angle", "epot_torsion", "epot_improper")] + \
               [("pairs", C.c_int64), ("born_pairs", C.c_int64), ("status", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libpdgpu.so not built ({LIB_PATH}); run pd.arcus_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "pdgpu_options_default": (None, [P(Options)]),
        "pdgpu_options_default_nonbonded": (None, [P(Options)]),
        "pdgpu_nlist_export_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _up, C.c_int64,
                                                 P(C.c_int64)]),
        "pdgpu_ctx_create": (C.c_int, [P(Options), P(C.c_void_p)]),
        "pdgpu_ctx_destroy": (None, [C.c_void_p]),
        "pdgpu_last_error": (C.c_char_p, [C.c_void_p]),
        "pdgpu_instance_add": (C.c_int, [C.c_void_p, P(Topology), _ip]),
        "pdgpu_instance_remove": (C.c_int, [C.c_void_p, C.c_int32]),
        "pdgpu_instance_natoms": (C.c_int, [C.c_void_p, C.c_int32]),
        "pdgpu_set_positions": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_get_positions": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_set_velocities": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_get_velocities": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_nlist_build": (C.c_int, [C.c_void_p, C.c_int32]),
        "pdgpu_nlist_export": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _up, C.c_int64, P(C.c_int64)]),
        "pdgpu_calc_forces": (C.c_int, [C.c_void_p, C.c_int32, P(Energies), _dp]),
        "pdgpu_calc_energies": (C.c_int, [C.c_void_p, C.c_int32, P(Energies), _dp]),
        "pdgpu_get_born_radii": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_get_forces": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
        "pdgpu_minimise_sd": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double, _ip]),
        "pdgpu_md_vv": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, P(Energies), _dp]),
        "pdgpu_md_vv_pd": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, P(Energies), _dp]),
        "pdgpu_md_langevin_pd": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double,
                                           C.c_double, C.c_uint64, _dp, P(Energies), _dp]),
        "pdgpu_md_beeman_pd": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, P(Energies),
                                         _dp]),
        "pdgpu_mime_encode": (C.c_int64, [_dp, _dp, _dp, C.c_int32, C.c_double, C.c_char_p, C.c_char_p,
                                          C.c_int64]),
        "pdgpu_snapshot_mime": (C.c_int64, [C.c_void_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_int64]),
        "pdgpu_dcd_header": (C.c_int64, [C.c_int32, C.c_char_p, C.c_int64]),
        "pdgpu_dcd_frame": (C.c_int64, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int64]),
        "pdgpu_dcd_create": (C.c_int, [C.c_char_p, C.c_int32]),
        "pdgpu_dcd_append": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p]),
        "pdgpu_traj_dcd": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]),
        "pdgpu_set_hydrogens": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p]),
        "pdgpu_btf_header": (C.c_int64, [P(BtfAtom), C.c_int32, C.c_int32, C.c_int32, P(C.c_int32), C.c_char_p,
                                         C.c_int64]),
        "pdgpu_btf_frame": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, _ip, C.c_int32, P(BtfState),
                                        C.c_char_p, C.c_int64]),
        "pdgpu_btf_append": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, C.c_int32, P(BtfState),
                                       C.c_char_p]),
        "pdgpu_set_thermostat": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
        "pdgpu_bondorder_build": (C.c_int, [C.c_int32, _ip, C.c_int32, _ip, _ip, C.c_int32, _ip, _ip, C.c_int32,
                                            _ip, _ip, _ip, _bp, _ip, _ip, _ip, _ip]),
        "pdgpu_fixed_terms": (C.c_int, [C.c_int32, _dp, C.c_int32, _ip, _ip, _dp, C.c_int32, _ip, _ip, _dp,
                                        P(Options), _dp, _dp]),
        "pdgpu_instance_bytes": (C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
        "pdgpu_bytes_free": (C.c_int64, [C.c_void_p]),
        "pdgpu_profile": (C.c_int, [C.c_void_p, C.c_int32]),
        "pdgpu_sd_start": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_double]),
        "pdgpu_sd_advance": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _ip]),
        "pdgpu_sd_finish": (C.c_int, [C.c_void_p, C.c_int32, P(Energies), _ip]),
        "pdgpu_profile_read": (C.c_int, [C.c_void_p, _dp, P(C.c_int64), C.c_int32]),
        "pdgpu_bonded_set": (C.c_int, [C.c_void_p, C.c_int32, P(Bonded)]),
        "pdgpu_bonded_bytes": (C.c_int64, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
        "pdgpu_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_options_nonbonded(**kw):
    """Options of a new FF_NonBonded (force switching, Cutoff 15 / 12) with
    components FF_NONBONDED | FF_BONDED; keywords override fields."""
    o = Options()
    lib.pdgpu_options_default_nonbonded(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def default_options(**kw):
    o = Options()
    lib.pdgpu_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def bondorder_build(resid, a12, b12, a13, b13, a14, b14):
    """PD BondOrder band + sparse list from undirected pair lists."""
    resid = _i32(resid)
    n = len(resid)
    arrs = [_i32(x) for x in (a12, b12, a13, b13, a14, b14)]
    mid = np.zeros(1, np.int32)
    args = [n, _p(resid, _ip), len(arrs[0]), _p(arrs[0], _ip), _p(arrs[1], _ip), len(arrs[2]), _p(arrs[2], _ip),
            _p(arrs[3], _ip), len(arrs[4]), _p(arrs[4], _ip), _p(arrs[5], _ip), _p(mid, _ip)]
    rc = lib.pdgpu_bondorder_build(*args, None, None, None, None, None)
    if rc != 0:
        raise PdgpuError(rc, "bondorder_build")
    m = int(mid[0])
    band = np.zeros(max(0, m * (n - 1)), np.int8)
    cap = len(arrs[0]) + len(arrs[2]) + len(arrs[4]) + 1
    si, sj, so = (np.zeros(cap, np.int32) for _ in range(3))
    ns = np.zeros(1, np.int32)
    rc = lib.pdgpu_bondorder_build(*args, _p(band, _bp), _p(ns, _ip), _p(si, _ip), _p(sj, _ip), _p(so, _ip))
    if rc != 0:
        raise PdgpuError(rc, "bondorder_build")
    k = int(ns[0])
    return m, band, si[:k].copy(), sj[:k].copy(), so[:k].copy()


def fixed_terms(gb_radius, a12, b12, d12, a13, b13, d13, options=None):
    """Still V_j and terms123 (calcFixedBornRadiiTerms)."""
    o = options if options is not None else default_options()
    g = _f64(gb_radius)
    n = len(g)
    a12, b12, a13, b13 = (_i32(x) for x in (a12, b12, a13, b13))
    d12, d13 = _f64(d12), _f64(d13)
    vj = np.zeros(n)
    t = np.zeros(n)
    rc = lib.pdgpu_fixed_terms(n, _p(g, _dp), len(a12), _p(a12, _ip), _p(b12, _ip), _p(d12, _dp), len(a13),
                               _p(a13, _ip), _p(b13, _ip), _p(d13, _dp), C.byref(o), _p(vj, _dp), _p(t, _dp))
    if rc != 0:
        raise PdgpuError(rc, "fixed_terms")
    return vj, t


class Context:
    """One pdgpu context (one GPU stream) holding resident instances."""

    def __init__(self, options=None, **kw):
        self.opt = options if options is not None else default_options(**kw)
        h = C.c_void_p()
        rc = lib.pdgpu_ctx_create(C.byref(self.opt), C.byref(h))
        if rc != 0:
            raise PdgpuError(rc, "pdgpu_ctx_create failed")
        self.h = h
        self._natoms = {}
        self._earr = None

    def close(self):
        if self.h:
            lib.pdgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise PdgpuError(rc, f"{what}: {lib.pdgpu_last_error(self.h).decode()}")

    def add(self, charge, radius, epsilon, gb_radius, vj, terms123, mid, band, mass=None,
            sparse_i=None, sparse_j=None, sparse_order=None, xyz=None):
        keep = [_f64(charge), _f64(radius), _f64(epsilon), _f64(gb_radius), _f64(vj), _f64(terms123)]
        n = len(keep[0])
        massa = _f64(mass) if mass is not None else None
        banda = np.ascontiguousarray(band, dtype=np.int8)
        e = np.zeros(0, np.int32)
        si = _i32(sparse_i if sparse_i is not None else e)
        sj = _i32(sparse_j if sparse_j is not None else e)
        so = _i32(sparse_order if sparse_order is not None else e)
        t = Topology(n, int(mid), *[_p(a, _dp) for a in keep], _p(massa, _dp), _p(banda, _bp), len(si), 0,
                     _p(si, _ip), _p(sj, _ip), _p(so, _ip))
        i = C.c_int32()
        self._check(lib.pdgpu_instance_add(self.h, C.byref(t), C.byref(i)), "instance_add")
        self._natoms[i.value] = n
        if xyz is not None:
            self.set_positions(i.value, xyz)
        return i.value

    def add_system(self, s, xyz=None, bonded=False):
        """s: mapping with PD per-atom arrays (see pdgpu.synth / golden fixtures).
        bonded=True also uploads its FF_Bonded term lists (BONDED_KEYS)."""
        i = self.add(s["charge"], s["radius"], s["epsilon"], s["gb_radius"], s["vj"], s["terms123"],
                     s["mid"], s["band"], s.get("mass"), s.get("sp_i"), s.get("sp_j"), s.get("sp_o"),
                     xyz=s["xyz"] if xyz is None else xyz)
        if bonded:
            self.set_bonded(i, s)
        return i

    def set_bonded(self, iid, d):
        """FF_Bonded term lists (J/molecule units) from a mapping with BONDED_KEYS."""
        fi = lambda k: _i32(np.asarray(d[k]).ravel())
        fd = lambda k: _f64(np.asarray(d[k]).ravel())
        a = {k: (fi(k) if k.endswith(("idx", "terms")) else fd(k)) for k in BONDED_KEYS}
        b = Bonded(len(a["bd_lk"]) // 2, len(a["an_tk"]) // 2, len(a["to_terms"]), len(a["im_terms"]),
                   _p(a["bd_idx"], _ip), _p(a["bd_lk"], _dp), _p(a["an_idx"], _ip), _p(a["an_tk"], _dp),
                   _p(a["to_idx"], _ip), _p(a["to_terms"], _ip), _p(a["to_vn"], _dp), _p(a["to_n"], _dp),
                   _p(a["to_gamma"], _dp), _p(a["im_idx"], _ip), _p(a["im_terms"], _ip), _p(a["im_vn"], _dp),
                   _p(a["im_n"], _dp), _p(a["im_gamma"], _dp))
        self._check(lib.pdgpu_bonded_set(self.h, iid, C.byref(b)), "bonded_set")

    @staticmethod
    def bonded_bytes(d):
        return lib.pdgpu_bonded_bytes(len(d["bd_lk"]) // 2, len(d["an_tk"]) // 2, len(d["to_terms"]),
                                      len(d["im_terms"]))

    def remove(self, iid):
        self._check(lib.pdgpu_instance_remove(self.h, iid), "instance_remove")
        self._natoms.pop(iid, None)

    def natoms(self, iid):
        return self._natoms[iid]

    def set_positions(self, iid, xyz):
        x = _f64(xyz).reshape(-1)
        assert len(x) == 3 * self._natoms[iid]
        self._check(lib.pdgpu_set_positions(self.h, iid, _p(x, _dp)), "set_positions")

    def get_positions(self, iid):
        x = np.zeros(3 * self._natoms[iid])
        self._check(lib.pdgpu_get_positions(self.h, iid, _p(x, _dp)), "get_positions")
        return x.reshape(-1, 3)

    def set_velocities(self, iid, v):
        x = _f64(v).reshape(-1)
        assert len(x) == 3 * self._natoms[iid]
        self._check(lib.pdgpu_set_velocities(self.h, iid, _p(x, _dp)), "set_velocities")

    def get_velocities(self, iid):
        x = np.zeros(3 * self._natoms[iid])
        self._check(lib.pdgpu_get_velocities(self.h, iid, _p(x, _dp)), "get_velocities")
        return x.reshape(-1, 3)

    def nlist_build(self, iid=PDGPU_ALL):
        self._check(lib.pdgpu_nlist_build(self.h, iid), "nlist_build")

    def nlist_export(self, iid, shadow=True):
        n = self._natoms[iid]
        rs = np.zeros(n + 1, np.int32)
        cnt = C.c_int64()
        rc = lib.pdgpu_nlist_export(self.h, iid, int(shadow), _p(rs, _ip), None, 0, C.byref(cnt))
        if rc not in (0, PDGPU_ERR_NOMEM):
            self._check(rc, "nlist_export")
        ent = np.zeros(max(1, cnt.value), np.uint32)
        self._check(lib.pdgpu_nlist_export(self.h, iid, int(shadow), _p(rs, _ip), _p(ent, _up), len(ent),
                                           C.byref(cnt)), "nlist_export")
        return rs, ent[:cnt.value]

    def _calc(self, fn, iid, want_forces, check, raw=False):
        if iid == PDGPU_ALL:
            if self._earr is None:
                n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
                self._earr = (Energies * n)()
            arr = self._earr
            rc = fn(self.h, iid, arr, None)
            if check and rc not in (0, PDGPU_ERR_NONFINITE):
                self._check(rc, "calc")
            if raw:  # the per-slot pdgpu_energies array, reused by the next call
                return arr
            return {i: arr[i].as_dict() for i in self._natoms}
        e = Energies()
        f = np.zeros(3 * self._natoms[iid]) if want_forces else None
        rc = fn(self.h, iid, C.byref(e), _p(f, _dp))
        if check and rc not in (0, PDGPU_ERR_NONFINITE):
            self._check(rc, "calc")
        return e.as_dict(), (f.reshape(-1, 3) if want_forces else None)

    def calc_forces(self, iid=PDGPU_ALL, forces=True, check=True, raw=False):
        return self._calc(lib.pdgpu_calc_forces, iid, forces, check, raw)

    def calc_energies(self, iid=PDGPU_ALL, forces=True, check=True, raw=False):
        return self._calc(lib.pdgpu_calc_energies, iid, forces, check, raw)

    def born_radii(self, iid):
        b = np.zeros(self._natoms[iid])
        self._check(lib.pdgpu_get_born_radii(self.h, iid, _p(b, _dp)), "get_born_radii")
        return b

    def forces(self, iid):
        f = np.zeros(3 * self._natoms[iid])
        self._check(lib.pdgpu_get_forces(self.h, iid, _p(f, _dp)), "get_forces")
        return f.reshape(-1, 3)

    def minimise_sd(self, iid, steps, step_size=2e5, update_nlist=1, grms_tol=0.0):
        if iid == PDGPU_ALL:
            n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
            out = np.zeros(n, np.int32)
        else:
            out = np.zeros(1, np.int32)
        rc = lib.pdgpu_minimise_sd(self.h, iid, steps, step_size, update_nlist, grms_tol, _p(out, _ip))
        if rc not in (0, PDGPU_ERR_NONFINITE):
            self._check(rc, "minimise_sd")
        return out if iid == PDGPU_ALL else int(out[0])

    def sd_start(self, iid, steps, step_size=2e5, update_nlist=1, grms_tol=0.0):
        self._check(lib.pdgpu_sd_start(self.h, iid, steps, step_size, update_nlist, grms_tol), "sd_start")

    def sd_advance(self, nsteps):
        n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
        done = np.zeros(n, np.int32)
        steps = np.zeros(n, np.int32)
        running = C.c_int32()
        self._check(lib.pdgpu_sd_advance(self.h, nsteps, _p(done, _ip), _p(steps, _ip), C.byref(running)),
                    "sd_advance")
        return done, steps, running.value

    def sd_finish(self, iid):
        e = Energies()
        st = C.c_int32()
        rc = lib.pdgpu_sd_finish(self.h, iid, C.byref(e), C.byref(st))
        if rc not in (0, PDGPU_ERR_NONFINITE):
            self._check(rc, "sd_finish")
        return e.as_dict(), st.value

    def md_vv(self, iid, steps, dt, update_nlist=10):
        """Velocity Verlet on the device. For one instance: (energies dict,
        kinetic energy J); for PDGPU_ALL: dicts indexed by instance id."""
        if iid == PDGPU_ALL:
            n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
            arr = (Energies * n)()
            ek = np.zeros(n)
            self._check(lib.pdgpu_md_vv(self.h, iid, steps, dt, update_nlist, arr, _p(ek, _dp)), "md_vv")
            return {i: arr[i].as_dict() for i in self._natoms}, {i: float(ek[i]) for i in self._natoms}
        e = Energies()
        ek = np.zeros(1)
        self._check(lib.pdgpu_md_vv(self.h, iid, steps, dt, update_nlist, C.byref(e), _p(ek, _dp)), "md_vv")
        return e.as_dict(), float(ek[0])

    def md_vv_pd(self, iid, steps, dt, update_nlist=10):
        """PD's MolecularDynamics with its VelocityVerlet update as written
        (protocols/md.cpp:788-838). Returns like md_vv."""
        if iid == PDGPU_ALL:
            n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
            arr = (Energies * n)()
            ek = np.zeros(n)
            self._check(lib.pdgpu_md_vv_pd(self.h, iid, steps, dt, update_nlist, arr, _p(ek, _dp)), "md_vv_pd")
            return {i: arr[i].as_dict() for i in self._natoms}, {i: float(ek[i]) for i in self._natoms}
        e = Energies()
        ek = np.zeros(1)
        self._check(lib.pdgpu_md_vv_pd(self.h, iid, steps, dt, update_nlist, C.byref(e), _p(ek, _dp)), "md_vv_pd")
        return e.as_dict(), float(ek[0])


    def _md_call(self, name, fn, iid):
        if iid == PDGPU_ALL:
            n = self.opt.max_instances if self.opt.max_instances > 0 else 8192
            arr = (Energies * n)()
            ek = np.zeros(n)
            self._check(fn(arr, _p(ek, _dp)), name)
            return {i: arr[i].as_dict() for i in self._natoms}, {i: float(ek[i]) for i in self._natoms}
        e = Energies()
        ek = np.zeros(1)
        self._check(fn(C.byref(e), _p(ek, _dp)), name)
        return e.as_dict(), float(ek[0])

    def md_langevin_pd(self, iid, steps, dt, temperature=300.0, friction=1e13, seed=1, noise=None,
                       update_nlist=10):
        """PD's MolecularDynamics with its Langevin integrator
        (protocols/md.cpp:906-1043). noise: None (device Philox stream keyed
        by seed) or [steps][natoms][6] deviates (n1x, n1y, n1z, n2x, n2y,
        n2z) for a single instance. Returns like md_vv."""
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            if iid == PDGPU_ALL or nz.size != steps * self._natoms[iid] * 6:
                raise ValueError("noise must be [steps][natoms][6] for one instance")
        return self._md_call("md_langevin_pd", lambda e, k: lib.pdgpu_md_langevin_pd(
            self.h, iid, steps, dt, update_nlist, temperature, friction, seed,
            _p(nz, _dp) if nz is not None else None, e, k), iid)

    def md_beeman_pd(self, iid, steps, dt, update_nlist=10):
        """PD's MolecularDynamics with its Beeman integrator
        (protocols/md.cpp:840-904). Returns like md_vv."""
        return self._md_call("md_beeman_pd", lambda e, k: lib.pdgpu_md_beeman_pd(
            self.h, iid, steps, dt, update_nlist, e, k), iid)
    def snapshot_mime(self, iid, name=""):
        """WorkSpace::save().writeMIME(name) of instance iid (text)."""
        nb = name.encode()
        need = lib.pdgpu_snapshot_mime(self.h, iid, nb, None, 0)
        if need < 0:
            self._check(int(need), "snapshot_mime")
        buf = C.create_string_buffer(need + 1)
        got = lib.pdgpu_snapshot_mime(self.h, iid, nb, buf, need + 1)
        if got != need:
            self._check(int(got) if got < 0 else PDGPU_ERR_HIP, "snapshot_mime")
        return buf.raw[:need].decode()

    def dcd_frame(self, iid):
        """One DCD frame of instance iid's positions (bytes)."""
        need = lib.pdgpu_dcd_frame(self.h, iid, None, 0)
        if need < 0:
            self._check(int(need), "dcd_frame")
        buf = C.create_string_buffer(need)
        got = lib.pdgpu_dcd_frame(self.h, iid, buf, need)
        if got != need:
            self._check(int(got) if got < 0 else PDGPU_ERR_HIP, "dcd_frame")
        return buf.raw

    def dcd_create(self, path, iid):
        self._check(lib.pdgpu_dcd_create(str(path).encode(), self._natoms[iid]), "dcd_create")

    def dcd_append(self, path, iid):
        self._check(lib.pdgpu_dcd_append(self.h, iid, str(path).encode()), "dcd_append")

    def btf_frame(self, iid, includes=BTF_ALL, backbone=None, state=None):
        """One BTF trajectory record of instance iid (bytes). backbone:
        (nres, 3) instance-local N, CA, C indices (-1 where absent); state:
        btf_state(...) for the energy record."""
        bb = np.ascontiguousarray(backbone if backbone is not None else np.zeros((0, 3)), dtype=np.int32)
        nres = len(bb)
        st = state if state is not None else btf_state()
        args = (self.h, iid, includes, _p(bb.ravel(), _ip), nres, C.byref(st))
        need = lib.pdgpu_btf_frame(*args, None, 0)
        if need < 0:
            self._check(int(need), "btf_frame")
        buf = C.create_string_buffer(need)
        got = lib.pdgpu_btf_frame(*args, buf, need)
        if got != need:
            self._check(int(got) if got < 0 else PDGPU_ERR_HIP, "btf_frame")
        return buf.raw

    def btf_append(self, path, iid, includes=BTF_ALL, backbone=None, state=None):
        bb = np.ascontiguousarray(backbone if backbone is not None else np.zeros((0, 3)), dtype=np.int32)
        st = state if state is not None else btf_state()
        self._check(lib.pdgpu_btf_append(self.h, iid, includes, _p(bb.ravel(), _ip), len(bb), C.byref(st),
                                         str(path).encode()), "btf_append")

    def set_thermostat(self, tau, temperature=300.0):
        """Berendsen thermostat of the PD MD calls (tau in s; 0: off)."""
        self._check(lib.pdgpu_set_thermostat(self.h, tau, temperature), "set_thermostat")

    def set_hydrogens(self, iid, flags):
        """Particle::isHydrogen flags of instance iid (used by Langevin when
        the options' langevin_on_hydrogens is 0)."""
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        if f.size != self._natoms[iid]:
            raise ValueError("one flag per atom")
        self._check(lib.pdgpu_set_hydrogens(self.h, iid, f.tobytes()), "set_hydrogens")

    def traj_dcd(self, iid, path, every):
        """Attach (path, every > 0) or detach (path None) a DCD trajectory
        written by the PD MD calls every `every` steps."""
        self._check(lib.pdgpu_traj_dcd(self.h, iid, None if path is None else str(path).encode(), int(every)),
                    "traj_dcd")

    def profile(self, enable=True):
        self._check(lib.pdgpu_profile(self.h, int(enable)), "profile")

    def profile_read(self):
        n = len(KERNEL_GROUPS)
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        self._check(lib.pdgpu_profile_read(self.h, _p(ms, _dp), cnt.ctypes.data_as(C.POINTER(C.c_int64)), n),
                    "profile_read")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_GROUPS)}

    def autotune(self, iid=PDGPU_ALL, reps=2):
        """Chooses wavefronts per CU for the enumeration, Born, pair and
        chain-rule kernels by timing them; returns the four choices."""
        out = np.zeros(4, np.int32)
        self._check(lib.pdgpu_autotune(self.h, iid, reps, _p(out, _ip)), "autotune")
        return dict(zip(("enum", "born", "pair", "chain"), (int(x) for x in out)))

    def bytes_free(self):
        return lib.pdgpu_bytes_free(self.h)

    def instance_bytes(self, natoms, mid, nsparse=0):
        return lib.pdgpu_instance_bytes(self.h, natoms, mid, nsparse)
