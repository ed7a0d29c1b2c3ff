"""ctypes binding of the CPU oracle (oracle/liboracle.so) for tests.

The oracle is the checker; nothing in the product imports this module.
Arrays use PD's units (Angstrom, e, sqrt(J/molecule), J/molecule, N, kg).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_bp = C.POINTER(C.c_int8)


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "cutoff", "inner_cutoff", "vdw_cutoff", "vdw_inner_cutoff",
        "vdw14_scaling", "elec14_scaling", "dielectric_solute", "dielectric_solvent",
        "still_cutoff", "exp_approx_threshold", "dielectric_offset", "born_radius_offset",
        "P1", "P2", "P3", "P4", "P5", "padding")] + [("elec_mode", C.c_int), ("vdw_mode", C.c_int)]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "epot", "epot_vdw", "epot_elec", "epot_pol", "epot_pol_self", "epot_pol_cross", "virial",
        "epot_bond", "epot_angle", "epot_torsion", "epot_improper")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Bonded(C.Structure):
    _fields_ = [("nbonds", C.c_int), ("nangles", C.c_int), ("ntorsions", C.c_int), ("nimpropers", C.c_int),
                ("bond_ij", _ip), ("bond_lk", _dp), ("angle_iaj", _ip), ("angle_tk", _dp),
                ("tors_iabj", _ip), ("tors_terms", _ip), ("tors_vn", _dp), ("tors_n", _dp), ("tors_gamma", _dp),
                ("impr_iabj", _ip), ("impr_terms", _ip), ("impr_vn", _dp), ("impr_n", _dp), ("impr_gamma", _dp)]


class System(C.Structure):
    _fields_ = [("n", C.c_int), ("charge", _dp), ("radius", _dp), ("epsilon", _dp),
                ("gb_radius", _dp), ("vj", _dp), ("terms123", _dp), ("mass", _dp),
                ("mid", C.c_int), ("band", _bp), ("nsparse", C.c_int),
                ("sp_i", _ip), ("sp_j", _ip), ("sp_o", _ip), ("bonded", C.POINTER(Bonded))]


class NList(C.Structure):
    _fields_ = [("row_start", _ip), ("entries", _up), ("cap", C.c_int64), ("count", C.c_int64)]


def _ensure_built():
    src = os.path.join(ORACLE_DIR, "pd_oracle.c")
    if (not os.path.exists(LIB_PATH)) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])


_ensure_built()
lib = C.CDLL(LIB_PATH)
lib.orc_const.restype = C.c_double
lib.orc_const.argtypes = [C.c_int]
lib.orc_default_params.argtypes = [C.POINTER(Params)]
lib.orc_bondorder.restype = C.c_int
lib.orc_bondorder.argtypes = [C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip,
                              _ip, _bp, _ip, _ip, _ip]
lib.orc_pair_bondorder.restype = C.c_int
lib.orc_pair_bondorder.argtypes = [C.POINTER(System), C.c_int, C.c_int]
lib.orc_nlist_build.restype = C.c_int64
lib.orc_nlist_build.argtypes = [C.POINTER(System), _dp, C.c_double, C.c_double, C.c_int, C.POINTER(NList)]
lib.orc_fixed_terms.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip, _dp, C.c_int, _ip, _ip, _dp,
                                C.POINTER(Params), _dp, _dp]
lib.orc_born_radii.argtypes = [C.POINTER(System), _dp, C.POINTER(NList), C.POINTER(Params), _dp]
lib.orc_gb_calc_forces.argtypes = [C.POINTER(System), _dp, C.POINTER(NList), C.POINTER(Params),
                                   _dp, C.POINTER(Energies), _dp, _dp]
lib.orc_nb_calc_forces.argtypes = [C.POINTER(System), _dp, C.POINTER(NList), C.POINTER(Params),
                                   _dp, C.POINTER(Energies)]
lib.orc_gb_calc_energies.argtypes = [C.POINTER(System), _dp, C.POINTER(NList), C.POINTER(Params),
                                     _dp, C.POINTER(Energies), _dp]
lib.orc_bonded_forces.argtypes = [C.POINTER(Bonded), _dp, _dp, C.POINTER(Energies)]
lib.orc_minimise_sd.restype = C.c_int
lib.orc_minimise_sd.argtypes = [C.POINTER(System), _dp, C.POINTER(Params), C.c_int, C.c_double, C.c_int,
                                _dp, C.POINTER(Energies), _dp]
lib.orc_softvdw_forces.restype = None
lib.orc_softvdw_forces.argtypes = [C.POINTER(System), _dp, C.POINTER(NList), C.c_double, C.c_double, C.c_double,
                                   _dp, C.POINTER(Energies)]
lib.orc_md_vv_pd.restype = C.c_int
lib.orc_md_vv_pd.argtypes = [C.POINTER(System), _dp, _dp, C.POINTER(Params), C.c_int, C.c_double, C.c_int,
                             C.c_double, C.c_double, _dp, _dp]
lib.orc_langevin_coef.restype = C.c_int
lib.orc_langevin_coef.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
lib.orc_langevin_noise.restype = None
lib.orc_langevin_noise.argtypes = [C.c_uint, C.c_int, C.c_int, C.POINTER(C.c_int8), _dp]
lib.orc_md_langevin_pd.restype = C.c_int
lib.orc_md_langevin_pd.argtypes = [C.POINTER(System), _dp, _dp, C.POINTER(Params), C.c_int, C.c_double, C.c_int,
                                   C.c_double, C.c_double, C.c_double, _dp, C.POINTER(C.c_int8), _dp, _dp]
lib.orc_md_beeman_pd.restype = C.c_int
lib.orc_md_beeman_pd.argtypes = [C.POINTER(System), _dp, _dp, _dp, C.POINTER(Params), C.c_int, C.c_double, C.c_int,
                                 C.c_double, C.c_double, _dp, _dp]
lib.orc_md_vv.restype = C.c_int
lib.orc_md_vv.argtypes = [C.POINTER(System), _dp, _dp, C.POINTER(Params), C.c_int, C.c_double, C.c_int,
                          _dp, _dp]

KCAL2J = lib.orc_const(0)
NA = lib.orc_const(1)
ECONV_J = lib.orc_const(3)
HALFECONV = lib.orc_const(6)
KB = lib.orc_const(7)
AMU = lib.orc_const(8)
J_PER_KCALMOL = KCAL2J / NA


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def default_params(**kw):
    p = Params()
    lib.orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class OSystem:
    """Owns numpy arrays and the ctypes view of one instance."""

    def __init__(self, charge, radius, epsilon, gb_radius, vj, terms123, mass, mid, band,
                 sp_i=None, sp_j=None, sp_o=None, bonded=None):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.charge, self.radius, self.epsilon = f(charge), f(radius), f(epsilon)
        self.gb_radius, self.vj, self.terms123 = f(gb_radius), f(vj), f(terms123)
        self.mass = f(mass if mass is not None else np.ones_like(self.charge))
        self.n = len(self.charge)
        self.mid = int(mid)
        self.band = np.ascontiguousarray(band, dtype=np.int8)
        e = np.zeros(0, np.int32)
        self.sp_i = np.ascontiguousarray(sp_i if sp_i is not None else e, dtype=np.int32)
        self.sp_j = np.ascontiguousarray(sp_j if sp_j is not None else e, dtype=np.int32)
        self.sp_o = np.ascontiguousarray(sp_o if sp_o is not None else e, dtype=np.int32)
        self.c = System(self.n, _ptr(self.charge, _dp), _ptr(self.radius, _dp), _ptr(self.epsilon, _dp),
                        _ptr(self.gb_radius, _dp), _ptr(self.vj, _dp), _ptr(self.terms123, _dp),
                        _ptr(self.mass, _dp), self.mid, _ptr(self.band, _bp), len(self.sp_i),
                        _ptr(self.sp_i, _ip), _ptr(self.sp_j, _ip), _ptr(self.sp_o, _ip),
                        C.pointer(bonded.c) if bonded is not None else None)
        self.bonded = bonded  # keeps the term arrays alive


class OBonded:
    """FF_Bonded term lists in the layout of a golden fixture or of
    pdgpu.synth (keys bd_idx, bd_lk, an_idx, an_tk, to_idx, to_terms, to_vn,
    to_n, to_gamma, im_idx, im_terms, im_vn, im_n, im_gamma)."""

    def __init__(self, d):
        fi = lambda k: np.ascontiguousarray(d[k], dtype=np.int32).ravel()
        fd = lambda k: np.ascontiguousarray(d[k], dtype=np.float64).ravel()
        self.arr = {k: (fi(k) if k.endswith(("idx", "terms")) else fd(k)) for k in (
            "bd_idx", "bd_lk", "an_idx", "an_tk", "to_idx", "to_terms", "to_vn", "to_n", "to_gamma",
            "im_idx", "im_terms", "im_vn", "im_n", "im_gamma")}
        a = self.arr
        self.c = Bonded(len(a["bd_lk"]) // 2, len(a["an_tk"]) // 2, len(a["to_terms"]), len(a["im_terms"]),
                        _ptr(a["bd_idx"], _ip), _ptr(a["bd_lk"], _dp), _ptr(a["an_idx"], _ip),
                        _ptr(a["an_tk"], _dp), _ptr(a["to_idx"], _ip), _ptr(a["to_terms"], _ip),
                        _ptr(a["to_vn"], _dp), _ptr(a["to_n"], _dp), _ptr(a["to_gamma"], _dp),
                        _ptr(a["im_idx"], _ip), _ptr(a["im_terms"], _ip), _ptr(a["im_vn"], _dp),
                        _ptr(a["im_n"], _dp), _ptr(a["im_gamma"], _dp))


def bonded_forces(bonded, xyz):
    """FF_Bonded::calcForces: (energies dict, forces (n,3) in N)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).ravel()
    f = np.zeros_like(xyz)
    e = Energies()
    lib.orc_bonded_forces(C.byref(bonded.c), _ptr(xyz, _dp), _ptr(f, _dp), C.byref(e))
    return e.as_dict(), f.reshape(-1, 3)


def nlist_build(sysm, xyz, list_cutoff, padding, shadow=True):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rs = np.zeros(sysm.n + 1, np.int32)
    cap = max(1024, sysm.n * 64)
    while True:
        ent = np.zeros(cap, np.uint32)
        nl = NList(_ptr(rs, _ip), _ptr(ent, _up), cap, 0)
        need = lib.orc_nlist_build(C.byref(sysm.c), _ptr(xyz, _dp), list_cutoff, padding, int(shadow), C.byref(nl))
        if need == 0:
            return rs.copy(), ent[:nl.count].copy()
        cap = int(need)


def _nl(rs, ent):
    rs = np.ascontiguousarray(rs, dtype=np.int32)
    ent = np.ascontiguousarray(ent, dtype=np.uint32)
    return NList(_ptr(rs, _ip), _ptr(ent, _up), len(ent), len(ent)), (rs, ent)


def born_radii(sysm, xyz, rs, ent, params):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nl, keep = _nl(rs, ent)
    out = np.zeros(sysm.n)
    lib.orc_born_radii(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(nl), C.byref(params), _ptr(out, _dp))
    return out


def gb_calc_forces(sysm, xyz, rs, ent, params):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nl, keep = _nl(rs, ent)
    f = np.zeros(3 * sysm.n)
    born = np.zeros(sysm.n)
    deda = np.zeros(sysm.n)
    e = Energies()
    lib.orc_gb_calc_forces(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(nl), C.byref(params), _ptr(f, _dp),
                           C.byref(e), _ptr(born, _dp), _ptr(deda, _dp))
    return e.as_dict(), f.reshape(-1, 3), born, deda


def nb_calc_forces(sysm, xyz, rs, ent, params):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nl, keep = _nl(rs, ent)
    f = np.zeros(3 * sysm.n)
    e = Energies()
    lib.orc_nb_calc_forces(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(nl), C.byref(params), _ptr(f, _dp), C.byref(e))
    return e.as_dict(), f.reshape(-1, 3)


def gb_calc_energies(sysm, xyz, rs, ent, params):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nl, keep = _nl(rs, ent)
    f = np.zeros(3 * sysm.n)
    born = np.zeros(sysm.n)
    e = Energies()
    lib.orc_gb_calc_energies(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(nl), C.byref(params), _ptr(f, _dp),
                             C.byref(e), _ptr(born, _dp))
    return e.as_dict(), f.reshape(-1, 3), born


def fixed_terms(gb_radius, a12, b12, d12, a13, b13, d13, params):
    gb_radius = np.ascontiguousarray(gb_radius, dtype=np.float64)
    n = len(gb_radius)
    a12, b12, a13, b13 = (np.ascontiguousarray(x, dtype=np.int32) for x in (a12, b12, a13, b13))
    d12, d13 = (np.ascontiguousarray(x, dtype=np.float64) for x in (d12, d13))
    vj = np.zeros(n)
    t = np.zeros(n)
    lib.orc_fixed_terms(n, _ptr(gb_radius, _dp), len(a12), _ptr(a12, _ip), _ptr(b12, _ip), _ptr(d12, _dp),
                        len(a13), _ptr(a13, _ip), _ptr(b13, _ip), _ptr(d13, _dp), C.byref(params),
                        _ptr(vj, _dp), _ptr(t, _dp))
    return vj, t


def bondorder(resid, a12, b12, a13, b13, a14, b14):
    resid = np.ascontiguousarray(resid, dtype=np.int32)
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in (a12, b12, a13, b13, a14, b14)]
    n = len(resid)
    mid = np.zeros(1, np.int32)
    lib.orc_bondorder(n, _ptr(resid, _ip), len(arrs[0]), _ptr(arrs[0], _ip), _ptr(arrs[1], _ip),
                      len(arrs[2]), _ptr(arrs[2], _ip), _ptr(arrs[3], _ip),
                      len(arrs[4]), _ptr(arrs[4], _ip), _ptr(arrs[5], _ip), _ptr(mid, _ip), None, None, None, None)
    m = int(mid[0])
    band = np.zeros(max(0, m * (n - 1)), np.int8)
    cap = len(arrs[0]) + len(arrs[2]) + len(arrs[4]) + 1
    spi, spj, spo = (np.zeros(cap, np.int32) for _ in range(3))
    nsp = lib.orc_bondorder(n, _ptr(resid, _ip), len(arrs[0]), _ptr(arrs[0], _ip), _ptr(arrs[1], _ip),
                            len(arrs[2]), _ptr(arrs[2], _ip), _ptr(arrs[3], _ip),
                            len(arrs[4]), _ptr(arrs[4], _ip), _ptr(arrs[5], _ip), _ptr(mid, _ip),
                            _ptr(band, _bp), _ptr(spi, _ip), _ptr(spj, _ip), _ptr(spo, _ip))
    return m, band, spi[:nsp].copy(), spj[:nsp].copy(), spo[:nsp].copy()


def minimise_sd(sysm, xyz, params, steps, step_size=2e5, update_nlist=1):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    trace = np.zeros(max(steps, 1))
    e = Energies()
    f = np.zeros(3 * sysm.n)
    rc = lib.orc_minimise_sd(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(params), steps, step_size, update_nlist,
                             _ptr(trace, _dp), C.byref(e), _ptr(f, _dp))
    return rc, xyz.reshape(-1, 3), trace[:steps], e.as_dict(), f.reshape(-1, 3)


def md_vv(sysm, xyz, vel, params, steps, dt, update_nlist=10):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    te = np.zeros(steps)
    tk = np.zeros(steps)
    rc = lib.orc_md_vv(C.byref(sysm.c), _ptr(xyz, _dp), _ptr(vel, _dp), C.byref(params), steps, dt, update_nlist,
                       _ptr(te, _dp), _ptr(tk, _dp))
    return rc, xyz.reshape(-1, 3), vel.reshape(-1, 3), te, tk


def md_vv_pd(sysm, xyz, vel, params, steps, dt, update_nlist=10, tau=0.0, temperature=300.0):
    """PD's MolecularDynamics with its VelocityVerlet update as written."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    te = np.zeros(steps)
    tk = np.zeros(steps)
    rc = lib.orc_md_vv_pd(C.byref(sysm.c), _ptr(xyz, _dp), _ptr(vel, _dp), C.byref(params), steps, dt,
                          update_nlist, tau, temperature, _ptr(te, _dp), _ptr(tk, _dp))
    return rc, xyz.reshape(-1, 3), vel.reshape(-1, 3), te, tk


def langevin_coef(gamma, dt):
    """(branch, c0, c1, c2, varv, varr, crv) of PD's Langevin update."""
    v = [C.c_double() for _ in range(6)]
    br = lib.orc_langevin_coef(gamma, dt, *[C.byref(x) for x in v])
    return (br,) + tuple(x.value for x in v)


def _i8(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int8)


def langevin_noise(seed, natoms, steps, skip=None):
    """The reference's normal deviates after srand(seed), [steps][natoms][6]
    (atoms with skip set draw none)."""
    out = np.zeros((steps, natoms, 6))
    sk = _i8(skip)
    lib.orc_langevin_noise(seed, natoms, steps, None if sk is None else _ptr(sk, C.POINTER(C.c_int8)),
                           _ptr(out, _dp))
    return out


def md_langevin_pd(sysm, xyz, vel, params, steps, dt, temperature, gamma, noise, update_nlist=10, hydrogen=None,
                   tau=0.0):
    """PD's MolecularDynamics with its Langevin integrator (hydrogen: the
    LangevinOnHydrogens = false variant's hydrogen flags; tau: Berendsen,
    which acts only in PD's VelocityVerlet fallback below gamma dt 1e-5)."""
    hy = _i8(hydrogen)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    noise = np.ascontiguousarray(noise, dtype=np.float64)
    te = np.zeros(steps)
    tk = np.zeros(steps)
    rc = lib.orc_md_langevin_pd(C.byref(sysm.c), _ptr(xyz, _dp), _ptr(vel, _dp), C.byref(params), steps, dt,
                                update_nlist, temperature, gamma, tau, _ptr(noise, _dp),
                                None if hy is None else _ptr(hy, C.POINTER(C.c_int8)), _ptr(te, _dp), _ptr(tk, _dp))
    return rc, xyz.reshape(-1, 3), vel.reshape(-1, 3), te, tk


def md_beeman_pd(sysm, xyz, vel, params, steps, dt, f_old=None, update_nlist=10, tau=0.0, temperature=300.0):
    """PD's MolecularDynamics with its Beeman integrator; f_old (N) in/out."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).copy()
    fo = np.zeros(xyz.size) if f_old is None else np.ascontiguousarray(f_old, dtype=np.float64).ravel().copy()
    te = np.zeros(steps)
    tk = np.zeros(steps)
    rc = lib.orc_md_beeman_pd(C.byref(sysm.c), _ptr(xyz, _dp), _ptr(vel, _dp), _ptr(fo, _dp), C.byref(params),
                              steps, dt, update_nlist, tau, temperature, _ptr(te, _dp), _ptr(tk, _dp))
    return rc, xyz.reshape(-1, 3), vel.reshape(-1, 3), fo.reshape(-1, 3), te, tk


def softvdw_forces(sysm, xyz, rs, ent, hardness=12.0, olap=0.6, vdw_cutoff=15.0):
    """FF_SoftVDW::calcForces_Soft over a shadow list (PD units)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nl, keep = _nl(rs, ent)
    f = np.zeros(3 * sysm.n)
    e = Energies()
    lib.orc_softvdw_forces(C.byref(sysm.c), _ptr(xyz, _dp), C.byref(nl), hardness, olap, vdw_cutoff, _ptr(f, _dp),
                           C.byref(e))
    return e.as_dict(), f.reshape(-1, 3)
