"""Synthetic peptide-like molecules for benchmarks and large-size tests.

Shapes follow the workloads PD is run on (polypeptides built residue by
residue with an all-atom forcefield, ~16 atoms per residue), but every
number here is this module's own: residue templates are generic
(backbone N-H, CA-HA, C=O plus a side chain of aliphatic carbons ending in
CH3 / OH / SH / COO- / NH3+ / a six-ring), parameters are round literature
values for C/N/O/S/H (LJ R_min/2 and epsilon, Still-type GB radii), and the
backbone uses textbook bond lengths and angles. Nothing is read from PD's
parameter files.

Output per molecule: a dict with PD-unit arrays (xyz A, charge e, radius
A, epsilon sqrt(J/molecule), gb_radius A, mass kg), the 1-2/1-3/1-4 pair
lists with equilibrium distances, PD's bond-order band (mid, band, sparse)
and the Still fixed terms (vj, terms123).
"""
import numpy as np

from . import bondorder_build, fixed_terms, default_options, J_PER_KCALMOL

AMU = 1.66053878e-27
DEG = np.pi / 180.0

# element classes: R_min/2 (A), epsilon (kcal/mol), GB radius (A), mass (amu)
ELEM = {
    "C": (1.90, 0.086, 1.70, 12.011),
    "N": (1.82, 0.170, 1.55, 14.007),
    "O": (1.66, 0.210, 1.50, 15.999),
    "S": (2.00, 0.250, 1.80, 32.06),
    "HN": (0.60, 0.0157, 1.10, 1.008),   # polar H
    "HC": (1.49, 0.0157, 1.20, 1.008),   # aliphatic H
}

# side chain: (number of chain carbons, terminal group, net charge)
RESIDUES = {
    "G": (0, "H", 0), "A": (1, "CH3", 0), "S": (1, "OH", 0), "C": (1, "SH", 0), "T": (2, "OH", 0),
    "V": (2, "CH3", 0), "P": (3, "CH3", 0), "L": (3, "CH3", 0), "I": (3, "CH3", 0), "M": (3, "SH", 0),
    "N": (2, "OH", 0), "D": (1, "COO", -1), "Q": (3, "OH", 0), "E": (2, "COO", -1), "K": (4, "NH3", 1),
    "R": (5, "NH3", 1), "H": (1, "RING", 0), "F": (1, "RING", 0), "Y": (2, "RING", 0), "W": (2, "RING", 0),
}
LETTERS = "ACDEFGHIKLMNPQRSTVWY"

# backbone geometry (A, degrees)
B_NCA, B_CAC, B_CN, B_CO, B_NH, B_CH, B_CC = 1.458, 1.525, 1.329, 1.229, 1.01, 1.09, 1.53
A_NCAC, A_CACN, A_CNCA, A_CACO, A_CANH = 111.2, 116.2, 121.7, 120.5, 119.0


def _place(a, b, c, bond, angle_deg, torsion_deg):
    """Natural-extension placement of d bonded to c (|cd|=bond, angle bcd,
    torsion abcd). Works on (..., 3) arrays."""
    bc = c - b
    bc = bc / np.linalg.norm(bc, axis=-1, keepdims=True)
    n = np.cross(b - a, bc)
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    m = np.cross(n, bc)
    th = np.asarray(angle_deg) * DEG
    ph = np.asarray(torsion_deg) * DEG
    d0 = -bond * np.cos(th)
    d1 = bond * np.sin(th) * np.cos(ph)
    d2 = bond * np.sin(th) * np.sin(ph)
    return c + bc * d0[..., None] + m * d1[..., None] + n * d2[..., None]


class _Template:
    """HA and side-chain atoms of one residue type in the (N, CA, C) frame.
    Local atom index k >= 4 is the (k-4)-th side atom; 2 is CA."""

    def __init__(self, letter):
        nc, term, charge = RESIDUES[letter]
        N = np.array([0.0, 0.0, 0.0])
        CA = np.array([B_NCA, 0.0, 0.0])
        C = _place(np.array([0.0, 1.0, 0.0]), N, CA, B_CAC, A_NCAC, 180.0)
        elems, pos, bonds, q = [], [], [], []

        def add(el, p, bond_to, charge_):
            elems.append(el)
            pos.append(np.asarray(p, dtype=float))
            q.append(charge_)
            k = len(elems) - 1 + 4
            if bond_to is not None:
                bonds.append((bond_to, k))
            return k

        def P(k):
            return pos[k - 4]

        self.ha_pos = _place(C, N, CA, B_CH, 109.5, -120.0)
        if nc == 0:
            add("HC", _place(C, N, CA, B_CH, 109.5, 120.0), 2, 0.06)
        else:
            prev2, prev, cur = C, N, CA
            last = 2
            chain = []
            tors = [120.0, -60.0, 180.0, 180.0, 180.0, 180.0]
            for k in range(nc):
                p = _place(prev2, prev, cur, B_CC, 110.5, tors[k])
                idx = add("C", p, last, 0.0)
                chain.append((idx, prev, cur))
                prev2, prev, cur = prev, cur, p
                last = idx
            for k, (idx, pp2, pp) in enumerate(chain):
                # hydrogens staggered around the torsion of the next heavy atom
                if k < nc - 1:
                    nxt = tors[k + 1]
                else:
                    nxt = {"CH3": 60.0, "RING": 90.0}.get(term, 180.0)
                hs = (nxt, nxt + 120.0, nxt + 240.0) if term == "CH3" and k == nc - 1 else (nxt + 120.0, nxt + 240.0)
                for h in hs:
                    add("HC", _place(pp2, pp, P(idx), B_CH, 109.5, h), idx, 0.06)
                    q[idx - 4] -= 0.06
            a_, b_, c_ = prev2, prev, cur
            if term == "OH":
                o = add("O", _place(a_, b_, c_, 1.43, 109.5, 180.0), last, -0.65)
                add("HN", _place(b_, c_, P(o), 0.96, 108.5, 180.0), o, 0.42)
                q[last - 4] += 0.23
            elif term == "SH":
                s_ = add("S", _place(a_, b_, c_, 1.81, 114.0, 180.0), last, -0.31)
                add("HN", _place(b_, c_, P(s_), 1.34, 96.0, 180.0), s_, 0.19)
                q[last - 4] += 0.12
            elif term == "COO":
                c2 = add("C", _place(a_, b_, c_, B_CC, 112.0, 180.0), last, 0.70)
                add("O", _place(b_, c_, P(c2), 1.25, 118.0, 90.0), c2, -0.85)
                add("O", _place(b_, c_, P(c2), 1.25, 118.0, -90.0), c2, -0.85)
            elif term == "NH3":
                n3 = add("N", _place(a_, b_, c_, 1.47, 111.0, 180.0), last, -0.30)
                for h in range(3):
                    add("HN", _place(b_, c_, P(n3), 1.01, 109.5, 60.0 + 120.0 * h), n3, 0.33)
                q[last - 4] += 0.31
            elif term == "RING":
                # planar six-ring (side 1.39 A) attached through r0
                r0p = _place(a_, b_, c_, 1.51, 114.0, 90.0)
                u = (r0p - c_) / np.linalg.norm(r0p - c_)
                w = np.cross(u, b_ - c_)
                w = w / np.linalg.norm(w)
                v = np.cross(w, u)
                ctr = r0p + 1.39 * u
                ring = []
                for k in range(6):
                    ang = np.pi / 3 * k
                    rk = ctr + 1.39 * (-u * np.cos(ang) + v * np.sin(ang))
                    ring.append(add("C", rk, ring[-1] if ring else last, 0.0 if k == 0 else -0.115))
                bonds.append((ring[-1], ring[0]))  # ring closure
                for k in range(1, 6):
                    d = P(ring[k]) - ctr
                    add("HC", ctr + (1.39 + 1.08) * d / np.linalg.norm(d), ring[k], 0.115)
        self.elems = elems
        self.charges = np.array(q, dtype=float)
        self.bonds = bonds
        self.net = charge
        self.R0 = _frame(N[None], CA[None], C[None])[0]
        self.local_ha = (self.ha_pos - CA) @ self.R0
        self.local = (np.array(pos).reshape(-1, 3) - CA) @ self.R0 if pos else np.zeros((0, 3))
        self.nside = len(elems)
        self.natoms = 6 + self.nside


def _frame(N, CA, C):
    e1 = C - CA
    e1 = e1 / np.linalg.norm(e1, axis=-1, keepdims=True)
    v = N - CA
    e2 = v - (v * e1).sum(-1, keepdims=True) * e1
    e2 = e2 / np.linalg.norm(e2, axis=-1, keepdims=True)
    e3 = np.cross(e1, e2)
    return np.stack([e1, e2, e3], axis=-1)  # columns are axes: world = R @ local


_TEMPLATES = {}


def template(letter):
    if letter not in _TEMPLATES:
        _TEMPLATES[letter] = _Template(letter)
    return _TEMPLATES[letter]


def splitmix64(seed, n):
    """n uniform 64-bit integers from a SplitMix64 stream."""
    out = np.empty(n, dtype=np.uint64)
    x = np.uint64(seed)
    with np.errstate(over="ignore"):
        for k in range(n):
            x = np.uint64(x + np.uint64(0x9E3779B97F4A7C15))
            z = x
            z = np.uint64((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
            z = np.uint64((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
            out[k] = z ^ (z >> np.uint64(31))
    return out


def random_sequence(seed, length):
    z = splitmix64(seed, length)
    return "".join(LETTERS[int(v % np.uint64(20))] for v in z)


def _residue_atoms(t, Ni, CAi, Ci, phi_i, psi_i, first):
    R = _frame(Ni[None], CAi[None], Ci[None])[0]
    side = CAi + t.local @ R.T
    ha = CAi + t.local_ha @ R.T
    # H on N and O on C lie in the peptide planes, so they follow phi / psi
    h = _place(Ci, CAi, Ni, B_NH, A_CANH, (0.0 if first else phi_i) + 180.0)
    o = _place(Ni, CAi, Ci, B_CO, A_CACO, psi_i + 180.0)
    return np.array([Ni, h, CAi, ha] + list(side) + [Ci, o])


def build(seq, phi=None, psi=None, seed=0, mode="extended", min_dist=2.0, tries=12, relax_iters=400):
    """Coordinates, parameters and topology of one molecule. Atom order per
    residue: N, H, CA, HA, side chain, C, O (PD's all-atom order).

    mode 'extended' places every residue at phi = psi = 180 (PD's builder).
    mode 'basins' samples phi/psi per residue and re-samples a residue
    whose atoms come closer than min_dist to atoms two or more residues
    back (or to the previous side chain), then relieves the remaining
    clashes with a soft-core push-apart relaxation (relax_clashes): the
    stand-in for PD's soft-core clash-relief minimisation."""
    L = len(seq)
    if mode == "extended" and phi is None:
        return _build_extended(seq)
    rng = np.random.default_rng(seed)
    if phi is None:
        if mode == "extended":
            phi, psi = np.full(L, 180.0), np.full(L, 180.0)
        else:
            phi, psi = np.zeros(L), np.zeros(L)
    phi = np.array(phi, dtype=float)
    psi = np.array(psi, dtype=float)

    def sample():
        if rng.random() < 0.6:
            return -63.0 + rng.normal(0, 15.0), -43.0 + rng.normal(0, 15.0)
        return -120.0 + rng.normal(0, 15.0), 130.0 + rng.normal(0, 15.0)

    placed = []        # per residue: atom array
    side_of = []       # per residue: slice of side-chain atoms
    N0 = np.array([0.0, 0.0, 0.0])
    CA0 = np.array([B_NCA, 0.0, 0.0])
    for i, letter in enumerate(seq):
        t = template(letter)
        older = np.concatenate(placed[:-1]) if len(placed) > 1 else np.zeros((0, 3))
        prev_side = placed[-1][4:4 + template(seq[i - 1]).nside] if i > 0 else np.zeros((0, 3))
        best, best_d = None, -1.0
        for attempt in range(tries if mode != "extended" else 1):
            if mode != "extended":
                phi[i], psi[i] = sample()
            if i == 0:
                Ni, CAi = N0, CA0
                Ci = _place(np.array([0.0, 1.0, 0.0]), Ni, CAi, B_CAC, A_NCAC, 180.0)
            else:
                Np, CAp, Cp = placed[-1][0], placed[-1][2], placed[-1][-2]
                Ni = _place(Np, CAp, Cp, B_CN, A_CACN, psi[i - 1])
                CAi = _place(CAp, Cp, Ni, B_NCA, A_CNCA, 180.0)
                Ci = _place(Cp, Ni, CAi, B_CAC, A_NCAC, phi[i])
            atoms = _residue_atoms(t, Ni, CAi, Ci, phi[i], psi[i], i == 0)
            if mode == "extended":
                best = atoms
                break
            ref = np.concatenate([older, prev_side])
            if len(ref):
                near = ref[((ref - CAi) ** 2).sum(-1) < 14.0 ** 2]
                d = np.sqrt(((atoms[:, None, :] - near[None, :, :]) ** 2).sum(-1)).min() if len(near) else 99.0
            else:
                d = 99.0
            if d > best_d:
                best, best_d, keep = atoms, d, (phi[i], psi[i])
            if d >= min_dist:
                break
        if mode != "extended":
            phi[i], psi[i] = keep
            # O of the previous residue follows psi[i-1], which is fixed already
        placed.append(best)
    xyz, el, q, bonds, res_of = [], [], [], [], []
    prev_c = None
    off = 0
    for i, letter in enumerate(seq):
        t = template(letter)
        c_i = off + 4 + t.nside
        xyz += list(placed[i])
        el += ["N", "HN", "C", "HC"] + t.elems + ["C", "O"]
        ca_q = 0.03 - (0.06 if letter == "G" else 0.0)
        q += [-0.42, 0.27, ca_q, 0.08] + list(t.charges) + [0.60, -0.56]
        bonds += [(off, off + 1), (off, off + 2), (off + 2, off + 3), (off + 2, c_i), (c_i, c_i + 1)]
        bonds += [(off + a, off + b) for a, b in t.bonds]
        if prev_c is not None:
            bonds.append((prev_c, off))
        prev_c = c_i
        res_of += [i] * t.natoms
        off += t.natoms
    xyz = np.array(xyz, dtype=float)
    bonds = np.array(bonds, dtype=np.int64)
    if mode != "extended" and relax_iters > 0:
        xyz = relax_clashes(xyz, bonds, min_dist, relax_iters)
    return _finish(xyz, el, np.array(q, dtype=float), np.array(res_of, dtype=np.int32), bonds)


_EXT = {}


def _extended_frames():
    """Residue-frame step of the phi = psi = omega = 180 chain: the (N, CA, C)
    frame of residue i+1 is a fixed rigid transform of residue i's."""
    if "step" not in _EXT:
        N0 = np.array([0.0, 0.0, 0.0])
        CA0 = np.array([B_NCA, 0.0, 0.0])
        C0 = _place(np.array([0.0, 1.0, 0.0]), N0, CA0, B_CAC, A_NCAC, 180.0)
        N1 = _place(N0, CA0, C0, B_CN, A_CACN, 180.0)
        CA1 = _place(CA0, C0, N1, B_NCA, A_CNCA, 180.0)
        C1 = _place(C0, N1, CA1, B_CAC, A_NCAC, 180.0)
        R0 = _frame(N0[None], CA0[None], C0[None])[0]
        R1 = _frame(N1[None], CA1[None], C1[None])[0]
        # world = R @ local + CA; residue 1 = step applied to residue 0
        _EXT["R0"], _EXT["t0"] = R0, CA0
        _EXT["dR"] = R1 @ R0.T
        _EXT["dt"] = CA1 - _EXT["dR"] @ CA0
    return _EXT


def _ext_local(letter):
    """All atoms of a residue type in its own frame for the extended chain."""
    key = "local_" + letter
    if key not in _EXT:
        ex = _extended_frames()
        t = template(letter)
        N0 = np.array([0.0, 0.0, 0.0])
        CA0 = np.array([B_NCA, 0.0, 0.0])
        C0 = _place(np.array([0.0, 1.0, 0.0]), N0, CA0, B_CAC, A_NCAC, 180.0)
        atoms = _residue_atoms(t, N0, CA0, C0, 180.0, 180.0, False)
        _EXT[key] = (atoms - ex["t0"]) @ ex["R0"]
    return _EXT[key]


def _build_extended(seq):
    ex = _extended_frames()
    L = len(seq)
    Rs = np.empty((L, 3, 3))
    ts = np.empty((L, 3))
    Rs[0], ts[0] = ex["R0"], ex["t0"]
    for i in range(1, L):
        Rs[i] = ex["dR"] @ Rs[i - 1]
        ts[i] = ex["dR"] @ ts[i - 1] + ex["dt"]
    xyz, el, q, bonds, res_of = [], [], [], [], []
    prev_c = None
    off = 0
    for i, letter in enumerate(seq):
        t = template(letter)
        xyz.append(_ext_local(letter) @ Rs[i].T + ts[i])
        c_i = off + 4 + t.nside
        el += ["N", "HN", "C", "HC"] + t.elems + ["C", "O"]
        ca_q = 0.03 - (0.06 if letter == "G" else 0.0)
        q += [-0.42, 0.27, ca_q, 0.08] + list(t.charges) + [0.60, -0.56]
        bonds += [(off, off + 1), (off, off + 2), (off + 2, off + 3), (off + 2, c_i), (c_i, c_i + 1)]
        bonds += [(off + a, off + b) for a, b in t.bonds]
        if prev_c is not None:
            bonds.append((prev_c, off))
        prev_c = c_i
        res_of += [i] * t.natoms
        off += t.natoms
    return _finish(np.concatenate(xyz), el, np.array(q, dtype=float), np.array(res_of, dtype=np.int32),
                   np.array(bonds, dtype=np.int64))


def relax_clashes(xyz, bonds, min_dist=2.0, iters=100):
    """Pushes apart atom pairs closer than min_dist that are not 1-2 / 1-3
    partners, and pulls 1-2 / 1-3 pairs back towards their built lengths."""
    from scipy.spatial import cKDTree
    n = len(xyz)
    p12, p13, _ = _pairs(n, bonds)
    excl = np.concatenate([p12, p13])
    ekey = np.sort(excl[:, 0].astype(np.int64) * n + excl[:, 1])
    x = xyz.copy()
    ref12 = np.linalg.norm(xyz[p12[:, 0]] - xyz[p12[:, 1]], axis=1)
    ref13 = np.linalg.norm(xyz[p13[:, 0]] - xyz[p13[:, 1]], axis=1)
    for _ in range(iters):
        pr = cKDTree(x).query_pairs(min_dist, output_type="ndarray")
        if len(pr):
            key = pr[:, 0].astype(np.int64) * n + pr[:, 1]
            pos = np.searchsorted(ekey, key)
            pos[pos >= len(ekey)] = 0
            pr = pr[ekey[pos] != key]
        if len(pr) == 0:
            break
        dv = x[pr[:, 0]] - x[pr[:, 1]]
        d = np.linalg.norm(dv, axis=1)
        d = np.maximum(d, 1e-3)
        push = (0.5 * (min_dist - d) / d)[:, None] * dv
        mv = np.zeros_like(x)
        np.add.at(mv, pr[:, 0], push)
        np.add.at(mv, pr[:, 1], -push)
        x += 0.5 * mv
        for pp, ref in ((p12, ref12), (p13, ref13)):
            dv = x[pp[:, 0]] - x[pp[:, 1]]
            d = np.maximum(np.linalg.norm(dv, axis=1), 1e-6)
            corr = (0.5 * (ref - d) / d)[:, None] * dv
            mv = np.zeros_like(x)
            np.add.at(mv, pp[:, 0], corr)
            np.add.at(mv, pp[:, 1], -corr)
            x += 0.5 * mv
    return x


def _pairs(n, bonds):
    """1-2, 1-3, 1-4 pair lists (i<j), each pair in its lowest order only."""
    nb = [[] for _ in range(n)]
    for a, b in bonds:
        nb[a].append(b)
        nb[b].append(a)
    p12 = {tuple(sorted((int(a), int(b)))) for a, b in bonds}
    p13 = set()
    for c in range(n):
        ns = nb[c]
        for x in range(len(ns)):
            for y in range(x + 1, len(ns)):
                p = tuple(sorted((ns[x], ns[y])))
                if p not in p12:
                    p13.add(p)
    p14 = set()
    for b, c in p12:
        for a in nb[b]:
            if a == c:
                continue
            for d in nb[c]:
                if d == b or d == a:
                    continue
                p = tuple(sorted((a, d)))
                if p not in p12 and p not in p13:
                    p14.add(p)
    f = lambda s: np.array(sorted(s), dtype=np.int32).reshape(-1, 2)
    return f(p12), f(p13), f(p14)


# FF_Bonded parameters of the generic forcefield (kcal/mol, A, rad): bond
# and angle force constants by whether a hydrogen takes part; equilibrium
# lengths and angles are the built geometry's. Torsions: amide C-N bonds a
# 2-fold term, every other rotatable bond a weak 3-fold term; impropers a
# 2-fold term on 3-coordinated C and N. Each dihedral's phase puts its
# minimum at the built angle (gamma = pi - n*phi0), the way PD's
# FF_Bonded::resetEquilibrium* take the current structure as equilibrium.
K_BOND_HEAVY, K_BOND_H = 300.0, 340.0
K_ANGLE_HEAVY, K_ANGLE_H = 60.0, 35.0
V_AMIDE, V_SP3 = 2.5, 0.15
V_IMPROPER_C, V_IMPROPER_N = 10.5, 1.1


def bonded_terms(xyz, el, bonds):
    """FF_Bonded term lists (BONDED_KEYS layout, PD units) for a molecule."""
    n = len(xyz)
    kj = J_PER_KCALMOL
    nb = [[] for _ in range(n)]
    for a, b in bonds:
        nb[a].append(int(b))
        nb[b].append(int(a))
    nb = [sorted(v) for v in nb]
    isH = np.array([e.startswith("H") for e in el])
    p12 = np.array(sorted({tuple(sorted((int(a), int(b)))) for a, b in bonds}), dtype=np.int32).reshape(-1, 2)
    l = np.linalg.norm(xyz[p12[:, 0]] - xyz[p12[:, 1]], axis=1)
    kb = np.where(isH[p12[:, 0]] | isH[p12[:, 1]], K_BOND_H, K_BOND_HEAVY) * kj
    ang = [(i, c, j) for c in range(n) for x, i in enumerate(nb[c]) for j in nb[c][x + 1:]]
    ang = np.array(ang, dtype=np.int32).reshape(-1, 3)
    u = xyz[ang[:, 0]] - xyz[ang[:, 1]]
    v = xyz[ang[:, 2]] - xyz[ang[:, 1]]
    th0 = np.arccos(np.clip((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1), -1, 1))
    ka = np.where(isH[ang[:, 0]] | isH[ang[:, 2]], K_ANGLE_H, K_ANGLE_HEAVY) * kj
    amide_c = {c for c in range(n) if el[c] == "C" and any(el[o] == "O" and len(nb[o]) == 1 for o in nb[c])}
    tors, tv, tn = [], [], []
    for b, c in p12:
        amide = (b in amide_c and el[c] == "N") or (c in amide_c and el[b] == "N")
        for a in nb[b]:
            if a == c:
                continue
            for d in nb[c]:
                if d == b or d == a:
                    continue
                tors.append((a, b, c, d))
                tv.append(V_AMIDE if amide else V_SP3)
                tn.append(2.0 if amide else 3.0)
    m = len(tors)
    z4 = lambda vals: np.concatenate([np.asarray(vals, dtype=float).reshape(-1, 1), np.zeros((len(vals), 3))], 1)
    imp, iv = [], []
    for c in range(n):
        if len(nb[c]) == 3 and el[c] in ("C", "N"):
            x, y, w = nb[c]
            imp.append((x, y, c, w))
            iv.append(V_IMPROPER_C if el[c] == "C" else V_IMPROPER_N)
    k = len(imp)
    tors = np.array(tors, dtype=np.int32).reshape(-1, 4)
    imp = np.array(imp, dtype=np.int32).reshape(-1, 4)
    tn = np.asarray(tn, dtype=float)
    tg = np.pi - tn * dihedral(xyz, tors)
    ig = np.pi - 2.0 * dihedral(xyz, imp)
    return {"bd_idx": p12.ravel(), "bd_lk": np.stack([l, kb], 1).ravel(),
            "an_idx": ang.ravel(), "an_tk": np.stack([th0, ka], 1).ravel(),
            "to_idx": tors.ravel(), "to_terms": np.ones(m, np.int32),
            "to_vn": (z4(tv) * kj).ravel(), "to_n": z4(tn).ravel(), "to_gamma": z4(tg).ravel(),
            "im_idx": imp.ravel(), "im_terms": np.ones(k, np.int32),
            "im_vn": (z4(iv) * kj).ravel(), "im_n": z4([2.0] * k).ravel(), "im_gamma": z4(ig).ravel()}


def dihedral(xyz, q):
    """IUPAC dihedral angles (rad) of the quartets q (m, 4)."""
    b1 = xyz[q[:, 1]] - xyz[q[:, 0]]
    b2 = xyz[q[:, 2]] - xyz[q[:, 1]]
    b3 = xyz[q[:, 3]] - xyz[q[:, 2]]
    m = np.cross(b1, b2)
    n = np.cross(b2, b3)
    return np.arctan2(np.linalg.norm(b2, axis=1) * (b1 * n).sum(1), (m * n).sum(1))


def _finish(xyz, el, q, resid, bonds):
    n = len(xyz)
    p12, p13, p14 = _pairs(n, bonds)
    d12 = np.linalg.norm(xyz[p12[:, 0]] - xyz[p12[:, 1]], axis=1)
    d13 = np.linalg.norm(xyz[p13[:, 0]] - xyz[p13[:, 1]], axis=1)
    rad = np.array([ELEM[e][0] for e in el])
    eps_kcal = np.array([ELEM[e][1] for e in el])
    gbr = np.array([ELEM[e][2] for e in el])
    mass = np.array([ELEM[e][3] for e in el]) * AMU
    eps = np.sqrt(eps_kcal * J_PER_KCALMOL)  # PD stores sqrt(J/molecule)
    mid, band, si, sj, so = bondorder_build(resid, p12[:, 0], p12[:, 1], p13[:, 0], p13[:, 1], p14[:, 0], p14[:, 1])
    vj, t123 = fixed_terms(gbr, p12[:, 0], p12[:, 1], d12, p13[:, 0], p13[:, 1], d13, default_options())
    return {"xyz": xyz, "charge": q, "radius": rad, "epsilon": eps, "gb_radius": gbr, "mass": mass,
            "resid": resid, "a12": p12[:, 0], "b12": p12[:, 1], "d12": d12, "a13": p13[:, 0], "b13": p13[:, 1],
            "d13": d13, "a14": p14[:, 0], "b14": p14[:, 1], "mid": mid, "band": band, "sp_i": si, "sp_j": sj,
            "sp_o": so, "vj": vj, "terms123": t123, "n": n, "el": list(el), "bonds": np.asarray(bonds)}


def peptide(seed, residues, mode="extended", bonded=False):
    """One molecule of `residues` residues with a seeded random sequence;
    bonded=True adds FF_Bonded term lists (bonded_terms)."""
    seq = random_sequence(seed, residues)
    m = build(seq, seed=seed, mode=mode)
    if bonded:
        m.update(bonded_terms(m["xyz"], m["el"], m["bonds"]))
    return m


def _peptide_job(args):
    seed, residues, mode, bonded = args
    return peptide(seed, residues, mode, bonded)


def many(seeds, residues, mode="extended", workers=None, bonded=False, start="fork"):
    """Builds many molecules in parallel worker processes. start="fork" is
    the fast choice before any GPU context exists in the process; a process
    that has used the GPU passes start="spawn" (fresh interpreters)."""
    import multiprocessing as mp
    import os
    jobs = list(zip(seeds, residues, [mode] * len(seeds), [bonded] * len(seeds)))
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(jobs) < 8:
        return [_peptide_job(j) for j in jobs]
    with mp.get_context(start).Pool(workers) as pool:
        return pool.map(_peptide_job, jobs, chunksize=max(1, len(jobs) // (4 * workers)))
