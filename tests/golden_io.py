"""Loads the golden fixtures written from the compiled reference."""
import glob
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

PARAM_NAMES = ("cutoff", "inner_cutoff", "vdw_cutoff", "vdw_inner_cutoff", "vdw14_scaling",
               "elec14_scaling", "dielectric_solute", "dielectric_solvent", "still_cutoff",
               "exp_approx_threshold", "dielectric_offset", "born_radius_offset",
               "P1", "P2", "P3", "P4", "P5", "padding")


def case_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))


def load(name):
    with np.load(os.path.join(GOLDEN_DIR, name + ".npz")) as z:
        d = {k: z[k] for k in z.files}
    d["name"] = name
    d["n"] = len(d["charge"])
    d["xyz"] = d["xyz"].reshape(-1, 3)
    for k in ("force", "en_force", "nb_force", "sd_force", "sd_xyz", "bf_force", "bfp_force", "bfp_xyz",
              "ff_force", "sdf_force", "sdf_xyz", "sv_force", "svsd_force", "svsd_xyz", "md_xyz0", "md_v0",
              "md_xyz", "md_v", "lg_xyz0", "lg_xyz", "lg_v", "bm_xyz0", "bm_xyz", "bm_v", "io_force", "lx_xyz0", "lx_xyz", "lx_v", "lh_xyz0", "lh_xyz", "lh_v", "vb_xyz0", "vb_xyz", "vb_v", "bb_xyz0", "bb_xyz", "bb_v"):
        if k in d:
            d[k] = d[k].reshape(-1, 3)
    d["mid"] = int(d["mid"][0])
    d["param_dict"] = dict(zip(PARAM_NAMES, (float(x) for x in d["params"])))
    for k in list(d):
        if isinstance(d[k], np.ndarray) and d[k].shape == (1,) and d[k].dtype == np.float64:
            d[k] = float(d[k][0])
    return d


def btf_atoms(g):
    """The per-atom system-definition inputs of a BTF file (pdgpu.btf_header's
    atom dicts) from a fixture's io_* records."""
    n = g["n"]
    names = g["io_names"].view(np.uint8).reshape(n, 4, 16)
    text = lambda row: bytes(row).split(b"\0")[0].decode()
    cov = g["io_cov12"].reshape(n, 6)
    ref = g["io_ref"].reshape(n, 3)
    return [dict(pdbname=text(names[i, 0]), type_name=text(names[i, 1]), rawname=text(names[i, 2]),
                 parentname=text(names[i, 3]), ir=int(g["resid"][i]), known_structure=int(g["io_known"][i]),
                 ref=ref[i], cov12=list(cov[i, :g["io_ncov12"][i]]), charge=g["charge"][i], radius=g["radius"][i],
                 epsilon=g["epsilon"][i], mass=g["mass"][i]) for i in range(n)]


def btf_undefined_bytes(g, length):
    """Positions in a BTF file that PD writes from uninitialised memory: the
    ID strings after "null" (BTF_Header, fileio/tratypes.cpp:24-27),
    parentname after its terminator (strcpy into a reused entry) and the
    cov12atom slots past n_cov12atoms (tratypes.cpp:144-171)."""
    mask = np.zeros(length, bool)
    for q in range(4):
        mask[36 + 32 * q + 5:36 + 32 * (q + 1)] = True
    n = g["n"]
    names = g["io_names"].view(np.uint8).reshape(n, 4, 16)
    for i in range(n):
        base = 18084 + 8 + 220 * i
        plen = len(bytes(names[i, 3]).split(b"\0")[0])
        mask[base + 32 + plen + 1:base + 40] = True
        mask[base + 56 + 4 * int(g["io_ncov12"][i]):base + 80] = True
    return mask
