"""Shared helpers for GPU tests: contexts configured like a golden case."""
import numpy as np

import golden_io as G
import pdgpu

OPT_MAP = {"cutoff": "cutoff", "inner_cutoff": "inner_cutoff", "vdw_cutoff": "vdw_cutoff",
           "vdw_inner_cutoff": "vdw_inner_cutoff", "vdw14_scaling": "vdw14_scaling",
           "elec14_scaling": "elec14_scaling", "dielectric_solute": "dielectric_solute",
           "dielectric_solvent": "dielectric_solvent", "still_cutoff": "gbsa_still_cutoff",
           "exp_approx_threshold": "exp_approx_threshold", "dielectric_offset": "dielectric_offset",
           "born_radius_offset": "born_radius_offset", "P1": "p1", "P2": "p2", "P3": "p3", "P4": "p4",
           "P5": "p5", "padding": "nlist_padding"}


def options_for(g, precision=pdgpu.FP64, **extra):
    kw = {OPT_MAP[k]: v for k, v in g["param_dict"].items()}
    kw.update(extra)
    return pdgpu.default_options(precision=precision, **kw)


def context_for(g, precision=pdgpu.FP64, **extra):
    return pdgpu.Context(options_for(g, precision, **extra))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def rms_rel(f, g):
    return float(np.sqrt(np.mean((f - g) ** 2)) / np.sqrt(np.mean(g ** 2)))


J2KCAL = 1.0 / pdgpu.J_PER_KCALMOL
