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


# Accuracy contract of the mixed path against FP64 (oracle or the FP64 GPU
# path), per instance: |dE| <= max(2e-5 |E|, 2e-5 kcal/mol x sqrt(P)), P the
# instance's pair count; forces: RMS relative error <= 1e-4. The FP32 pair
# terms carry a random relative error of ~2e-7 each, so the absolute error
# of a sum of P terms grows like sqrt(P): an instance whose total nearly
# cancels (|E| ~ 1 kcal/mol from components of ~1e4) is held to the absolute
# bound. Block-relative FP32 coordinates resolve ~1e-6 A, and the r^-12
# repulsion multiplies a relative error in r by 12: an energy dominated by
# clashes (~1 A contacts) reaches ~1e-5 relative. Measured over the 4,096
# conformers of the batch workload: 99.4% within 1e-5 relative (median
# 5.6e-7), the worst near-cancelling total at 4.3e-6 kcal/mol x sqrt(P),
# the worst clash-dominated one (+1.7e5 kcal/mol) at 1.1e-5 relative.
MIXED_REL = 2e-5
MIXED_ABS_PER_SQRT_PAIR = 2e-5  # kcal/mol
MIXED_FORCE_RMS = 1e-4


def mixed_energy_bound(e_ref, pairs):
    """Largest |E_mixed - E_ref| (J/molecule) the contract allows."""
    return max(MIXED_REL * abs(e_ref), MIXED_ABS_PER_SQRT_PAIR * np.sqrt(max(pairs, 1)) * pdgpu.J_PER_KCALMOL)
