"""The C++ ForcefieldBase-shaped adapter (pd.arcus_amd/host/ff_gb_hip.*)
driven like a PD protocol drives a component, through the C ABI, on a
reference fixture: its bond orders and Still fixed terms come from the
covalent lists (host setup code), its energies and forces must match the
compiled reference."""
import os
import struct
import subprocess

import numpy as np
import pytest

import golden_io as G
from gpu_helpers import rel, rms_rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "pd.arcus_amd", "lib", "ff_gb_hip_demo")


def write_input(g, path, precision, bonded=False):
    p = g["param_dict"]
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", g["n"], len(g["a12"]), len(g["a13"]), len(g["a14"]), precision))
        f.write(struct.pack("<9d", p["cutoff"], p["inner_cutoff"], p["vdw_cutoff"], p["vdw_inner_cutoff"],
                            p["vdw14_scaling"], p["elec14_scaling"], p["dielectric_solvent"], p["padding"],
                            p["exp_approx_threshold"]))
        for i in range(g["n"]):
            f.write(struct.pack("<8di", *g["xyz"][i], g["charge"][i], g["radius"][i], g["epsilon"][i],
                                g["gb_radius"][i], g["mass"][i], int(g["resid"][i])))
        for a, b, d in zip(g["a12"], g["b12"], g["d12"]):
            f.write(struct.pack("<2id", int(a), int(b), d))
        for a, b, d in zip(g["a13"], g["b13"], g["d13"]):
            f.write(struct.pack("<2id", int(a), int(b), d))
        for a, b in zip(g["a14"], g["b14"]):
            f.write(struct.pack("<2i", int(a), int(b)))
        if bonded:
            bi, bl = g["bd_idx"].reshape(-1, 2), g["bd_lk"].reshape(-1, 2)
            ai, at = g["an_idx"].reshape(-1, 3), g["an_tk"].reshape(-1, 2)
            f.write(struct.pack("<4i", len(bl), len(at), len(g["to_terms"]), len(g["im_terms"])))
            for (i, j), (l, k) in zip(bi, bl):
                f.write(struct.pack("<2i2d", int(i), int(j), l, k))
            for (i, a, j), (t0, k) in zip(ai, at):
                f.write(struct.pack("<3i2d", int(i), int(a), int(j), t0, k))
            for tag in ("to", "im"):
                idx = g[tag + "_idx"].reshape(-1, 4)
                vn, nn, gm = (g[tag + s_].reshape(-1, 4) for s_ in ("_vn", "_n", "_gamma"))
                for q in range(len(idx)):
                    f.write(struct.pack("<5i", *[int(x) for x in idx[q]], int(g[tag + "_terms"][q])))
                    f.write(struct.pack("<12d", *vn[q], *nn[q], *gm[q]))


@pytest.mark.parametrize("name", ["c1_seed2", "tri_ala"])
def test_cpp_adapter_matches_reference(name, tmp_path):
    assert os.path.exists(DEMO), "C++ adapter not built"
    g = G.load(name)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(g, inp, 0)
    r = subprocess.run([DEMO, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, dtype=np.float64)
    n3 = 3 * g["n"]
    e1, f1 = d[:6], d[6:6 + n3].reshape(-1, 3)
    e2, f2 = d[6 + n3:13 + n3], d[13 + n3:].reshape(-1, 3)
    for v, k in zip(e1, ("epot", "epot_vdw", "epot_elec", "epot_pol", "epot_pol_self", "epot_pol_cross")):
        assert rel(v, g[k]) <= 1e-10, k
    assert rms_rel(f1, g["force"]) <= 1e-9
    assert rel(e2[0], g["en_epot"]) <= 1e-10
    assert rel(e2[3], g["en_pol"]) <= 1e-10
    assert rel(e2[6], g["en_virial"]) <= 1e-10
    assert rms_rel(f2, g["en_force"]) <= 1e-9


@pytest.mark.parametrize("name", ["tri_ala", "c1_seed3"])
def test_cpp_bonded_and_gb_forcefield(name, tmp_path):
    """FF_Bonded_HIP + FF_GeneralizedBorn_Still_HIP in one forcefield, as a
    PD Forcefield holding both components evaluates them."""
    g = G.load(name)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(g, inp, 0, bonded=True)
    r = subprocess.run([DEMO, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, dtype=np.float64)
    n3 = 3 * g["n"]
    e3 = d[13 + 2 * n3:21 + 2 * n3]
    f3 = d[21 + 2 * n3:].reshape(-1, 3)
    for v, k in zip(e3[:7], ("epot", "bond", "angle", "torsion", "vdw", "elec", "pol")):
        assert rel(v, g["ff_" + k]) <= 1e-10, k
    assert rms_rel(f3, g["ff_force"]) <= 1e-9
