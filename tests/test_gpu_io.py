"""Output formats written from device state through the C ABI:
pdgpu_dcd_create / pdgpu_dcd_append and pdgpu_dcd_frame (positions
converted to float on the device) against the DCD file the compiled
reference's OutTra_NAMD wrote for the same three frames (every byte except
PD's uninitialised title bytes), and pdgpu_snapshot_mime against
WorkSpace::save().writeMIME (positions and velocities bit-exact, forces and
energy to the FP64 path's bounds: 1e-9 force RMS, 1e-10 energy), and
pdgpu_btf_header + pdgpu_btf_append against the BTF file of OutTra_BTF
(every include, three frames)."""
import numpy as np
import pytest

import golden_io as G
import pdgpu
from gpu_helpers import context_for, rel, rms_rel
from test_io_formats import IO_CASES, TITLE_GARBAGE, decode_6bit, split_mime

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", IO_CASES)
def test_dcd_file_matches_reference(name, tmp_path):
    g = G.load(name)
    n = len(g["mass"])
    frames = g["io_dcd_xyz"].reshape(3, n, 3)
    path = tmp_path / "t.dcd"
    with context_for(g, pdgpu.MIXED) as ctx:
        i = ctx.add_system(g, bonded=True)
        ctx.dcd_create(path, i)
        for fr in frames:
            ctx.set_positions(i, fr)
            ctx.dcd_append(path, i)
        one = ctx.dcd_frame(i)
    ours = path.read_bytes()
    ref = g["io_dcd"].view(np.uint8).tobytes()
    assert len(ours) == len(ref)
    keep = np.ones(len(ref), bool)
    keep[TITLE_GARBAGE] = False
    assert np.array_equal(np.frombuffer(ours, np.uint8)[keep], np.frombuffer(ref, np.uint8)[keep])
    assert one == ours[-len(one):]


@pytest.mark.parametrize("name", IO_CASES)
def test_snapshot_mime_of_device_state(name):
    g = G.load(name)
    n = len(g["mass"])
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz"], bonded=True)
        ctx.set_velocities(i, g["md_v"])
        ctx.calc_forces(i)
        text = ctx.snapshot_mime(i, name)
    nm, body = split_mime(text)
    assert nm == name
    rec = decode_6bit(body)
    assert np.frombuffer(rec[0:4], np.int32)[0] == n
    assert rel(np.frombuffer(rec[4:12], np.float64)[0], float(g["io_epot"])) <= 1e-10
    a = np.frombuffer(rec[128:128 + 72 * n], np.float64).reshape(n, 9)
    assert np.array_equal(a[:, :3], g["md_xyz"]) and np.array_equal(a[:, 3:6], g["md_v"])
    assert rms_rel(a[:, 6:], g["io_force"]) <= 1e-9


@pytest.mark.parametrize("name", [c for c in IO_CASES if "md_dcd" in G.load(c)])
def test_md_trajectory_matches_reference(name, tmp_path):
    """PD's VelocityVerlet run with UpdateTra = 10 and an OutTra_NAMD in the
    workspace: the device run writes the same frames (steps 10, 20, ...),
    to FP64 rounding of the positions (the float frames then agree to a few
    ulp); header bytes as in the DCD file test."""
    g = G.load(name)
    n = len(g["mass"])
    steps = int(g["md_steps"])
    path = tmp_path / "md.dcd"
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz0"], bonded=True)
        ctx.set_velocities(i, g["md_v0"].reshape(-1, 3))
        ctx.traj_dcd(i, path, 10)
        ctx.md_vv_pd(i, steps, float(g["md_dt"]), int(g["md_update_nlist"]))
        ctx.traj_dcd(i, None, 0)
    ours = path.read_bytes()
    ref = g["md_dcd"].view(np.uint8).tobytes()
    frames = (steps - 1) // 10
    assert len(ours) == len(ref) == 276 + frames * 3 * (8 + 4 * n)
    keep = np.ones(276, bool)
    keep[TITLE_GARBAGE] = False
    assert np.array_equal(np.frombuffer(ours[:276], np.uint8)[keep], np.frombuffer(ref[:276], np.uint8)[keep])
    a = np.frombuffer(ours[276:], np.uint8).reshape(frames, 3, 8 + 4 * n)
    b = np.frombuffer(ref[276:], np.uint8).reshape(frames, 3, 8 + 4 * n)
    assert np.array_equal(a[:, :, :4], b[:, :, :4]) and np.array_equal(a[:, :, -4:], b[:, :, -4:])
    fa = a[:, :, 4:-4].copy().view(np.float32)
    fb = b[:, :, 4:-4].copy().view(np.float32)
    assert np.abs(fa - fb).max() <= 1e-6 * max(1.0, np.abs(fb).max())


@pytest.mark.parametrize("name", [c for c in IO_CASES if "io_btf" in G.load(c)])
def test_btf_file_matches_reference(name, tmp_path):
    """The whole BTF file: header as in the CPU test, then three frames of
    device state. Positions are the set positions' float casts, so those
    bytes are equal. phi/psi, the energy record (kcal/mol) and the force
    vectors p + 5e8 f go through the FP64 path (energies within 1e-10,
    forces within 1e-9 of the reference), so each of those floats may sit
    one float spacing from the reference's; the step, state and zero
    fields are equal."""
    g = G.load(name)
    n = g["n"]
    ref = np.frombuffer(g["io_btf"].view(np.uint8).tobytes(), np.uint8)
    bb = g["io_backbone"].reshape(-1, 3)
    nres = len(bb)
    date = np.frombuffer(ref[24:36].tobytes(), np.int32)
    st = g["io_state"]
    state = pdgpu.btf_state(step=int(st[0]), crms=st[1], drms=st[2], ekin=st[3], etot=st[4], data=st[5:15])
    path = tmp_path / "t.tra"
    path.write_bytes(pdgpu.btf_header(G.btf_atoms(g), nres, pdgpu.BTF_ALL, date))
    frames = g["io_dcd_xyz"].reshape(3, n, 3)
    with context_for(g, pdgpu.FP64) as ctx:
        i = ctx.add_system(g, xyz=g["md_xyz"], bonded=True)
        ctx.calc_forces(i)
        for fr in frames:
            ctx.set_positions(i, fr)
            ctx.btf_append(path, i, pdgpu.BTF_ALL, bb, state)
        one = ctx.btf_frame(i, pdgpu.BTF_ALL, bb, state)
    ours = np.frombuffer(path.read_bytes(), np.uint8)
    assert len(ours) == len(ref)
    hlen = 18084 + 8 + 220 * n + 8
    undefined = G.btf_undefined_bytes(g, hlen)
    assert np.array_equal(ours[:hlen][~undefined], ref[:hlen][~undefined])
    bs = 4 + 12 * n + 8 * nres + 256 + 12 * n
    assert np.frombuffer(ref[16:20].tobytes(), np.int32)[0] == bs
    for k in range(3):
        a, b = ours[hlen + k * bs:hlen + (k + 1) * bs], ref[hlen + k * bs:hlen + (k + 1) * bs]
        assert bytes(a[:4]) == bytes(b[:4]) == b"TRAE"
        fa, fb = np.frombuffer(a[4:].tobytes(), np.float32), np.frombuffer(b[4:].tobytes(), np.float32)
        pos = slice(0, 3 * n)
        assert np.array_equal(fa[pos], fb[pos])
        ulp = np.spacing(np.abs(fb).astype(np.float32))
        assert np.all(np.abs(fa.astype(np.float64) - fb) <= ulp), k
        e0 = 3 * n + 2 * nres
        assert np.array_equal(fa[e0:e0 + 5], fb[e0:e0 + 5])  # step, time, cRMS, dRMS, etot
        assert np.array_equal(fa[e0 + 18:e0 + 64], fb[e0 + 18:e0 + 64])  # zeros and data1..10
    assert one == ours[-bs:].tobytes()
