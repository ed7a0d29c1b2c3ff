"""PD's result hand-off formats written by the library's host code
(pdgpu_mime_encode, pdgpu_dcd_header, pdgpu_btf_header), against the bytes
the compiled reference wrote (golden io_mime: WorkSpace::save().writeMIME
after a VelocityVerlet run; io_dcd: OutTra_NAMD with three frames; io_btf:
OutTra_BTF with every include and the same three frames).

PD leaves some bytes uninitialised: 100 bytes of the MIME header
(workspace/snapshot.cpp:321-330 allocates it with new[]) and each DCD title
record past its 25 characters (fileio/psfdcd.cpp: an 80-byte fwrite of a
stack buffer), and in BTF the ID strings after "null", parentname past its
terminator and unused cov12atom slots. Those bytes are zero here and are
masked in the comparison; every other byte must match."""
import numpy as np
import pytest

import golden_io as G
import pdgpu

IO_CASES = [c for c in G.case_names() if "io_mime" in G.load(c)]
CODE = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"


def decode_6bit(text):
    """PD's decoder (tools/io.cpp:165-185): characters below 32 are
    skipped, four characters make three bytes, low bits first."""
    vals = [CODE.index(ch) if ch in CODE else 63 for ch in text if ord(ch) >= 32]
    out = bytearray()
    for k in range(0, len(vals), 4):
        a, b, c, d = (vals[k:k + 4] + [0, 0, 0])[:4]
        out += bytes([(a | (b << 6)) & 255, ((b >> 2) | (c << 4)) & 255, ((d << 2) | (c >> 4)) & 255])
    return bytes(out)


def split_mime(text):
    lines = text.split("\n")
    assert lines[0].startswith("!PSP ")
    end = lines.index("!ENDPSP")
    return lines[0][5:], "\n".join(lines[1:end])


def ref_text(g):
    return g["io_mime"].view(np.uint8).tobytes().decode()


@pytest.mark.parametrize("name", IO_CASES)
def test_mime_matches_reference(name):
    g = G.load(name)
    n = len(g["mass"])
    ours = pdgpu.mime_encode(g["md_xyz"], g["md_v"], g["io_force"], float(g["io_epot"]), name)
    ref = ref_text(g)
    # the reference's record with its uninitialised header bytes zeroed,
    # coded again by the library, is the library's text character for character
    rname, rbody = split_mime(ref)
    rec = bytearray(decode_6bit(rbody))
    length = 128 + 72 * n
    assert len(rec) == 3 * ((length + 2) // 3)
    rec[28:128] = bytes(100)
    again = pdgpu.mime_encode(np.frombuffer(bytes(rec[128:length]), np.float64).reshape(n, 9)[:, 0:3],
                              np.frombuffer(bytes(rec[128:length]), np.float64).reshape(n, 9)[:, 3:6],
                              np.frombuffer(bytes(rec[128:length]), np.float64).reshape(n, 9)[:, 6:9],
                              np.frombuffer(bytes(rec[4:12]), np.float64)[0], rname)
    assert again == ours
    # and the reference's text differs from ours only inside the groups
    # that code header bytes 27..128 (characters 36..171 of the body)
    body = rbody.replace("\n", "")
    ob = split_mime(ours)[1].replace("\n", "")
    assert len(body) == len(ob)
    assert body[:36] == ob[:36] and body[172:] == ob[172:]
    # line structure: 64 characters per line, the last one partial, then
    # the coder's closing newline
    lines = split_mime(ours)[1].split("\n")
    assert all(len(line) == 64 for line in lines[:-2]) and len(lines[-2]) <= 64 and lines[-1] == ""
    assert ours.endswith("\n!ENDPSP\n") and ours.startswith(f"!PSP {name}\n")


def test_mime_record_layout_round_trip():
    rng = np.random.default_rng(5)
    p, v, f = (rng.normal(size=(7, 3)) for _ in range(3))
    t = pdgpu.mime_encode(p, v, f, -1.25e-19, "x")
    rec = decode_6bit(split_mime(t)[1])
    assert np.frombuffer(rec[0:4], np.int32)[0] == 7
    assert np.frombuffer(rec[4:12], np.float64)[0] == -1.25e-19
    assert rec[12:128] == bytes(116)
    a = np.frombuffer(rec[128:128 + 7 * 72], np.float64).reshape(7, 9)
    assert np.array_equal(a[:, :3], p) and np.array_equal(a[:, 3:6], v) and np.array_equal(a[:, 6:], f)


TITLE_GARBAGE = list(range(125, 180)) + list(range(205, 260))


@pytest.mark.parametrize("name", IO_CASES)
def test_dcd_header_matches_reference(name):
    g = G.load(name)
    n = len(g["mass"])
    ref = g["io_dcd"].view(np.uint8).tobytes()
    ours = pdgpu.dcd_header(n)
    assert len(ours) == 276
    keep = np.ones(276, bool)
    keep[TITLE_GARBAGE] = False
    r = np.frombuffer(ref[:276], np.uint8)
    o = np.frombuffer(ours, np.uint8)
    assert np.array_equal(r[keep], o[keep])
    assert np.frombuffer(ours[268:272], np.int32)[0] == n
    # the frames that follow are 3 records of n floats each
    assert len(ref) == 276 + 3 * 3 * (8 + 4 * n)


@pytest.mark.parametrize("name", [c for c in IO_CASES if "io_btf" in G.load(c)])
def test_btf_header_matches_reference(name):
    """pdgpu_btf_header against the start of the BTF file the compiled
    reference's OutTra_BTF wrote (create(AllIncludes)): header, system
    definition and TRASTART equal in every byte PD defines (the date is the
    reference run's, read from its file)."""
    g = G.load(name)
    ref = np.frombuffer(g["io_btf"].view(np.uint8).tobytes(), np.uint8)
    nres = len(g["io_backbone"]) // 3
    date = np.frombuffer(ref[24:36].tobytes(), np.int32)
    h = np.frombuffer(pdgpu.btf_header(G.btf_atoms(g), nres, pdgpu.BTF_ALL, date), np.uint8)
    assert len(h) == 18084 + 8 + 220 * g["n"] + 8
    undefined = G.btf_undefined_bytes(g, len(h))
    assert undefined.sum() < 0.1 * len(h)
    assert np.array_equal(h[~undefined], ref[:len(h)][~undefined])
    # every undefined byte is zero in ours
    assert not h[undefined].any()


def test_btf_header_rejects_bad_input():
    a = [dict(pdbname="N", type_name="N3", rawname="N", parentname="ALA", ir=0, known_structure=0, ref=(0, 0, 0),
              cov12=list(range(7)), charge=0.1, radius=1.8, epsilon=0.1, mass=14.0)]
    with pytest.raises(pdgpu.PdgpuError):
        pdgpu.btf_header(a, 1)
    a[0]["cov12"] = [1]
    with pytest.raises(pdgpu.PdgpuError):
        pdgpu.btf_header(a, 1, includes=32)
    assert len(pdgpu.btf_header(a, 1, includes=pdgpu.BTF_ENERGY)) == 18084 + 8 + 220 + 8
