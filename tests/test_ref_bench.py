"""The compiled reference's CPU timing driver (oracle/ref_bench.cpp) and the
oracle port agree on the batch workload's PD-built conformers: same pair
counts within the cutoff and energies within 1e-10, so the reference/port
time ratio in the CPU baseline compares the same work. Runs where the
reference tree and PD's parameter file exist (the build container)."""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/mmlib"
PARAM = "/root/reference/param/amber99aa.ff"

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(PARAM)),
                                reason="reference tree not present")


def test_reference_and_port_agree_on_batch_conformers(monkeypatch):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "ref_bench"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_cpu_baseline as rcb
    with tempfile.TemporaryDirectory() as d:
        lengths = os.path.join(d, "lengths.txt")
        with open(lengths, "w") as f:
            f.write("12\n3\n150\n40\n")  # 3 residues: the smallest the grammar helper takes
        ref = rcb.run_ref(0, 4, 1e9, lengths, d)
        assert [row[0] for row in ref] == [0, 1, 2, 3]
        monkeypatch.setattr(rcb, "DUMP", d)
        port = rcb._port_job(([0, 1, 2, 3], 1e9))
    for a, b in zip(ref, port):
        assert a[1] == b[1]  # atoms
        assert a[2] == b[2] > 0  # unordered pairs within 10 A
        assert abs(a[4] - b[4]) <= 1e-10 * abs(a[4])
