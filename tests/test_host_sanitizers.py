"""The library's host-only code under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/host_sanitize_check.cpp is compiled with
g++ together with pd.arcus_amd/csrc/pdgpu_setup.cpp and pdgpu_prep.cpp
(option defaults, bond-order band, fixed GB terms, parameter-file reader,
sequence grammar, peptide builder, term assembly) and run; any sanitizer
report fails the run. With PD's parameter file present it also builds
peptides of up to 150 residues and feeds the grammar malformed sequences."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pd.arcus_amd", "csrc")
PARAM = "/root/reference/param/amber99aa.ff"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_sanitize_check"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host_sanitize_check.cpp"), os.path.join(CSRC, "pdgpu_setup.cpp"),
           os.path.join(CSRC, "pdgpu_prep.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    args = [str(exe)] + ([PARAM] if os.path.exists(PARAM) else [])
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks passed" in r.stdout
    assert "runtime error" not in r.stderr
