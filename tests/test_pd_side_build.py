"""The PD-side binding (pd.arcus_amd/pd_side/gbff_hip.h) builds inside PD:
compiled against the reference's own headers, linked with the compiled
reference and libpdgpu.so (oracle/Makefile, target pd_side). Runs where
the reference tree is present (the build container); skipped elsewhere.
Also checks that INTEGRATION.md shows the header verbatim."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/mmlib"
HEADER = os.path.join(ROOT, "pd.arcus_amd", "pd_side", "gbff_hip.h")


def test_integration_guide_shows_the_header_verbatim():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", doc, flags=re.S)
    assert open(HEADER).read() in blocks


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_pd_side_binding_compiles_and_links_against_pd():
    lib = os.path.join(ROOT, "pd.arcus_amd", "lib", "libpdgpu.so")
    if not os.path.exists(lib):
        pytest.skip("libpdgpu.so not built")
    r = subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "ref", "pd_side"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = os.path.join(ROOT, "oracle", "_ref", "gbff_hip_check")
    assert os.access(exe, os.X_OK)
    # the binary resolves every libpdgpu symbol the header uses
    r = subprocess.run(["ldd", exe], capture_output=True, text=True)
    assert "libpdgpu.so" in r.stdout and "not found" not in r.stdout
