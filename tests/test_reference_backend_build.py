"""The reference-side HIP backend (highperformanceseismicstacking_amd/reference_backend/hip)
compiles against the reference's own common/ headers and links with the reference's
search drivers, data classes, reader, traveltime sources and libhss.so into one program
with no undefined name (oracle/ref/Makefile, target ref_hip_backend).  Needs the reference
tree, so it runs in the build container only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree absent")
def test_reference_backend_compiles_and_links_against_the_reference():
    if not os.path.exists(os.path.join(ROOT, "highperformanceseismicstacking_amd", "libhss.so")):
        pytest.fail("libhss.so is not built; run __graft_entry__.build()")
    out = os.path.join(ROOT, "oracle", "_ref", "ref_hip_backend")
    r = subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle", "ref"), "../_ref/ref_hip_backend"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert os.access(out, os.X_OK)
    # every symbol of the hss C ABI the backend calls resolves in libhss.so
    nm = subprocess.run(["nm", "-u", out], capture_output=True, text=True).stdout
    assert "hss_grid_semblance" in nm and "hss_de_semblance" in nm and "hss_stretch_free" in nm
