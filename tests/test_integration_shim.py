"""The reference-side shim (tapsim_modified_amd/integration/grid_3d_gpu.h)
compiles in the reference's C++98 dialect against the reference's own
headers and links with the reference's objects (built from its sources by
oracle/Makefile.ref) and libtapsim_gpu.so.  Compile and link only: running
it needs the GPU."""
import glob
import os
import subprocess

import pytest

from tests.conftest import ROOT

REF = "/root/reference"
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "tapsim")) or not os.path.isdir(OBJ),
                    reason="needs the reference sources and oracle/_ref (oracle/Makefile.ref)")
def test_shim_compiles_and_links_with_the_reference(tmp_path):
    inc = ["-I" + os.path.join(REF, d) for d in ("tapsim", "vector", "predicates", "tetgen")]
    inc += ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "tapsim_modified_amd", "integration")]
    obj = str(tmp_path / "shim_link.o")
    r = subprocess.run(["g++", "-std=gnu++98", "-O1", "-Wall", "-Wno-unused-function", "-c",
                        os.path.join(ROOT, "tests", "integration", "shim_link.cpp"), "-o", obj] + inc,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # the reference's objects except the fixture driver's own
    # (the two predicate builds go through their archive, as the reference links them)
    refobjs = sorted(p for p in glob.glob(os.path.join(OBJ, "*.o"))
                     if not p.endswith("ref_driver.o") and "predicates_" not in os.path.basename(p))
    exe = str(tmp_path / "shim_link")
    r = subprocess.run(["g++", "-o", exe, obj] + refobjs +
                       [os.path.join(ROOT, "oracle", "_ref", "libpredicates.a"),
                        "-L" + os.path.join(ROOT, "tapsim_modified_amd"), "-ltapsim_gpu", "-lpthread",
                        "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,-rpath," + os.path.join(ROOT, "tapsim_modified_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    nm = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    for sym in ("tsg_grid_create", "tsg_relax_threshold", "tsg_local_relax", "tsg_field_o2", "tsg_evaporate"):
        assert sym in nm, sym
