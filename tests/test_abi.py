"""CPU checks of the boundary: the C-ABI library loads and exports every
function include/clowns_pcg.h declares (no compute call: no GPU here), the
Fortran shim object defines PCGPME with the reference's link name, and the
Python mirror validates layouts before calling."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "clowns_pcg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(clowns_\w+)\s*\(", src)))


def test_header_declares_the_drop_in():
    names = declared_functions()
    assert "clowns_pcgpme" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    import clowns3d
    lib = clowns3d.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", clowns3d.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert b"gfx950" in lib.clowns_version()


def test_library_targets_gfx950_only():
    import clowns3d
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", clowns3d.LIB_PATH], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    archs = set(re.findall(r"gfx\w+", out.stdout))
    assert archs == {"gfx950"}, archs


def test_fortran_shim_defines_pcgpme():
    obj = os.path.join(PKG, "lib", "pcgpme_hip.o")
    assert os.path.exists(obj)
    out = subprocess.run(["nm", obj], capture_output=True, text=True).stdout
    assert re.search(r"\bT pcgpme_\b", out), out  # PCGPME, default Fortran mangling
    assert re.search(r"\bU clowns_pcgpme\b", out), out
    assert re.search(r"\bT pcgpme_will_\b", out), out  # PCGPME_WILL
    assert re.search(r"\bU clowns_pcgpme_will\b", out), out


def test_python_mirror_rejects_bad_layouts():
    import clowns3d
    N, NA, NL = 10, 20, 6
    AL = np.zeros((NA, NL))  # C order: wrong
    LL = np.zeros((NA, NL), np.int32, order="F")
    with pytest.raises(ValueError):
        clowns3d.pcgpme(AL, NL, NA, LL, np.zeros(N), N, 64, np.zeros(N), np.zeros(N + 64), np.zeros(2), 5.0, 10)
    with pytest.raises(TypeError):
        clowns3d.pcgpme(np.asfortranarray(AL), NL, NA, LL.astype(np.int64, order="F"), np.zeros(N), N, 64,
                        np.zeros(N), np.zeros(N + 64), np.zeros(2), 5.0, 10)
