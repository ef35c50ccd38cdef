"""GPU: the Fortran drop-in.  pcgpme_shim_test (Fortran) calls PCGPME with
cal_poisson_equation's arguments (src2LOWNS/1sub-momentum.F90:559-564) through
fortran/pcgpme_hip.F90 and the ISO_C_BINDING interface to clowns_pcgpme; the
solve must converge and leave the reference's scaled/sorted row in AL/LL/D:
an interior row of the uniform grid (dx=dy=1, dz=0.5) scales to D=1,
AL = [-1/3, -1/3, -1/12 x4] in slot order [z-, z+, y-, y+, x-, x+]."""
import os
import re
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def test_fortran_shim_drop_in():
    exe = os.path.join(PKG, "lib", "pcgpme_shim_test")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    m = re.search(r"N=\s*(\d+)\s+ITER=\s*(\d+)\s+IER=\s*(-?\d+)\s+EPS=\s*(\S+)\s+RES=\s*(\S+)", out.stdout)
    N, it, ier, eps, res = int(m[1]), int(m[2]), int(m[3]), float(m[4]), float(m[5])
    assert N == 24 * 20 * 12
    assert ier == 0 and 0 < it < 1000 and eps <= 1e-10
    assert res < 1e-8
    row = [float(v) for v in re.search(r"ROW\s+(.*)", out.stdout)[1].split()]
    assert row[0] == 1.0
    assert abs(row[1] + 1 / 3) < 1e-15 and abs(row[2] + 1 / 3) < 1e-15
    assert all(abs(v + 1 / 12) < 1e-15 for v in row[3:])
    ll = [int(v) for v in re.search(r"LL\s+(.*)", out.stdout)[1].split()]
    r = 12 + 24 * ((10 - 1) + 20 * (6 - 1))  # 1-based row of the printed cell
    assert ll == [r - 480, r + 480, r - 24, r + 24, r - 1, r + 1]
