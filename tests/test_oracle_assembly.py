"""Pins the C oracle of the system assembly (oracle/oracle_assembly.c):
classification (set_nflag), row numbering, init_matrix/cal_matrix_element
coefficients and the RHS, bit for bit against the flang build of the
reference's own routines (oracle/_ref/libref_assembly.so, driven by
oracle/ref_glue_assembly.F90) on small instances of the benchmark
configurations; plus closed-form checks of the uniform-grid coefficients."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REF_ASM_SO, oracle, ptr
from clowns3d import configs


def build_oracle(g, lda=None):
    A = g.arrays()
    nc = g.ncell
    lda = lda or nc
    out = dict(nf=np.zeros(nc, np.int8), nb=np.zeros(nc, np.int8), row2cell=np.zeros(nc, np.int32),
               mno=np.zeros(nc, np.int32), AL=np.zeros((6, lda)), LL=np.zeros((6, lda), np.int32),
               D=np.zeros(lda), B=np.zeros(lda))
    o = oracle()
    o.oracle_build_system.restype = ctypes.c_int
    N = o.oracle_build_system(g.nx, g.ny, g.nz, ptr(A["solid"]), ptr(A["F"]), ptr(A["fb"]), ptr(A["a"]),
                              ptr(A["un"]), ptr(A["dx"]), ptr(A["xf"]), g.stride, ctypes.c_double(g.dt),
                              ptr(out["nf"]), ptr(out["nb"]), ptr(out["row2cell"]), ptr(out["mno"]), lda,
                              ptr(out["AL"]), ptr(out["LL"]), ptr(out["D"]), ptr(out["B"]))
    assert N > 0
    out["N"] = N
    out["lda"] = lda
    return out


_refasm = None


def ref_asm():
    global _refasm
    if _refasm is None:
        if not os.path.exists(REF_ASM_SO):
            pytest.skip("flang build of the reference assembly (oracle/_ref) not present")
        _refasm = ctypes.CDLL(REF_ASM_SO, mode=os.RTLD_LAZY)
    return _refasm


def build_ref(g):
    A = g.arrays()
    nc = g.ncell
    lda = nc
    out = dict(nf=np.zeros(nc, np.int8), nb=np.zeros(nc, np.int8), row2cell=np.zeros(nc, np.int32),
               AL=np.zeros((6, lda)), LL=np.zeros((6, lda), np.int32), D=np.zeros(lda), B=np.zeros(lda))
    n = ctypes.c_int(0)
    mp = ctypes.c_double(0)
    ref_asm().ref_build_system(g.nx, g.ny, g.nz, ptr(A["solid"]), ptr(A["F"]), ptr(A["fb"]), ptr(A["a"]),
                               ptr(A["un"]), ptr(A["dx"]), ptr(A["xf"]), g.stride, ctypes.c_double(g.dt),
                               ptr(out["nf"]), ptr(out["nb"]), ptr(out["row2cell"]), lda, ptr(out["AL"]),
                               ptr(out["LL"]), ptr(out["D"]), ptr(out["B"]), ctypes.byref(n), ctypes.byref(mp))
    out["N"] = n.value
    out["mp"] = mp.value
    return out


GRIDS = {
    "wavy20x16x12": lambda: configs.config_wavy_box(20, 16, 12, (6, 11, 5, 10, 0, 4), seed=7),
    "wavy32x24x16": lambda: configs.config_wavy_box(32, 24, 16, (8, 20, 8, 16, 0, 6), seed=8),
    "bathy28x24x16": lambda: configs.config_bathymetry(28, 24, 16, nbumps=6, seed=9, width=(2, 5)),
    "solitary40x8x12": lambda: configs.config_solitary(40, 8, 12, t=0.7, seed=10),
    "objects40x24x16": lambda: configs.config_objects(40, 24, 16, seed=11),
}


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_assembly_matches_reference(name):
    g = GRIDS[name]()
    r = build_ref(g)
    o = build_oracle(g)
    N = r["N"]
    assert o["N"] == N
    assert np.array_equal(o["nf"], r["nf"])
    assert np.array_equal(o["nb"], r["nb"])
    assert np.array_equal(o["row2cell"][:N], r["row2cell"][:N])
    assert np.array_equal(o["LL"][:, :N], r["LL"][:, :N])
    assert np.array_equal(o["AL"][:, :N], r["AL"][:, :N])
    assert np.array_equal(o["D"][:N], r["D"][:N])
    assert np.array_equal(o["B"][:N], r["B"][:N])
    # every class of row occurs
    nf = o["nf"][o["row2cell"][:N]]
    assert (nf == 1).any() and (nf == 2).any()


def test_uniform_coefficients_closed_form():
    """dx=dy=1, dz=0.5, a=fb=1: interior fluid row ALU_x=-0.25, ALU_z=-1,
    D=3 (1sub-matrix.F90:50-61, mp = 0.5^2); a surface row has D=1, the real
    coefficient towards the fluid below -(fss-1/2)/(fss+1/2), the others -1e-20."""
    g = configs.config_wavy_box(20, 16, 12, (6, 11, 5, 10, 0, 4), seed=7)
    o = build_oracle(g)
    N, X, Y = o["N"], g.X, g.Y
    cells = o["row2cell"][:N]
    nf = o["nf"][cells]
    k = cells // (X * Y)
    interior = (nf == 1) & (k >= 2) & (k <= 3)
    rows = np.nonzero(interior)[0]
    ok = 0
    for r in rows:
        if (o["LL"][:, r] > 0).all():
            assert list(o["AL"][:, r]) == [-0.25, -0.25, -0.25, -0.25, -1.0, -1.0]
            assert o["D"][r] == 3.0
            ok += 1
    assert ok > 100
    surf = np.nonzero(nf == 2)[0]
    for r in surf[:50]:
        assert o["D"][r] == 1.0
        c = cells[r]
        b = o["nb"][c]
        if b == -3 and o["LL"][4, r] > 0:
            zs = g.arrays()["F"][c] * 0.5  # sp - x_low = F*dz for b = -3 with fb = 1
            fss = zs / 0.5
            expect = (fss - 0.5) * 0.5 / (fss * 0.5 + 0.5 * 0.5) * (-1.0)
            assert abs(o["AL"][4, r] - expect) <= 1e-15
            others = [o["AL"][s, r] for s in (0, 1, 2, 3) if o["LL"][s, r] > 0]
            assert all(v == -1e-20 for v in others)


def oracle_rhs_dmax(g, o, un):
    """oracle: RHS b (per row) and DMAX of a pass for the current velocities."""
    A = g.arrays()
    N = o["N"]
    orc = oracle()
    B = np.zeros(N)
    orc.oracle_rhs.restype = ctypes.c_double
    Fc = A["F"].copy()
    Fc[o["nf"] == 0] = np.maximum(Fc[o["nf"] == 0], 0.0)
    mp = min(g.h) ** 2
    dmax = orc.oracle_rhs(g.nx, g.ny, g.nz, ptr(A["solid"]), ptr(o["nf"]), ptr(Fc), ptr(A["fb"]), ptr(A["a"]),
                          ptr(un), ptr(A["dx"]), g.stride, ctypes.c_double(g.dt), ctypes.c_double(mp), N,
                          ptr(o["row2cell"]), ptr(B))
    return B, dmax


def oracle_apply_correction(g, o, xe, p, un):
    """oracle: p -= xe and estimate_velocity2, in place."""
    A = g.arrays()
    for r in range(o["N"]):
        p[o["row2cell"][r]] -= xe[r]
    oracle().oracle_velocity_correction(g.nx, g.ny, g.nz, ptr(A["solid"]), ptr(o["nf"]), ptr(A["a"]),
                                        ptr(A["dx"]), g.stride, ctypes.c_double(g.dt), ptr(o["mno"]), ptr(xe),
                                        ptr(un))


def oracle_correction_pass(g, o, xe, p, un):
    oracle_apply_correction(g, o, xe, p, un)
    return oracle_rhs_dmax(g, o, un)


@pytest.mark.parametrize("name", ["wavy20x16x12", "bathy28x24x16", "objects40x24x16"])
def test_correction_pass_matches_reference(name):
    """p -= xe, estimate_velocity2 (1sub-momentum.F90:590-651) and the RHS of
    the next pass, bit for bit against the reference routines; on the grid
    with object cells, pressure-boundary cells' own faces and the faces of
    flow cells next to them are corrected too."""
    g = GRIDS[name]()
    r = build_ref(g)
    o = build_oracle(g)
    N = o["N"]
    rng = np.random.default_rng(5)
    xe = rng.uniform(-1, 1, N)
    A = g.arrays()
    un_r = A["un"].ravel().copy()
    p_r = rng.uniform(0, 1, g.ncell)
    p_o = p_r.copy()
    un_o = A["un"].copy()
    Br = np.zeros(N)
    dm = ctypes.c_double(0)
    ref_asm().ref_correction_pass(g.nx, g.ny, g.nz, ptr(xe), ptr(un_r), ptr(p_r), ptr(Br), ctypes.byref(dm))
    Bo, dmo = oracle_correction_pass(g, o, xe, p_o, un_o)
    sol = A["solid"] == 0
    assert np.array_equal(p_o[sol], p_r[sol])
    un_r = un_r.reshape(3, -1)
    case = A["solid"] != 1  # every cell the case numbers, object cells included
    assert np.array_equal(un_o[:, case], un_r[:, case])
    moved = (un_o != A["un"]).any(axis=0)
    if (A["solid"] == 4).any():
        assert moved[A["solid"] == 4].any()  # pressure-boundary cells' faces were corrected
    assert np.array_equal(Bo, Br)
    assert dmo == dm.value


def test_object_cells_change_the_system():
    """The object-cell grid pins what it should: read as cells outside the
    case (the sentinel), its object cells would give another right-hand side
    (their own face fluxes drop out), so the bit-exact match above checks
    the object handling itself."""
    g = GRIDS["objects40x24x16"]()
    r = build_ref(g)
    g2 = GRIDS["objects40x24x16"]()
    g2.solid[g2.solid != 0] = 1
    o2 = build_oracle(g2)
    N = r["N"]
    assert o2["N"] == N
    assert not np.array_equal(o2["B"][:N], r["B"][:N])
    codes = set(np.unique(g.solid).tolist())
    assert {4, 5, -1, -128} <= codes
