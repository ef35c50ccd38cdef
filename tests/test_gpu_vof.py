"""GPU: the device VOF update (clowns_solver_vof) against the C oracle
(oracle_vof, itself bit-exact against the flang build of the reference's
cal_vof routines): face fluxes, new F and the reduced dt bit for bit, with
the same counts of corrected and slid cells.

Cases: the synthetic configurations in the surface-stage modes (kaidan, PLIC
face fluxes given, quvwf_normal from PLIC segment volumes), a stressed field (corrections and slides), an
outflow that makes cal_new_f shrink dt, a chain of surface cells whose
corrections feed each other (the serial order of quvwf_modify), and the
256x256x64 configuration (volume conservation to rounding and the oracle)."""
import numpy as np
import pytest

from _vof import VOF_OK, call_oracle_vof, plic_faces, vof_case

pytestmark = pytest.mark.gpu

CASES = [("wavy20x16x12", "step"), ("wavy32x24x16", "stress"), ("bathy28x24x16", "step"),
         ("bathy28x24x16", "stress"), ("solitary40x8x12", "stress"), ("wavy20x16x12", "retry")]


def run_gpu(g, c, Fp, dt, mode, qs=None, mask=None):
    from clowns3d import DeviceGrid, Solver
    sol = Solver(0)
    dg = DeviceGrid(g, classified=c)
    Fn, qf, info = dg.vof(sol, Fp, mode=mode, qs=qs, qs_mask=mask, dt=dt)
    out = Fn.cpu().numpy(), qf.cpu().numpy(), info
    sol.close()
    return out


def compare(g, c, Fp, dt, mode, qs=None, mask=None):
    o = call_oracle_vof(g, c["arrays"], c["nf"], c["nb"], c["F"], Fp, dt, mode, qs, mask)
    assert o["rc"] == VOF_OK
    Fn, qf, info = run_gpu(g, c, Fp, dt, mode, qs, mask)
    print("oracle", {k: o[k] for k in ("retries", "modified", "slid", "dt")}, "gpu", info, flush=True)
    assert info["dt"] == o["dt"] and info["retries"] == o["retries"]
    assert info["modified"] == o["modified"] and info["slid"] == o["slid"]
    assert np.array_equal(qf, o["qf"])
    assert np.array_equal(Fn, o["F"])
    return o, info


@pytest.mark.parametrize("name,scenario", CASES)
@pytest.mark.parametrize("mode", [0, 1])
def test_vof_matches_oracle(name, scenario, mode):
    from clowns3d import assembly
    g, A, Fp, dt = vof_case(name, scenario, seed=CASES.index((name, scenario)))
    c = assembly.classify(g)
    qs, mask = plic_faces(g, A, c["nf"], seed=5) if mode == 0 else (None, None)
    o, info = compare(g, c, Fp, dt, mode, qs, mask)
    if scenario == "stress":
        assert o["modified"] > 0 and o["slid"] > 0
    if scenario == "retry":
        assert info["retries"] > 0


def test_vof_modify_chain_serial_order():
    """A vertical stack of overfull surface cells, each pointing up (nfb =
    -3): every correction raises the next cell's inflow, so each cell's
    decision waits for the one below -- one device round per cell."""
    from clowns3d import assembly
    g, A, Fp, dt = vof_case("wavy20x16x12", "step", seed=7)
    c = assembly.classify(g)
    X, Y = g.X, g.Y
    nf, nb = c["nf"].copy(), c["nb"].copy()
    stack = [5 + X * (6 + Y * k) for k in range(2, 9)]
    for q in stack:
        assert c["arrays"]["solid"][q] == 0
        nf[q], nb[q] = 2, -3
    c = dict(c, nf=nf, nb=nb)
    Fp = Fp.copy()
    Fp[stack] = 1.0 + 1e-3
    o, info = compare(g, c, Fp, dt, 1)
    assert o["modified"] >= len(stack)
    assert info["rounds"] >= len(stack)


def _cell_volumes(g):
    dx = g.dx
    return (dx[0, :g.X][None, None, :] * dx[1, :g.Y][None, :, None] * dx[2, :g.Z][:, None, None]).ravel()


def test_vof_config2_size():
    """256x256x64 (4.2M cells): bit-exact with the oracle; volume is
    conserved to rounding (closed box, no clamps in this field)."""
    from clowns3d import assembly, configs
    g = configs.config2()
    rng = np.random.default_rng(11)
    g.un[:] = rng.uniform(-0.1, 0.1, size=g.un.shape)
    c = assembly.classify(g)
    A = c["arrays"]
    Fp = c["F"].copy()
    o, info = compare(g, c, Fp, g.dt, 1)
    vol = _cell_volumes(g) * A["fb"]
    before, after = (vol * Fp).sum(), (vol * o["F"]).sum()
    print("volume", before, after, (after - before) / before, "ms", info["ms"])
    assert abs(after - before) <= 1e-11 * before


def test_vof_rejects_open_ring():
    """A ring cell that is not solid would send the neighbour reads out of
    the box: the entry refuses the flags before any flux kernel runs."""
    from clowns3d import DeviceGrid, Solver, assembly
    from clowns3d._lib import CLOWNS_EBADARG, ClownsError
    g, A, Fp, dt = vof_case("wavy20x16x12", "step", seed=1)
    c = assembly.classify(g)
    nf = c["nf"].copy()
    nf[g.X * g.Y * 3] = 0  # i = 0, j = 0, k = 3
    c = dict(c, nf=nf)
    sol = Solver(0)
    dg = DeviceGrid(g, classified=c)
    with pytest.raises(ClownsError) as e:
        dg.vof(sol, Fp, mode=1, dt=dt)
    assert e.value.code == CLOWNS_EBADARG
    sol.close()


def test_vof_rejects_object_cells():
    """Object cells of the case (boundary objects, appended cells) are not
    covered by the VOF update: the entry refuses the grid before any flux
    kernel runs."""
    from clowns3d import DeviceGrid, Solver, configs
    from clowns3d._lib import CLOWNS_EBADARG, ClownsError
    g = configs.config_objects(20, 12, 10, seed=3)
    sol = Solver(0)
    dg = DeviceGrid(g)
    with pytest.raises(ClownsError) as e:
        dg.vof(sol, dg.t["F"].clone(), mode=1, dt=g.dt)
    assert e.value.code == CLOWNS_EBADARG
    sol.close()


@pytest.mark.parametrize("name,scenario", [("bathy28x24x16", "stress"), ("wavy32x24x16", "step"),
                                           ("solitary40x8x12", "stress")])
def test_vof_plic_volumes_match_oracle(name, scenario):
    """quvwf_normal on the device from PLIC segment volumes
    (clowns_solver_vof_plic) equals the oracle's bit for bit.  Parity with
    the reference unpinned: the oracle's plic_flux is itself checked only
    against its formula (the reference's PLIC state is out of scope)."""
    from clowns3d import DeviceGrid, Solver, assembly
    from _vof import plic_volumes
    g, A, Fp, dt = vof_case(name, scenario, seed=9)
    c = assembly.classify(g)
    cells, vol = plic_volumes(g, A, c["nf"], seed=4, frac=0.8)
    o = call_oracle_vof(g, A, c["nf"], c["nb"], c["F"], Fp, dt, 2, plic=(cells, vol))
    assert o["rc"] == VOF_OK
    sol = Solver(0)
    dg = DeviceGrid(g, classified=c)
    Fn, qf, info = dg.vof(sol, Fp, dt=dt, plic=(cells, vol))
    print("oracle", {k: o[k] for k in ("retries", "modified", "slid", "dt")}, "gpu", info, flush=True)
    assert np.array_equal(qf.cpu().numpy(), o["qf"]) and np.array_equal(Fn.cpu().numpy(), o["F"])
    assert info["modified"] == o["modified"] and info["slid"] == o["slid"] and info["dt"] == o["dt"]
    sol.close()


def test_vof_plic_rejects_solid_cell():
    from clowns3d import DeviceGrid, Solver, assembly
    from clowns3d._lib import CLOWNS_EBADARG, ClownsError
    g, A, Fp, dt = vof_case("wavy20x16x12", "step", seed=1)
    c = assembly.classify(g)
    cells = np.array([np.nonzero(c["nf"] == 2)[0][0], 0], np.int32)  # cell 0: the solid ring
    vol = np.zeros((2, 6))
    sol = Solver(0)
    dg = DeviceGrid(g, classified=c)
    with pytest.raises(ClownsError) as e:
        dg.vof(sol, Fp, dt=dt, plic=(cells, vol))
    assert e.value.code == CLOWNS_EBADARG
    sol.close()
