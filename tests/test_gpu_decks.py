"""GPU: a case read from the reference's file formats runs the device
pressure-correction loop exactly as the same Grid built in memory does
(bit-identical pressure and velocities), and the data file written with the
solved pressure (clowns_deck_write_data) reads back to the same values."""
import numpy as np
import pytest

from _deckfiles import write_grid_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sparse", [False, True])
def test_case_files_to_pressure_and_back(tmp_path, sparse):
    import torch
    from clowns3d import DeviceGrid, Solver, assembly, decks
    from test_oracle_assembly import GRIDS
    g0 = GRIDS["wavy32x24x16"]()
    g0.un[:, g0.solid == 1] = 0.0  # the files carry velocities of numbered cells only
    c = assembly.classify(g0)
    paths = write_grid_case(tmp_path, g0, c["nf"], c["nb"], np.zeros(g0.ncell), sparse=sparse, be=True)
    g1, _, d = decks.read_case(*paths)
    out = []
    for g in (g0, g1):
        sol = Solver(0)
        dg = DeviceGrid(g)
        p, info = dg.pressure_correction(sol, d0=1e-5, eps=1e-10)
        torch.cuda.synchronize()
        assert info["ier"] == 0 and info["passes"] >= 1
        out.append((p.cpu().numpy(), dg.t["un"].view(3, -1).cpu().numpy(), dg.t["nf"].cpu().numpy(),
                    dg.t["nb"].cpu().numpy()))
        sol.close()
    (p0, u0, _, _), (p1, u1, nf, nb) = out
    assert np.array_equal(p0, p1) and np.array_equal(u0, u1)
    res = tmp_path / "solved.data"
    decks.write_data(res, d, nf, nb, u1, g1.F, p1, 1.0 + g1.dt, g1.dt, big_endian=True)
    g2, p2, d2 = decks.read_case(paths[0], res, paths[2])
    num = g1.solid.ravel() == 0
    if sparse:  # cells without state are not listed and read back as zero
        num &= ~((nf == 0) & ((u1 ** 2).sum(0) == 0) & (g1.F.ravel() < 1e-13))
    assert np.array_equal(p2.ravel()[num], p1[num])
    assert np.array_equal(g2.un.reshape(3, -1)[:, num], u1[:, num])
    assert d2.t0 == 1.0 + g1.dt and d2.dt0 == g1.dt
