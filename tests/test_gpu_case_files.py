"""GPU: a case in the reference's own file formats, from the files to the
corrected velocities, against the reference's reader.

The same files -- a wavy free surface with object cells of every kind the
reader numbers (pressure-boundary cells, inflow boundary objects, appended
obstacle cells), in the dense and the sparse form -- go two ways:
- the product: clowns_deck_open/read (read_data's numbering, set_dx) into
  the box grid, then on the GPU the classification (set_nflag), the row
  numbering, the assembly (init_matrix, cal_matrix_element, the RHS) and one
  pressure-correction pass (solve, p -= xe, estimate_velocity2, DMAX);
- the flang build of the reference's read_data and read_fdata
  (src2LOWNS/data_io.F90:87-364) and set_dx (src2LOWNS/1sub-main.F90:53-74):
  its numbered cells placed where its own (i, j, k) say, then the oracle's
  classification, assembly and pass on that grid.
Bar: the row numbering, AL, LL, D and B bit for bit; DMAX before the pass
exactly; p (starting from the file's pressure) and un within the solve
tolerance (the two solves differ in dot-product summation order only)."""
import numpy as np
import pytest
import torch

from _deckfiles import write_grid_case
from conftest import call_oracle
from test_decks import reference_grid
from test_oracle_assembly import build_oracle, oracle_correction_pass, oracle_rhs_dmax

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sparse", [False, True])
def test_object_case_files_to_correction_pass(tmp_path, sparse):
    from clowns3d import DeviceGrid, Solver, assembly, configs, decks
    g0 = configs.config_objects(40, 24, 16, seed=11)
    c = assembly.classify(g0)
    p0 = np.where(g0.solid.ravel() == 1, 0.0, np.random.default_rng(7).standard_normal(g0.ncell) * 50)
    paths = write_grid_case(tmp_path, g0, c["nf"], c["nb"], p0, sparse=sparse)
    g1, p1, d = decks.read_case(*paths)
    gr, pr = reference_grid(paths, g1)
    print(f"objects case ({'sparse' if sparse else 'dense'}): inne {d.inne}, inne2 {d.inne2}", flush=True)

    # the oracle on the reference reader's grid
    o = build_oracle(gr)
    N = o["N"]
    A = gr.arrays()
    un_o = A["un"].copy()
    p_o = pr.ravel().copy()
    B, dm0 = oracle_rhs_dmax(gr, o, un_o)
    assert np.array_equal(B, o["B"][:N])
    sysd = dict(AL=np.asfortranarray(o["AL"][:, :N].T.copy()), LL=np.asfortranarray(o["LL"][:, :N].T.copy()),
                D=o["D"][:N].copy(), B=B.copy(), N=N, NA=N, NL=6)
    r = call_oracle(sysd, eps=1e-10)
    assert r["IER"] == 0
    _, dm1 = oracle_correction_pass(gr, o, r["X"][:N].copy(), p_o, un_o)

    # the product from the files, classified and assembled on the GPU
    sol = Solver(0)
    dg = DeviceGrid.on_device(g1, sol)
    assert dg.N == N
    assert np.array_equal(dg.row2cell.cpu().numpy(), o["row2cell"][:N])
    AL, LL, D, Bd = dg.assemble()
    torch.cuda.synchronize()
    assert np.array_equal(LL.cpu().numpy(), o["LL"][:, :N])
    assert np.array_equal(AL.cpu().numpy(), o["AL"][:, :N])
    assert np.array_equal(D.cpu().numpy(), o["D"][:N])
    assert np.array_equal(Bd.cpu().numpy(), o["B"][:N])
    p = torch.from_numpy(p1.ravel().copy()).to(dg.device)
    p, info = dg.pressure_correction(sol, p=p, d0=1e-5, eps=1e-10, max_passes=1)
    torch.cuda.synchronize()
    print(f"  N {N}: oracle DMAX {dm0:.6e} -> {dm1:.6e}, ITER {r['ITER']} | gpu {info}", flush=True)
    assert info["passes"] == 2 and info["ier"] == 3000  # one solve, then ik = 2 > max_passes
    assert info["dmax_first"] == dm0
    assert abs(info["dmax_last"] - dm1) <= 1e-6 * dm1
    case = A["solid"] != 1
    pg = p.cpu().numpy()
    ug = dg.t["un"].view(3, -1).cpu().numpy()
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(pg[case], p_o[case]) < 1e-8
    assert rel(ug[:, case], un_o[:, case]) < 1e-8
    sol.close()
