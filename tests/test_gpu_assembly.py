"""GPU assembly against the oracle.

clowns_assemble (init_matrix + cal_matrix_element + RHS on the device) must
equal the oracle's restatement (on the test grids and on configurations 1-3
at full size, there after the device classification) -- itself bit-exact against the flang build of
the reference routines -- bit for bit: AL, LL, D, B.  The fused
clowns_solver_assemble must leave the same scaled/sorted operator as
assemble -> set_matrix (PCGPME's scaling and sort), and its solve must match
the oracle's PCGPME on the oracle-built system within the reference's own
thread-count spread."""
import numpy as np
import pytest
import torch

from clowns3d import DeviceGrid, Solver
from conftest import call_oracle
from test_oracle_assembly import GRIDS, build_oracle
from clowns3d import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_device_assembly_bit_exact(name):
    g = GRIDS[name]()
    o = build_oracle(g)
    N = o["N"]
    dg = DeviceGrid(g)
    AL, LL, D, B = dg.assemble()
    torch.cuda.synchronize()
    assert dg.N == N
    assert np.array_equal(LL.cpu().numpy(), o["LL"][:, :N])
    assert np.array_equal(AL.cpu().numpy(), o["AL"][:, :N])
    assert np.array_equal(D.cpu().numpy(), o["D"][:N])
    assert np.array_equal(B.cpu().numpy(), o["B"][:N])


@pytest.mark.parametrize("config", [1, 2, 3])
def test_device_assembly_full_size(config):
    """The benchmark configurations at full size (77,776 / 2.45M / 13.07M
    rows): classification and row numbering on the GPU, then the assembly,
    bit for bit against the oracle's."""
    g = configs.CONFIGS[config]()
    o = build_oracle(g)
    N = o["N"]
    sol = Solver(0)
    dg = DeviceGrid.on_device(g, sol)
    assert dg.N == N and not dg.last_fallback
    assert np.array_equal(dg.row2cell.cpu().numpy(), o["row2cell"][:N])
    AL, LL, D, B = dg.assemble()
    torch.cuda.synchronize()
    print(f"configuration {config}: N={N}")
    assert np.array_equal(AL.cpu().numpy(), o["AL"][:, :N])
    assert np.array_equal(LL.cpu().numpy(), o["LL"][:, :N])
    assert np.array_equal(D.cpu().numpy(), o["D"][:N])
    assert np.array_equal(B.cpu().numpy(), o["B"][:N])
    sol.close()


@pytest.mark.parametrize("name", ["wavy32x24x16", "bathy28x24x16", "objects40x24x16"])
def test_fused_assembly_and_solve(name):
    g = GRIDS[name]()
    o = build_oracle(g)
    N = o["N"]
    dg = DeviceGrid(g)
    # unfused: assemble -> set_matrix
    AL, LL, D, B = dg.assemble()
    s1 = Solver(0)
    s1.set_matrix(AL, LL, D)
    A1, L1, D1, W1, _ = s1.get_scaled()
    # fused
    s2 = Solver(0)
    b = dg.assemble_into(s2, keep_raw=True)
    A2, L2, D2, W2, _ = s2.get_scaled()
    lda2 = A2.shape[1]
    assert torch.equal(A1[:, :N], A2[:, :N]) and torch.equal(L1[:, :N], L2[:, :N])
    assert torch.equal(D1, D2) and torch.equal(W1, W2) and torch.equal(b, B)
    x, info = s2.solve(b)
    sysd = dict(AL=np.asfortranarray(o["AL"][:, :N].T.copy()), LL=np.asfortranarray(o["LL"][:, :N].T.copy()),
                D=o["D"][:N].copy(), B=o["B"][:N].copy(), N=N, NA=N, NL=6)
    ref = call_oracle(sysd)
    tight = call_oracle(sysd, eps=2e-15)["X"][:N]
    x = x.cpu().numpy()
    e_gpu = np.linalg.norm(x - tight) / np.linalg.norm(tight)
    e_ref = np.linalg.norm(ref["X"][:N] - tight) / np.linalg.norm(tight)
    print(name, "N", N, "iters", info["iters"], ref["ITER"], "err", e_gpu, e_ref, "lda", lda2)
    assert info["ier"] == 0 and ref["IER"] == 0
    assert e_gpu <= max(1e-9, 3 * e_ref)
