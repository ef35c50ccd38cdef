"""GPU: the drop-in entry points on systems whose arrays span several of the
32 MB pinned chunks the host transfers go through (each way, a chunk is
filled or drained by host threads while the other one's DMA runs).

- clowns_pcgpme at 5.5M rows (each AL/LL column 22-44 MB): X, ITER and the
  arrays the reference leaves behind (AL, D scaled, LL sorted, B the final
  scaled residual; pcgpme_para.F90:77-112,250-257) equal, bit for bit, the
  resident solver's solve of the same system and its scaled operator.
- clowns_pcgpme_will at 2304x2048 (4.7M rows): B equals the resident
  solve_will's output bit for bit.
- Back-to-back calls of different sizes reuse the chunks (a small solve
  after the large one equals a fresh solver's)."""
import numpy as np
import pytest
import torch

from conftest import host_config_system

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_pcgpme_dropin_multi_chunk():
    import clowns3d
    from clowns3d import Solver, configs
    g = configs.config_wavy_box(384, 384, 64, (144, 240, 144, 240, 0, 24))
    s = host_config_system(g)
    N, NA = s["N"], s["NA"]
    assert N * 8 > (32 << 20)  # a column spans more than one chunk
    print(f"drop-in: N={N}", flush=True)
    AL, LL, D, B = (s[k].copy(order="F") for k in ("AL", "LL", "D", "B"))
    X = np.zeros(N + 64)
    EPS = np.array([1e-10, 0.0])
    it, ier = clowns3d.pcgpme(AL, 6, NA, LL, D, N, 64, B, X, EPS, 5.0, 10000)
    assert ier == 0

    sol = Solver(0)
    sol.set_matrix(_dev(s["AL"][:N].T), _dev(s["LL"][:N].T), _dev(s["D"]))
    x, info = sol.solve(_dev(s["B"]), eps=1e-10)
    ALs, LLs, Ds, _, Rs = sol.get_scaled(residual=True)
    torch.cuda.synchronize()
    print(f"  ITER {it} / {info['iters']}, ERR {EPS[0]:.3e}", flush=True)
    assert it == info["iters"] and EPS[0] == info["err"]
    assert np.array_equal(X[:N], x.cpu().numpy())
    assert np.array_equal(AL[:N], ALs[:, :N].cpu().numpy().T)
    assert np.array_equal(LL[:N], LLs[:, :N].cpu().numpy().T)
    assert np.array_equal(D, Ds.cpu().numpy())
    assert np.array_equal(B, Rs.cpu().numpy())
    # the padding rows of the caller's arrays are not touched
    assert np.array_equal(AL[N:], s["AL"][N:]) and np.array_equal(LL[N:], s["LL"][N:])
    sol.close()

    # a small system through the same chunks, against a fresh solver
    gs = configs.config_wavy_box(32, 24, 16, (8, 20, 8, 16, 0, 6), seed=8)
    ss = host_config_system(gs)
    n = ss["N"]
    ALb, LLb, Db, Bb = (ss[k].copy(order="F") for k in ("AL", "LL", "D", "B"))
    Xb = np.zeros(n + 64)
    EPSb = np.array([1e-10, 0.0])
    itb, ierb = clowns3d.pcgpme(ALb, 6, ss["NA"], LLb, Db, n, 64, Bb, Xb, EPSb, 5.0, 10000)
    ref = Solver(0)
    ref.set_matrix(_dev(ss["AL"][:n].T), _dev(ss["LL"][:n].T), _dev(ss["D"]))
    xr, ir = ref.solve(_dev(ss["B"]), eps=1e-10)
    assert ierb == 0 and itb == ir["iters"]
    assert np.array_equal(Xb[:n], xr.cpu().numpy())
    ref.close()


def test_pcgpme_will_dropin_multi_chunk():
    import clowns3d
    from clowns3d import Solver, configs
    s = configs.dispersion_2dh(2304, 2048, depth=4000.0, dx=(400.0, 500.0))
    N, NA = s["N"], s["NA"]
    assert N * 8 > (32 << 20)
    print(f"PCGPME_WILL drop-in: N={N}", flush=True)
    B = s["B"].copy()
    ier = clowns3d.pcgpme_will(s["AL"].copy(order="F"), 4, NA, s["LL"].copy(order="F"), s["D"].copy(), N, 64, B,
                               1e-10, 5.0, 1000)
    assert ier == 0
    sol = Solver(0)
    sol.set_matrix(_dev(s["AL"][:N].T), _dev(s["LL"][:N].T), _dev(s["D"]))
    out, info = sol.solve_will(_dev(s["B"]), eps=1e-10, iters=1000)
    torch.cuda.synchronize()
    assert np.array_equal(B, out.cpu().numpy())
    sol.close()
