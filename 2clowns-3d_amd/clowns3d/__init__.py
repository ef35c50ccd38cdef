"""clowns3d -- Python side of the MI355X pressure-Poisson solve of 2CLOWNS-3D.

The compute path is the gfx950 library ``lib/libclowns_pcg.so`` (C-ABI in
include/clowns_pcg.h); this package only binds it (ctypes), generates the
synthetic workloads and exposes the adjoint as a torch.autograd.Function.
"""
from ._lib import ClownsError, Info, LIB_PATH, lib  # noqa: F401
from .pcgpme import Solver, pcgpme, pcgpme_will  # noqa: F401
from . import configs  # noqa: F401,E402
from .assembly import DeviceGrid, classify, set_dx  # noqa: F401,E402
from .autograd import OpenRatioSolve, PCGSolve, pcg_solve, pcg_solve_open_ratios  # noqa: F401,E402
from . import decks  # noqa: F401,E402
