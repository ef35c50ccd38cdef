"""Synthetic PCGPME inputs for tests (numpy only, vectorised).

``box_system`` lays a 7-point operator over an nx*ny*nz box in the
reference's k-major row order with the unsorted slot order
[x-, x+, y-, y+, z-, z+] of init_matrix (src2LOWNS/1sub-matrix.F90:33-64):
fluid rows carry -mp*2/dx/(dx+dxm)*a/fb off-diagonals and D = -sum(ALU)
(1sub-matrix.F90:50-61, 82-87); the top layer are free-surface rows with
D = 1, one real off-diagonal towards the fluid below and tangential
entries rate = -1e-20 (src2LOWNS/1sub-momentum.F90:473,488-543).  Face
open ratios and fb are seeded random in [lo, 1] so rows are non-uniform
and the operator is non-symmetric, as the reference's is.
"""
import numpy as np


def box_system(nx, ny, nz, seed=0, lo=0.3, na_extra=60, dx=1.0, dz=0.5, rhs=True):
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    NA = N + na_extra
    mp = min(dx, dz) ** 2
    shp = (nz, ny, nx)
    fb = rng.uniform(lo, 1.0, size=shp)
    a_lo = [rng.uniform(lo, 1.0, size=shp) for _ in range(3)]  # low-face ratios x, y, z
    f_surf = rng.uniform(0.55, 0.95, size=(ny, nx))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    row = (i + nx * (j + ny * k)).astype(np.int64)
    AL = np.zeros((NA, 6), order="F")
    LL = np.zeros((NA, 6), dtype=np.int32, order="F")
    D = np.zeros(N)
    surface = k == nz - 1
    fluid = ~surface
    h = [dx, dx, dz]
    strides = [1, nx, nx * ny]
    coords = [i, j, k]
    dims = [nx, ny, nz]
    de = np.zeros(shp)
    for l in range(3):
        for s, side in enumerate((-1, 1)):
            slot = 2 * l + s
            c2 = coords[l] + side
            ok = (c2 >= 0) & (c2 < dims[l])
            nb = row + side * strides[l]
            # the shared face is the low face of the cell with the larger index
            if side == -1:
                a = a_lo[l]
            else:
                a = np.zeros(shp)
                sl = [slice(None)] * 3
                sl[2 - l] = slice(1, None)
                src = a_lo[l][tuple(sl)]
                dl = [slice(None)] * 3
                dl[2 - l] = slice(0, -1)
                a[tuple(dl)] = src
            alu = -mp * (2.0 / h[l] / (h[l] + h[l]) * a / fb)
            # rows are numbered in (k, j, i) order, so row.ravel() == arange(N)
            mf = ok & fluid
            ms = ok & surface
            de = np.where(mf, de - alu, de)
            if l == 2 and side == -1:
                f = f_surf[None, :, :]
                asurf = -(f - 0.5) * dz / (f * dz + dz * 0.5)
            else:
                asurf = -1.0e-20
            AL[:N, slot] = np.where(mf, alu, np.where(ms, asurf, 0.0)).ravel()
            LL[:N, slot] = np.where(ok, nb + 1, 0).ravel()
    D[:] = np.where(surface, 1.0, de).ravel()
    B = rng.uniform(-1.0, 1.0, size=N) if rhs else np.zeros(N)
    return dict(AL=AL, LL=LL, D=D, B=B, N=N, NA=NA, NL=6)
