"""Synthetic 2CLOWNS-3D grids of the benchmark configurations.

A grid is a box of cells i=0..nx+1, j=0..ny+1, k=0..nz+1 whose outer ring is
solid (closed walls), stored flat with index c = i + X*(j + Y*k), X = nx+2,
Y = ny+2 -- the reference's k-major row order (src2LOWNS/nfd2015.F90:69-78).
Fields, as the reference holds them per cell (src2LOWNS/0mod-arrays.F90):
  solid  int8   0 = flow cell, 1 = wall cell outside the case (nf = -1,
                nfb = 0, so mn = 0), other values an object cell of the case
                (nf = -1, numbered, its own F, fb, a, un) with that object
                code: for a boundary object with nfb = b (1..199) b itself
                for 2..126 (4: a pressure boundary), -2 for 1, 124 - b for
                127..199; -1 / -128 the cells appended with nfb 300 / 200
  F      f64    VOF fill
  fb     f64    cell open ratio
  a      f64    (3, ncell) open ratio of the LOW face of the cell in x, y, z
  un     f64    (3, ncell) velocity on that face
  dx, xf f64    (3, stride) cell sizes and low-face coordinates per axis
How F, fb and a follow from the water level and the bed follows the grid
preprocessor (srcMakeGrid/set_landform.F90:25-92): a_x = min(fb_i, fb_{i-1})
and a_z = 1 between open cells, 0 against a solid cell; in a partial bed cell
F = (h - z_{k+1} + fb dz)/(dz fb).
"""
import math

import numpy as np


class Grid:
    def __init__(self, nx, ny, nz, dx=1.0, dy=1.0, dz=0.5, dt=0.01):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.X, self.Y, self.Z = nx + 2, ny + 2, nz + 2
        self.ncell = self.X * self.Y * self.Z
        self.dt = dt
        self.stride = max(nx, ny, nz) + 2
        self.h = (dx, dy, dz)
        self.dx = np.zeros((3, self.stride))
        self.xf = np.zeros((3, self.stride))
        for l, (n, h) in enumerate(zip((nx, ny, nz), self.h)):
            self.dx[l, : n + 2] = h
            # x(l, i) = (i-1)*h: low face of cell i; the ring cell 0 starts at -h
            self.xf[l, : n + 2] = (np.arange(n + 2) - 1) * h
        self.solid = np.zeros((self.Z, self.Y, self.X), np.int8)
        self.solid[0, :, :] = self.solid[-1, :, :] = 1
        self.solid[:, 0, :] = self.solid[:, -1, :] = 1
        self.solid[:, :, 0] = self.solid[:, :, -1] = 1
        self.F = np.zeros((self.Z, self.Y, self.X))
        self.fb = np.ones((self.Z, self.Y, self.X))
        self.a = np.zeros((3, self.Z, self.Y, self.X))
        self.un = np.zeros((3, self.Z, self.Y, self.X))
        self.name = ""

    # cell-centre coordinates of the interior, shaped (Y-2, X-2)
    def centres(self):
        hx, hy, _ = self.h
        xc = (np.arange(1, self.nx + 1) - 0.5) * hx
        yc = (np.arange(1, self.ny + 1) - 0.5) * hy
        return np.meshgrid(xc, yc, indexing="xy")

    def set_faces(self):
        """a on every low face: min of the two cells' fb in x and y, 1 in z,
        0 where either side is solid (set_landform.F90:25-41)."""
        open_ = self.solid == 0
        fb = np.where(open_, self.fb, 0.0)
        self.a[:] = 0.0
        both = open_[:, :, 1:] & open_[:, :, :-1]
        self.a[0][:, :, 1:] = np.where(both, np.minimum(fb[:, :, 1:], fb[:, :, :-1]), 0.0)
        both = open_[:, 1:, :] & open_[:, :-1, :]
        self.a[1][:, 1:, :] = np.where(both, np.minimum(fb[:, 1:, :], fb[:, :-1, :]), 0.0)
        both = open_[1:, :, :] & open_[:-1, :, :]
        self.a[2][1:, :, :] = np.where(both, 1.0, 0.0)

    def set_water(self, zs):
        """F from the water level zs(y, x) over the open part of each cell."""
        dz = self.h[2]
        k = np.arange(1, self.nz + 1)[:, None, None]
        zlo, zhi = (k - 1) * dz, k * dz
        zs3 = zs[None, :, :]
        fb = self.fb[1:-1, 1:-1, 1:-1]
        with np.errstate(divide="ignore", invalid="ignore"):
            part = (zs3 - zhi + fb * dz) / (dz * fb)
        F = np.where(zlo >= zs3, 0.0, np.where(zhi <= zs3, 1.0,
                     np.where(fb > (zhi - zs3) / dz, np.clip(part, 0.0, 1.0), 0.0)))
        F = np.where(self.solid[1:-1, 1:-1, 1:-1] == 1, 0.0, F)
        self.F[1:-1, 1:-1, 1:-1] = F

    def set_velocity(self, seed, amp=0.1):
        rng = np.random.default_rng(seed)
        self.un[:] = rng.uniform(-amp, amp, size=self.un.shape)

    def arrays(self):
        """Flat C-contiguous copies in the box order c = i + X*(j + Y*k).
        Solid cells carry F = fb = 0 and closed low faces, as the reference's
        sentinel cell 0 does (src2LOWNS/data_io.F90:274-277)."""
        sol = self.solid == 1
        self.F[sol] = 0.0
        self.fb[sol] = 0.0
        self.a[:, sol] = 0.0
        return dict(solid=np.ascontiguousarray(self.solid.ravel()),
                    F=np.ascontiguousarray(self.F.ravel()),
                    fb=np.ascontiguousarray(self.fb.ravel()),
                    a=np.ascontiguousarray(self.a.reshape(3, -1)),
                    un=np.ascontiguousarray(self.un.reshape(3, -1)),
                    dx=np.ascontiguousarray(self.dx), xf=np.ascontiguousarray(self.xf))


def wavy_surface(g):
    """z_s = 0.6 H + 0.3 sin(2 pi x / Lx) cos(2 pi y / Ly) at cell centres."""
    H = g.nz * g.h[2]
    xc, yc = g.centres()
    Lx, Ly = g.nx * g.h[0], g.ny * g.h[1]
    return 0.6 * H + 0.3 * np.sin(2 * math.pi * xc / Lx) * np.cos(2 * math.pi * yc / Ly)


def config_wavy_box(nx, ny, nz, box, seed=2024):
    """Configurations 1 and 2: wavy free surface at 0.6 H, a solid box obstacle
    over cells box = (i0, i1, j0, j1, k0, k1) (0-based, half-open), face open
    ratios 1 elsewhere, RHS from a seeded U(-0.1, 0.1) face velocity field."""
    g = Grid(nx, ny, nz)
    i0, i1, j0, j1, k0, k1 = box
    g.solid[1 + k0:1 + k1, 1 + j0:1 + j1, 1 + i0:1 + i1] = 1
    g.fb[:] = 1.0
    g.set_faces()
    g.set_water(wavy_surface(g))
    g.set_velocity(seed)
    g.name = f"wavy{nx}x{ny}x{nz}"
    return g


def open_object_faces(g, codes=(4, 5)):
    """Open (a = 1) every face between a flow cell and an object cell whose
    code is in `codes` (inflow and outflow boundaries are open to the flow;
    set_faces closes every face against a solid or object cell)."""
    for l, ax in enumerate((2, 1, 0)):  # a[l] is indexed (k, j, i): axis of l
        obj = np.isin(g.solid, codes)
        flow = g.solid == 0
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(None, -1)
        hi[ax] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        # face between cell m (low side) and cell c = m + e_l is a[l][c]
        pair = (obj[lo] & flow[hi]) | (flow[lo] & obj[hi])
        g.a[l][hi] = np.where(pair, 1.0, g.a[l][hi])


def config_objects(nx=40, ny=24, nz=16, seed=2030):
    """Wavy free surface with object cells of the case (boundary objects,
    src2LOWNS/data_io.F90:171-199): the high-x wall a pressure-boundary
    column (code 4, open to the flow), the low-x wall inflow boundary
    objects (codes 5, 7, -2, -30: nfb 5, 7, 1, 154; open), a 2x2x2 pressure-boundary block in the water and an
    obstacle whose cells are appended object cells (codes -1 and -128,
    nfb 300 and 200) carrying seeded fill and open ratios; every object
    cell keeps its own F, fb, a and un, as the case reader gives them."""
    g = config_wavy_box(nx, ny, nz, (nx // 3, nx // 3 + 4, ny // 3, ny // 3 + 4, 0, 3), seed)
    rng = np.random.default_rng(seed + 1)
    i0, j0 = nx // 3 + 1, ny // 3 + 1
    obst = g.solid[1:4, j0:j0 + 4, i0:i0 + 4]
    obst[:] = np.where(rng.random(obst.shape) < 0.5, -1, -128).astype(np.int8)
    g.F[1:4, j0:j0 + 4, i0:i0 + 4] = rng.uniform(0.0, 1.0, obst.shape)
    g.fb[1:4, j0:j0 + 4, i0:i0 + 4] = rng.uniform(0.2, 1.0, obst.shape)
    g.solid[1:-1, 1:-1, -1] = 4
    g.solid[1:-1, 1:-1, 0] = 5
    g.solid[1:-1, 2:-1:4, 0] = 7     # other boundary objects: nfb 7, 1 and 154
    g.solid[1:-1, 3:-1:4, 0] = -2
    g.solid[1:-1, 4:-1:4, 0] = -30
    g.F[:, :, -1] = g.F[:, :, -2]
    g.F[:, :, 0] = g.F[:, :, 1]
    ib, jb = 2 * nx // 3, 2 * ny // 3
    g.solid[2:4, jb:jb + 2, ib:ib + 2] = 4
    g.F[2:4, jb:jb + 2, ib:ib + 2] = 1.0
    open_object_faces(g, codes=(4, 5, 7, -2, -30))
    g.name = f"objects{nx}x{ny}x{nz}"
    return g


def config1(seed=2024):
    return config_wavy_box(64, 64, 32, (24, 40, 24, 40, 0, 12), seed)


def config2(seed=2024):
    return config_wavy_box(256, 256, 64, (96, 160, 96, 160, 0, 24), seed)


def config_bathymetry(nx, ny, nz, nbumps=40, seed=2025, width=(5, 30), height=(0.1, 0.4)):
    """Configuration 3: bed = sum of seeded Gaussian bumps (height U(0.1,0.4) H,
    width U(5,30) cells); partial bed cells get fb = (z_top - z_b)/dz clipped
    to [0.05, 1] (below 0.05 the cell is solid), a_x, a_y = min(fb) and a_z = 1
    between open cells; wavy free surface at 0.6 H."""
    g = Grid(nx, ny, nz)
    rng = np.random.default_rng(seed)
    H = nz * g.h[2]
    xc, yc = g.centres()
    zb = np.zeros_like(xc)
    for _ in range(nbumps):
        cx, cy = rng.uniform(0, nx * g.h[0]), rng.uniform(0, ny * g.h[1])
        hgt = rng.uniform(*height) * H
        w = rng.uniform(*width) * g.h[0]
        zb += hgt * np.exp(-((xc - cx) ** 2 + (yc - cy) ** 2) / (2 * w * w))
    dz = g.h[2]
    k = np.arange(1, nz + 1)[:, None, None]
    frac = (k * dz - zb[None]) / dz
    fb = np.clip(frac, 0.0, 1.0)
    solid_in = fb < 0.05
    g.solid[1:-1, 1:-1, 1:-1] |= solid_in.astype(np.int8)
    g.fb[1:-1, 1:-1, 1:-1] = np.where(solid_in, 0.0, np.clip(fb, 0.05, 1.0))
    g.fb[g.solid == 1] = 0.0
    g.set_faces()
    g.set_water(wavy_surface(g))
    g.set_velocity(seed + 1)
    g.name = f"bathy{nx}x{ny}x{nz}"
    return g


def config3(seed=2025):
    return config_bathymetry(512, 512, 96, seed=seed)


def config_solitary(nx, ny, nz, t, seed=2026, amp_frac=0.1, x0_frac=0.2, g_acc=9.80665):
    """Configuration 5 at time t: flat bed, still-water depth H = 0.6 * height,
    surface lifted by a solitary wave eta = A sech^2(kappa (x - x0 - c t)),
    A = 0.1 H, c = sqrt(g H), kappa = sqrt(3 A / (4 H^3))."""
    g = Grid(nx, ny, nz)
    g.fb[:] = 1.0
    g.set_faces()
    solitary_water(g, t, amp_frac, x0_frac, g_acc)
    g.set_velocity(seed)
    g.name = f"solitary{nx}x{ny}x{nz}@t={t:g}"
    return g


def solitary_water(g, t, amp_frac=0.1, x0_frac=0.2, g_acc=9.80665):
    """Set F of grid g for the solitary wave at time t (static geometry kept)."""
    H = 0.6 * g.nz * g.h[2]
    A = amp_frac * H
    c = math.sqrt(g_acc * H)
    kap = math.sqrt(3 * A / (4 * H ** 3))
    xc, _ = g.centres()
    x0 = x0_frac * g.nx * g.h[0]
    eta = A / np.cosh(kap * (xc - x0 - c * t)) ** 2
    g.set_water(H + eta)
    g.name = f"solitary{g.nx}x{g.ny}x{g.nz}@t={t:g}"


def solitary_fill_device(g, F_dev, t, amp_frac=0.1, x0_frac=0.2, g_acc=9.80665):
    """F of the solitary wave at time t computed on the GPU into the flat box
    tensor F_dev (same expressions as set_water; solid cells stay 0)."""
    import torch
    H = 0.6 * g.nz * g.h[2]
    A = amp_frac * H
    c = math.sqrt(g_acc * H)
    kap = math.sqrt(3 * A / (4 * H ** 3))
    dev = F_dev.device
    hx, _, dz = g.h
    xc = (torch.arange(1, g.nx + 1, device=dev, dtype=torch.float64) - 0.5) * hx
    x0 = x0_frac * g.nx * hx
    zs = H + A / torch.cosh(kap * (xc - x0 - c * t)) ** 2          # (nx,)
    k = torch.arange(1, g.nz + 1, device=dev, dtype=torch.float64)[:, None, None]
    zlo, zhi = (k - 1) * dz, k * dz
    zs3 = zs[None, None, :]
    F3 = F_dev.view(g.Z, g.Y, g.X)
    fb = torch.ones((), dtype=torch.float64, device=dev)
    part = (zs3 - zhi + fb * dz) / (dz * fb)
    F = torch.where(zlo >= zs3, 0.0, torch.where(zhi <= zs3, 1.0,
                    torch.where(fb > (zhi - zs3) / dz, part.clamp(0.0, 1.0), 0.0)))
    F = F.expand(g.nz, g.ny, g.nx)
    F3[1:-1, 1:-1, 1:-1] = F
    return F_dev


def config5(t=0.0, seed=2026):
    return config_solitary(1024, 512, 80, t=t, seed=seed)


def dispersion_2dh(nx, ny, seed=2027, dx=(8.0, 12.0), depth=10.0, hump=None, a_disp=0.4,
                   b_disp=1.0 / 15.0, g_acc=9.80665, nbumps=12):
    """PCGPME_WILL inputs: the 2DH linear-dispersion matrix as disp_param
    builds it (src2LOWNS/1sub-2DH-initial_reset.f90:472-505) on an nx*ny
    grid of cells i=1..nx, j=1..ny, and a right-hand side PSI of the shape
    disp_adjust gives it (1sub-2DH-momentum.f90:247-315, linear case
    PSI = (A - B) g ZK^2 hxx).

    Rows are the cells with bed ZK < 0, numbered i fastest (the reference
    numbers them in its cell-list order, 1sub-2DH-initial_reset.f90:229-235).
    Per row: x neighbours are skipped on the first and last column and y
    neighbours on the first and last row; each existing neighbour (x-, x+,
    y-, y+ order) takes the next slot, LL packed with zeros after the last,
    AL = A*2*ZK*ZK/DX/(DX + DX_neighbour) and DIAG = -1 - sum(AL).  The bed
    slopes from `depth` offshore to dry land over the last ~10% in x, with
    seeded Gaussian bumps; spacings are seeded in `dx` per column/row.  The
    free surface is a Gaussian hump (centre, radius, amplitude) cut to
    exactly zero beyond three radii, so PSI is exactly zero there and the
    solution decays through the output pass's 1e-14 limit.
    Returns dict(AL, LL (NA,4) Fortran order, D, B, N, NA=N, NL=4, mnd)."""
    rng = np.random.default_rng(seed)
    X, Y = nx + 2, ny + 2
    dx1 = np.broadcast_to(rng.uniform(dx[0], dx[1], size=X)[None, :], (Y, X)).copy()
    dx2 = np.broadcast_to(rng.uniform(dx[0], dx[1], size=Y)[:, None], (Y, X)).copy()
    x = np.concatenate([[0.0], np.cumsum(dx1[0, 1:])])  # cell positions (approximate centres)
    y = np.concatenate([[0.0], np.cumsum(dx2[1:, 0])])
    Lx, Ly = x[-1], y[-1]
    xx, yy = np.meshgrid(x, y)
    zk = -depth * (1.0 - 1.1 * xx / Lx)
    for _ in range(nbumps):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        w = rng.uniform(0.03, 0.1) * min(Lx, Ly)
        zk = zk + rng.uniform(0.05, 0.25) * depth * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / w ** 2)
    ii, jj = np.meshgrid(np.arange(X), np.arange(Y))
    inner = (ii >= 1) & (ii <= nx) & (jj >= 1) & (jj <= ny)
    wet = inner & (zk < 0.0)
    mnd = np.zeros((Y, X), dtype=np.int64)
    mnd[wet] = np.arange(1, int(wet.sum()) + 1)  # row-major over (j, i): i fastest
    N = int(wet.sum())
    # surface hump and PSI
    if hump is None:
        hump = (0.35 * Lx, 0.5 * Ly, 0.08 * min(Lx, Ly), 0.1 * depth)
    cx, cy, rad, amp = hump
    r2 = ((xx - cx) ** 2 + (yy - cy) ** 2) / rad ** 2
    eta = np.where(r2 <= 9.0, amp * np.exp(-r2), 0.0)
    hxx = np.zeros((Y, X))
    for d, (sp, ax) in enumerate(((dx1, 1), (dx2, 0))):
        em, ep = np.roll(eta, 1, axis=ax), np.roll(eta, -1, axis=ax)
        sm, spp = np.roll(sp, 1, axis=ax), np.roll(sp, -1, axis=ax)
        hm, hp = 0.5 * (sm + sp), 0.5 * (sp + spp)
        term = 2.0 * (hm * ep - (hm + hp) * eta + hp * em) / (hm * hp * (hm + hp))
        edge = (ii == 1) | (ii == nx) if d == 0 else (jj == 1) | (jj == ny)
        hxx = hxx + np.where(edge, 0.0, term)
    psi = (a_disp - b_disp) * g_acc * zk * zk * hxx
    # matrix
    rows = mnd[wet] - 1
    ci, cj = ii[wet], jj[wet]
    zc = zk[wet]
    AL = np.zeros((N, 4), order="F")
    LL = np.zeros((N, 4), dtype=np.int32, order="F")
    slot = np.zeros(N, dtype=np.int64)
    cand = []
    for d, (di, dj) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
        sp = dx1 if d < 2 else dx2
        skip = ((ci == 1) | (ci == nx)) if d < 2 else ((cj == 1) | (cj == ny))
        nb = mnd[cj + dj, ci + di]
        ok = (~skip) & (nb != 0)
        own = sp[cj, ci]
        other = sp[cj + dj, ci + di]
        val = a_disp * 2.0 * zc * zc / own / (own + other)
        cand.append((ok, nb, val))
    for ok, nb, val in cand:
        r = rows[ok]
        s = slot[ok]
        AL[r, s] = val[ok]
        LL[r, s] = nb[ok]
        slot[ok] += 1
    D = -1.0 - (((AL[:, 0] + AL[:, 1]) + AL[:, 2]) + AL[:, 3])
    B = np.ascontiguousarray(psi[wet])
    return dict(AL=AL, LL=LL, D=D, B=B, N=N, NA=N, NL=4, mnd=mnd)


CONFIGS = {1: config1, 2: config2, 3: config3}
