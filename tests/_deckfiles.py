"""Synthetic 2CLOWNS-3D case files for the deck tests: a box with a solid
ring and a solid block inside, written as Fortran unformatted sequential
records (4-byte length, payload, length) in either byte order, in the dense
(".xyz" grid, nf/nfb in the data file) or the sparse (".xyzn" grid listing
the nf = -1 cells, data file listing cells with state) form that
src2LOWNS/data_io.F90 reads (read_data :87-300, read_fdata :304-330).

Along with the files, the generator returns what a reader must produce,
per box cell c = i + X*(j + Y*k) (i, j, k counted from the reference's
is-1, js-1, ks-1): solid, U, F, P, a, fb, the flags nf/nb the reference
keeps for each numbered cell, and the coordinates.
"""
import struct

import numpy as np


def _rec(f, fmt, vals, be):
    e = ">" if be else "<"
    payload = struct.pack(e + fmt, *vals) if fmt else b""
    m = struct.pack(e + "i", len(payload))
    f.write(m + payload + m)


# object codes of nf = -1 cells (the box model's solid values), from nfb
def object_code(nfb):
    if nfb == -1:
        return -127
    if nfb <= 0:
        return 1
    if nfb >= 200:
        return -1 if nfb == 300 else -128
    if nfb == 1:
        return -2
    if nfb >= 127:
        return 124 - nfb
    return nfb


def object_nfb(code):
    """nfb of an object cell from its box code (the inverse of object_code)."""
    if code == -127:
        return -1
    if code == -1:
        return 300
    if code == -128:
        return 200
    if code == -2:
        return 1
    if code <= -3:
        return 124 - code
    return code


def make_case(tmp, nx, ny, nz, seed, sparse, be=False, lo=(3, 3, 3), t0=12.5, with_dt0=True, obstacle=True,
              objects=False, fdata=True, sparse_frac=0.6):
    rng = np.random.default_rng(seed)
    X, Y, Z = nx + 2, ny + 2, nz + 2
    nc = X * Y * Z
    is_, js, ks = lo
    ie, je, ke = is_ + nx, js + ny, ks + nz
    solid = np.zeros((Z, Y, X), np.int8)
    solid[0], solid[-1] = 1, 1
    solid[:, 0], solid[:, -1] = 1, 1
    solid[:, :, 0], solid[:, :, -1] = 1, 1
    if obstacle:
        solid[1:1 + max(1, nz // 2), Y // 3:Y // 3 + 2, X // 2:X // 2 + 2] = 1
    solid = solid.ravel()
    nf = np.where(solid == 1, -1, rng.integers(0, 3, nc)).astype(np.int32)
    nfb = np.where(solid == 1, 0, rng.integers(0, 2, nc)).astype(np.int32)
    if objects:
        # boundary objects (1 <= nfb < 200), appended objects (200, 300) and
        # nfb = -1 cells (data_io.F90:168-199), inside and on the walls
        free = np.nonzero(solid == 0)[0]
        kinds = [4, 5, 1, 150, 7, 200, 300, 300, -1, 4, 200]
        for c, b in zip(rng.choice(free, len(kinds), replace=False), kinds):
            nf[c], nfb[c] = -1, b
        wall = np.nonzero(solid == 1)[0]
        for c, b in zip(rng.choice(wall, 4, replace=False), (4, 5, 300, 1)):
            nfb[c] = b
    solid = np.array([0 if nf[c] >= 0 else object_code(int(nfb[c])) for c in range(nc)], np.int8)
    # numbering (data_io.F90:168-199): j outer, k, i inner over flow cells and
    # boundary objects, then the same sweep over appended objects
    sweep = [i + X * (j + Y * k) for j in range(Y) for k in range(Z) for i in range(X)]
    main = [c for c in sweep if nf[c] >= 0 or 1 <= nfb[c] < 200]
    app = [c for c in sweep if nf[c] < 0 and nfb[c] >= 200]
    cells = np.array(main + app, np.int64)
    inne = len(main)
    inne2 = len(cells)
    # coordinates X(l, ls-2 .. le+1): n+4 faces, cell sizes on a 0.01 grid
    coords = []
    for n, base in zip((nx, ny, nz), (-3.0, 10.0, -1.5)):
        h = rng.integers(20, 90, n + 3) / 100.0
        coords.append(np.concatenate([[base], base + np.cumsum(h)]))
    U = np.zeros((3, nc))
    F = np.zeros(nc)
    P = np.zeros(nc)
    # without an open-ratio file: cells 1..inne open, appended ones closed
    opened = np.zeros(nc, bool)
    opened[main] = True
    A = np.where(opened, 1.0, 0.0)[None].repeat(3, 0)
    FB = np.where(opened, 1.0, 0.0)
    nf_keep = nf.copy()
    nb_keep = nfb.copy()
    for c in app:  # read_data resets the appended cells' nfb (data_io.F90:190-194)
        nb_keep[c] = -1 if nfb[c] == 300 else 0
    listed = np.ones(inne, bool)
    if sparse:
        listed = rng.random(inne) < sparse_frac
    for q, c in enumerate(main):
        if listed[q]:
            U[:, c] = rng.standard_normal(3)
            F[c] = rng.random()
            P[c] = rng.standard_normal() * 100
    for c in app:
        U[:, c] = rng.standard_normal(3)
    if sparse:
        # the data records carry the flags of the cells they list (an object
        # keeps its own); unlisted flow cells keep the grid's nf = 0, nfb = 0
        for q, c in enumerate(main):
            if nf[c] < 0:
                continue
            if listed[q]:
                nf_keep[c] = int(rng.integers(0, 3))
                nb_keep[c] = int(rng.integers(0, 2))
            else:
                nf_keep[c] = nb_keep[c] = 0
    a_listed = rng.random(inne) < 0.5
    for q, c in enumerate(main):
        if not sparse or a_listed[q]:
            A[:, c] = rng.integers(0, 101, 3) / 100.0
            FB[c] = rng.integers(1, 101) / 100.0
    app_listed = rng.random(len(app)) < 0.7
    for q, c in enumerate(app):
        if not sparse or app_listed[q]:
            A[:, c] = rng.integers(0, 101, 3) / 100.0
    g, nu, dt0 = 9.8, 1.0e-6, 0.0025
    ext = (is_, js, ks, ie, je, ke)
    gname = tmp / ("case.xyzn" if sparse else "case.xyz")
    dname = tmp / "case.data"
    fname = tmp / "case.fdata0"
    with open(gname, "wb") as f:
        _rec(f, "6i", ext, be)
        for xs in coords:
            _rec(f, f"{len(xs)}d", xs, be)
        if sparse:
            sol = [(i, j, k) for i in range(X) for j in range(Y) for k in range(Z) if nf[i + X * (j + Y * k)] == -1]
            _rec(f, "i", [len(sol)], be)
            for i, j, k in sol:
                c = i + X * (j + Y * k)
                _rec(f, "5i", [i + is_ - 1, k + ks - 1, j + js - 1, nf[c], nfb[c]], be)
    with open(dname, "wb") as f:
        _rec(f, "3d", [t0, g, nu], be)
        if not sparse:
            # (((NF3(I,k,j),K=ks-1,KE),J=js-1,je),I=is-1,IE): k fastest, i slowest
            order = [i + X * (j + Y * k) for i in range(X) for j in range(Y) for k in range(Z)]
            _rec(f, f"{nc}i", nf[order], be)
            _rec(f, f"{nc}i", nfb[order], be)
            for c in main:
                _rec(f, "5d", [*U[:, c], F[c], P[c]], be)
            for c in app:
                _rec(f, "3d", [*U[:, c]], be)
        else:
            idx = np.nonzero(listed)[0]
            _rec(f, "2i", [len(idx), len(app)], be)
            for q in idx:
                c = main[q]
                _rec(f, "i5d2i", [q + 1, *U[:, c], F[c], P[c], nf_keep[c], nb_keep[c]], be)
            for q, c in enumerate(app):
                _rec(f, "i3d", [inne + q + 1, *U[:, c]], be)
        if with_dt0:
            _rec(f, "d", [dt0], be)
    if fdata:
        with open(fname, "wb") as f:
            if not sparse:
                for c in main:
                    _rec(f, "4d", [*A[:, c], FB[c]], be)
                for c in app:
                    _rec(f, "3d", [*A[:, c]], be)
            else:
                idx = np.nonzero(a_listed)[0]
                aidx = np.nonzero(app_listed)[0]
                _rec(f, "2i", [len(idx), len(aidx)], be)
                for q in idx:
                    c = main[q]
                    _rec(f, "i4d", [q + 1, *A[:, c], FB[c]], be)
                for q in aidx:
                    _rec(f, "i3d", [inne + q + 1, *A[:, app[q]]], be)
    return dict(grid=gname, data=dname, fdata=fname if fdata else None, ext=ext, nx=nx, ny=ny, nz=nz,
                X=X, Y=Y, Z=Z, nc=nc, solid=solid, U=U, F=F, P=P, A=A, FB=FB, nf=nf_keep, nb=nb_keep,
                cells=cells, inne=inne, inne2=inne2, coords=coords, t0=t0, g=g, nu=nu, dt0=dt0)


def write_grid_case(tmp, g, nf, nb, p, sparse=False, be=False, lo=(3, 3, 3), t0=1.0, g_acc=9.8, nu=1.0e-6):
    """Case files holding a configs.Grid as it is: grid file (faces
    X(l, ls-2 .. le+1) continuing g.xf), data file (nf/nb the flags of the
    flow cells; cells with g.solid != 0 get nf = -1 and, for object cells,
    the nfb of their code; U, F, p; dt0 = g.dt) and open-ratio file (g.a,
    g.fb).  Object cells follow read_data's numbering (data_io.F90:168-199):
    boundary objects among the flow cells, appended objects (nfb >= 200)
    after them with U and a0 only.  Returns the three paths."""
    X, Y, Z = g.X, g.Y, g.Z
    nc = g.ncell
    sol = np.asarray(g.solid).ravel()
    nf = np.where(sol != 0, -1, np.asarray(nf).ravel()).astype(np.int32)
    nfb = np.where(sol == 0, np.asarray(nb).ravel(), 0).astype(np.int32)
    for c in np.nonzero((sol != 0) & (sol != 1))[0]:
        nfb[c] = object_nfb(int(sol[c]))
    un = g.un.reshape(3, -1)
    F = g.F.ravel()
    a = g.a.reshape(3, -1)
    fb = g.fb.ravel()
    p = np.asarray(p).ravel()
    is_, js, ks = lo
    ext = (is_, js, ks, is_ + g.nx, js + g.ny, ks + g.nz)
    sweep = [i + X * (j + Y * k) for j in range(Y) for k in range(Z) for i in range(X)]
    cells = [c for c in sweep if nf[c] >= 0 or 1 <= nfb[c] < 200]
    app = [c for c in sweep if nf[c] < 0 and nfb[c] >= 200]
    inne = len(cells)
    gname = tmp / ("case.xyzn" if sparse else "case.xyz")
    dname, fname = tmp / "case.data", tmp / "case.fdata0"
    with open(gname, "wb") as f:
        _rec(f, "6i", ext, be)
        for l, n in enumerate((g.nx, g.ny, g.nz)):
            h = g.xf[l, 1] - g.xf[l, 0]
            faces = np.concatenate([[g.xf[l, 0] - h], g.xf[l, :n + 2], [g.xf[l, n + 1] + h]])
            _rec(f, f"{n + 4}d", faces, be)
        if sparse:
            sol = [(i, j, k) for i in range(X) for j in range(Y) for k in range(Z) if nf[i + X * (j + Y * k)] == -1]
            _rec(f, "i", [len(sol)], be)
            for i, j, k in sol:
                c = i + X * (j + Y * k)
                _rec(f, "5i", [i + is_ - 1, k + ks - 1, j + js - 1, -1, nfb[c]], be)
    with open(dname, "wb") as f:
        _rec(f, "3d", [t0, g_acc, nu], be)
        if not sparse:
            order = [i + X * (j + Y * k) for i in range(X) for j in range(Y) for k in range(Z)]
            _rec(f, f"{nc}i", nf[order], be)
            _rec(f, f"{nc}i", nfb[order], be)
            for c in cells:
                _rec(f, "5d", [*un[:, c], F[c], p[c]], be)
            for c in app:
                _rec(f, "3d", [*un[:, c]], be)
        else:
            _rec(f, "2i", [len(cells), len(app)], be)
            for q, c in enumerate(cells):
                _rec(f, "i5d2i", [q + 1, *un[:, c], F[c], p[c], nf[c], nfb[c]], be)
            for q, c in enumerate(app):
                _rec(f, "i3d", [inne + q + 1, *un[:, c]], be)
        _rec(f, "d", [g.dt], be)
    with open(fname, "wb") as f:
        if not sparse:
            for c in cells:
                _rec(f, "4d", [*a[:, c], fb[c]], be)
            for c in app:
                _rec(f, "3d", [*a[:, c]], be)
        else:
            _rec(f, "2i", [len(cells), len(app)], be)
            for q, c in enumerate(cells):
                _rec(f, "i4d", [q + 1, *a[:, c], fb[c]], be)
            for q, c in enumerate(app):
                _rec(f, "i3d", [inne + q + 1, *a[:, c]], be)
    return gname, dname, fname
