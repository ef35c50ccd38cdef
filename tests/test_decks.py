"""The case readers and the data-file writer (clowns_deck_open/read/
write_data) against the reference's own read_data, read_fdata and
output_data (src2LOWNS/data_io.F90:87-364, :708-783) built with flang into
oracle/_ref/libref_decks.so, on synthetic dense (".xyz") and sparse
(".xyzn") cases.

Bar, all bit-exact:
- every field the reader fills (U, F, P, a, fb, solid, coordinates, header,
  dt0) equals what the fixture wrote, in both byte orders;
- the reference's reader numbers the same cells in the same order and reads
  the same values into them;
- the writer's file is byte-identical to output_data's on the same state,
  and reading it back returns the same fields;
- cells with nf = -1 follow read_data's numbering (data_io.F90:168-199):
  boundary objects numbered with the flow cells, appended objects after
  them (U and a0 only), nfb = 0 and -1 cells not numbered; the reader gives
  each its object code, and the writer puts nfb back as output_data does;
- truncated files give CLOWNS_EDECK.
The reference's readers and writer are compared in both byte orders: the
little-endian form through oracle/_ref/libref_decks.so (flang's runtime in
a library keeps the host order), the big-endian form -- the one the
reference's own build writes -- through oracle/_ref/ref_decks_be, the same
routines under a main program compiled with -fconvert=big-endian.
"""
import ctypes
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from _deckfiles import make_case, object_code, write_grid_case
from conftest import REF_ASM_SO, ROOT, ptr

REF_DECKS_SO = os.path.join(ROOT, "oracle", "_ref", "libref_decks.so")
REF_DECKS_BE = os.path.join(ROOT, "oracle", "_ref", "ref_decks_be")
CASES = [("dense", False), ("sparse", True)]


def _ref_lib():
    if not os.path.exists(REF_DECKS_SO):
        pytest.skip("flang build of the reference (oracle/_ref) not present")
    return ctypes.CDLL(REF_DECKS_SO, mode=os.RTLD_LAZY)


def ref_read(case, data=None):
    """The reference's read_data + read_fdata on the case; returns per
    numbered cell its box index and values."""
    lib = _ref_lib()
    gp, dp = str(case["grid"]).encode(), str(data or case["data"]).encode()
    fp = str(case["fdata"]).encode()
    nmax = case["nc"]
    S = max(case["nx"], case["ny"], case["nz"]) + 8
    ext = np.zeros(6, np.int32)
    counts = np.zeros(2, np.int32)
    hdr = np.zeros(4)
    xo = np.zeros((3, S))
    cells = np.zeros((nmax, 3), np.int32)
    uo = np.zeros((nmax, 3))
    fo, po, fbo = np.zeros(nmax), np.zeros(nmax), np.zeros(nmax)
    ao = np.zeros((nmax, 3))
    flags = np.zeros((nmax, 2), np.int32)
    lib.ref_read_deck(gp, len(gp), dp, len(dp), fp, len(fp), ptr(ext), ptr(counts), ptr(hdr), ptr(xo), S, nmax,
                      ptr(cells), ptr(uo), ptr(fo), ptr(po), ptr(ao), ptr(fbo), ptr(flags))
    return _ref_dict(case, ext, counts, hdr, xo, cells, uo, fo, po, ao, fbo, flags)


def _ref_dict(case, ext, counts, hdr, xo, cells, uo, fo, po, ao, fbo, flags):
    n = int(counts[1])  # every numbered cell: 1..inne and the appended ones
    is_, js, ks = ext[:3]
    X, Y = case["X"], case["Y"]
    i, j, k = cells[:n, 0] - (is_ - 1), cells[:n, 1] - (js - 1), cells[:n, 2] - (ks - 1)
    return dict(ext=ext, inne=int(counts[0]), inne2=n, hdr=hdr, xo=xo, box=i + X * (j + Y * k),
                U=uo[:n].T, F=fo[:n], P=po[:n], A=ao[:n].T, FB=fbo[:n], flags=flags[:n])


def _hex(v):
    return struct.pack(">d", v).hex()


def ref_read_be(case, data=None, write=None):
    """read_data + read_fdata of the reference built with big-endian records
    (oracle/_ref/ref_decks_be, a main program compiled with
    -fconvert=big-endian) on big-endian case files; the same dict as
    ref_read.  write = (path, t, dt): output_data of the state read, written
    big-endian."""
    if not os.path.exists(REF_DECKS_BE):
        pytest.skip("flang build of the reference (oracle/_ref) not present")
    nmax = case["nc"]
    S = max(case["nx"], case["ny"], case["nz"]) + 8
    out = str(case["grid"]) + ".ref_be.txt"
    cmd = [REF_DECKS_BE, str(case["grid"]), str(data or case["data"]), str(case["fdata"]), str(nmax), str(S), out]
    if write is not None:
        cmd += [str(write[0]), _hex(write[1]), _hex(write[2])]
    subprocess.run(cmd, check=True, capture_output=True, timeout=60)
    with open(out) as f:
        vals = f.read().split()
    pos = 0

    def ints(n):
        nonlocal pos
        v = np.array([int(x) for x in vals[pos:pos + n]], np.int64)
        pos += n
        return v

    def reals(n):
        nonlocal pos
        v = np.array([struct.unpack(">d", bytes.fromhex(x))[0] for x in vals[pos:pos + n]])
        pos += n
        return v

    ext = ints(6)
    counts = ints(2)
    hdr = reals(4)
    xo = reals(3 * S).reshape(3, S)
    n = min(int(counts[1]), nmax)
    cells, flags = np.zeros((n, 3), np.int64), np.zeros((n, 2), np.int64)
    uo, ao = np.zeros((n, 3)), np.zeros((n, 3))
    fo, po, fbo = np.zeros(n), np.zeros(n), np.zeros(n)
    for q in range(n):
        cells[q] = ints(3)
        uo[q] = reals(3)
        fo[q], po[q] = reals(2)
        ao[q] = reals(3)
        fbo[q] = reals(1)[0]
        flags[q] = ints(2)
    assert pos == len(vals)
    return _ref_dict(case, ext, counts, hdr, xo, cells, uo, fo, po, ao, fbo, flags)


def ref_write(path, t, dt):
    """output_data on the state the last ref_read left."""
    lib = _ref_lib()
    p = str(path).encode()
    lib.ref_write_data(p, len(p), ctypes.c_double(t), ctypes.c_double(dt))


def ours(case, data=None, fdata="fixture"):
    from clowns3d import decks
    fd = case["fdata"] if fdata == "fixture" else fdata
    return decks.read_case(case["grid"], data or case["data"], fd)


def _check_fields(case, g, p, d):
    assert (d.nx, d.ny, d.nz) == (case["nx"], case["ny"], case["nz"])
    assert (d.is_, d.js, d.ks, d.ie, d.je, d.ke) == case["ext"]
    assert d.inne == case["inne"] and d.inne2 == case["inne2"]
    assert (d.t0, d.g, d.nu) == (case["t0"], case["g"], case["nu"])
    assert np.array_equal(g.solid.ravel(), case["solid"])
    assert np.array_equal(g.un.reshape(3, -1), case["U"])
    assert np.array_equal(g.F.ravel(), case["F"])
    assert np.array_equal(p.ravel(), case["P"])
    if case["fdata"] is not None:
        assert np.array_equal(g.a.reshape(3, -1), case["A"])
        assert np.array_equal(g.fb.ravel(), case["FB"])
    for l, n in enumerate((case["nx"], case["ny"], case["nz"])):
        # x[l, c] = X(l, ls-1+c) for the box cells c = 0 .. n+1
        assert np.array_equal(g.xf[l, :n + 2], case["coords"][l][1:n + 3])


@pytest.mark.parametrize("name,sparse", CASES)
@pytest.mark.parametrize("be", [False, True])
def test_reader_matches_fixture(tmp_path, name, sparse, be):
    from clowns3d import set_dx
    case = make_case(tmp_path, 7, 5, 4, seed=3 + sparse, sparse=sparse, be=be)
    g, p, d = ours(case)
    _check_fields(case, g, p, d)
    assert d.big_endian == int(be) and d.sparse == int(sparse)
    assert d.has_dt0 == 1 and d.dt0 == case["dt0"] and g.dt == case["dt0"]
    for l, n in enumerate((case["nx"], case["ny"], case["nz"])):
        _, dx = set_dx(case["coords"][l])
        assert np.array_equal(g.dx[l, :n + 2], dx[1:n + 3])


@pytest.mark.parametrize("name,sparse", CASES)
def test_reader_without_fdata_defaults(tmp_path, name, sparse):
    case = make_case(tmp_path, 5, 4, 3, seed=11, sparse=sparse, fdata=False)
    g, p, d = ours(case, fdata=None)
    open_ = case["solid"] == 0
    assert np.array_equal(g.a.reshape(3, -1), np.where(open_, 1.0, 0.0)[None].repeat(3, 0))
    assert np.array_equal(g.fb.ravel(), np.where(open_, 1.0, 0.0))


# The first cell indices stay >= 3: the reference allocates its coordinate
# arrays from index 1 (alloc_dxdydz, src2LOWNS/0mod-arrays.F90:200-206) and
# reads the faces from ls-2 (data_io.F90:113-116), so a case starting below 3
# writes outside them (found with an AddressSanitizer build of the reader).
@pytest.mark.parametrize("name,sparse", CASES)
@pytest.mark.parametrize("shape,lo,objects", [((7, 5, 4), (3, 3, 3), False), ((12, 9, 6), (5, 3, 4), False),
                                              ((9, 7, 5), (3, 4, 5), True)])
@pytest.mark.parametrize("be", [False, True], ids=["le", "be"])
def test_reader_matches_reference(tmp_path, name, sparse, shape, lo, objects, be):
    case = make_case(tmp_path, *shape, seed=sum(shape) + sparse, sparse=sparse, lo=lo, objects=objects, be=be)
    _check_fields(case, *ours(case))
    r = ref_read_be(case) if be else ref_read(case)
    g, p, d = ours(case)
    assert tuple(r["ext"]) == case["ext"]
    assert r["inne"] == d.inne and r["inne2"] == d.inne2
    assert (d.inne2 > d.inne) == objects
    assert np.array_equal(r["box"], case["cells"])  # same numbering, same order
    assert tuple(r["hdr"]) == (d.t0, d.g, d.nu, d.dt0)
    c = r["box"]
    assert np.array_equal(r["U"], g.un.reshape(3, -1)[:, c])
    assert np.array_equal(r["F"], g.F.ravel()[c])
    assert np.array_equal(r["P"], p.ravel()[c])
    assert np.array_equal(r["A"], g.a.reshape(3, -1)[:, c])
    assert np.array_equal(r["FB"], g.fb.ravel()[c])
    assert np.array_equal(r["flags"][:, 0], case["nf"][c])
    assert np.array_equal(r["flags"][:, 1], case["nb"][c])
    for l, n in enumerate(shape):
        assert np.array_equal(r["xo"][l, 1:n + 3], g.xf[l, :n + 2])
    # every cell the reference does not number is outside the case in the box
    # model (1, or -127 for its second sentinel); numbered object cells carry
    # their codes
    rest = np.setdiff1d(np.arange(case["nc"]), c)
    sol = g.solid.ravel()
    assert np.all((sol[rest] == 1) | (sol[rest] == -127))
    assert np.all(sol[c] != 1) and np.all((sol[c] == 0) == (r["flags"][:, 0] >= 0))


def test_dense_restart_time_zero(tmp_path):
    """t0 = 0: the fields are still read, dt0 is not (data_io.F90:283-291)."""
    case = make_case(tmp_path, 6, 5, 3, seed=2, sparse=False, t0=0.0)
    g, p, d = ours(case)
    _check_fields(case, g, p, d)
    assert d.has_dt0 == 0 and d.dt0 == 0.0
    r = ref_read(case)
    c = r["box"]
    assert r["hdr"][3] == 0.0
    assert np.array_equal(r["P"], p.ravel()[c]) and np.array_equal(r["U"], g.un.reshape(3, -1)[:, c])


def writer_flags(case):
    """nf/nb as the classification gives them: the reference's flags, with
    each nf = -1 cell's object code as its nb."""
    nf = np.where(case["solid"] != 0, -1, case["nf"]).astype(np.int8)
    nb = np.where(case["solid"] != 0, np.where(case["solid"] == 1, 0, case["solid"]), case["nb"]).astype(np.int8)
    return nf, nb


@pytest.mark.parametrize("name,sparse", CASES)
@pytest.mark.parametrize("objects", [False, True])
@pytest.mark.parametrize("be", [False, True], ids=["le", "be"])
def test_writer_byte_identical_to_reference(tmp_path, name, sparse, objects, be):
    from clowns3d import decks
    case = make_case(tmp_path, 8, 6, 5, seed=21 + sparse, sparse=sparse, objects=objects, be=be)
    g, p, d = ours(case)
    nf, nb = writer_flags(case)
    rng = np.random.default_rng(4)
    # a new (solved) pressure on the open cells
    p2 = p.ravel() + rng.standard_normal(case["nc"]) * np.where(case["solid"] == 0, 1.0, 0.0)
    t, dt = 37.25, 0.00125
    # put the new pressure into the reference's state through a data file
    # it reads: ours, written once, read by the reference, which then writes
    # it out again with output_data
    mine = tmp_path / "mine.data"
    theirs = tmp_path / "theirs.data"
    decks.write_data(mine, d, nf, nb, g.un.reshape(3, -1), g.F.ravel(), p2, t, dt, big_endian=be)
    if be:
        r = ref_read_be(case, data=mine, write=(theirs, t, dt))
    else:
        ref_read(case)
        r = ref_read(case, data=mine)
        ref_write(theirs, t, dt)
    c = r["box"]
    U = g.un.reshape(3, -1)
    expect = np.where(np.arange(len(c)) < d.inne, p2[c], 0.0)  # appended cells carry no P
    if sparse:  # cells without state are not listed and read back as zero
        dropped = (nf[c] == 0) & ((U[:, c] ** 2).sum(0) == 0) & (g.F.ravel()[c] < 1e-13)
        assert dropped.any() and not dropped.all()
        expect = np.where(dropped, 0.0, expect)
    assert np.array_equal(r["P"], expect)
    assert np.array_equal(r["U"], g.un.reshape(3, -1)[:, r["box"]])
    raw = mine.read_bytes()
    assert raw[:4] == (24).to_bytes(4, "big" if be else "little")
    assert raw == theirs.read_bytes()


@pytest.mark.parametrize("name,sparse", CASES)
def test_writer_round_trip_big_endian(tmp_path, name, sparse):
    from clowns3d import decks
    case = make_case(tmp_path, 6, 7, 4, seed=31 + sparse, sparse=sparse, be=True)
    g, p, d = ours(case)
    out = tmp_path / "out.data"
    decks.write_data(out, d, case["nf"], case["nb"], g.un.reshape(3, -1), g.F, p * 2.0, 5.0, 0.5, big_endian=True)
    raw = out.read_bytes()
    assert raw[:4] == (24).to_bytes(4, "big")
    g2, p2, d2 = ours(case, data=out)
    assert d2.t0 == 5.0 and d2.dt0 == 0.5 and d2.big_endian == 1
    assert np.array_equal(p2, p * 2.0)
    assert np.array_equal(g2.un, g.un) and np.array_equal(g2.F, g.F)


def test_sparse_writer_drops_stateless_cells(tmp_path):
    """The .xyzn data form lists only cells with nf != 0, U != 0 or F >=
    1e-13 (data_io.F90:745-759)."""
    from clowns3d import decks
    case = make_case(tmp_path, 5, 4, 3, seed=8, sparse=True)
    g, p, d = ours(case)
    nc = case["nc"]
    un = np.zeros((3, nc))
    F = np.zeros(nc)
    nf = np.where(case["solid"] == 1, -1, 0).astype(np.int8)
    nb = np.zeros(nc, np.int8)
    c = case["cells"]
    F[c[0]] = 1e-13            # kept: F >= 1e-13
    F[c[1]] = 0.99e-13         # dropped
    un[1, c[2]] = 1e-100       # kept: |U|^2 != 0
    nf[c[3]] = 2               # kept: nf != 0
    un[0, c[4]] = 1e-300       # dropped: the rule tests |U|^2, which underflows to 0
    out = tmp_path / "out.data"
    decks.write_data(out, d, nf, nb, un, F, np.arange(nc, dtype=float), 1.0, 0.1, big_endian=False)
    raw = out.read_bytes()
    icc = int.from_bytes(raw[32 + 4:32 + 8], "little")
    assert icc == 3
    g2, p2, _ = ours(case, data=out)
    kept = [c[0], c[2], c[3]]
    assert np.array_equal(p2.ravel()[kept], np.array(kept, float))
    assert p2.ravel()[c[1]] == 0.0 and p2.ravel()[c[4]] == 0.0


def reference_grid(paths, like):
    """The box grid the reference itself makes of case files: the flang
    build of read_data + read_fdata places each numbered cell's values at
    the box position of its own (i, j, k); a cell's code comes from the
    reference's flags (flow cell 0; boundary object its nfb's code; appended
    cell -1 or -128 from the nfb read_data resets it to); every cell the
    reference does not number is outside the case (1); faces and dx come
    from the flang build of set_dx (src2LOWNS/1sub-main.F90:53-74) on the
    faces the reference read.  `like` gives the box size.  Returns (grid, p)."""
    from clowns3d import configs
    nx, ny, nz = like.nx, like.ny, like.nz
    r = ref_read(dict(grid=paths[0], data=paths[1], fdata=paths[2], nc=like.ncell, nx=nx, ny=ny, nz=nz,
                      X=like.X, Y=like.Y))
    g = configs.Grid(nx, ny, nz)
    nc = g.ncell
    sol = np.ones(nc, np.int8)
    F, P, fb = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    un, a = np.zeros((3, nc)), np.zeros((3, nc))
    for q in range(r["inne2"]):
        c = r["box"][q]
        f, b = (int(v) for v in r["flags"][q])
        sol[c] = 0 if f >= 0 else (object_code(b) if q < r["inne"] else (-1 if b == -1 else -128))
        un[:, c], F[c], P[c], a[:, c], fb[c] = r["U"][:, q], r["F"][q], r["P"][q], r["A"][:, q], r["FB"][q]
    shape = (g.Z, g.Y, g.X)
    g.solid[:] = sol.reshape(shape)
    g.F[:], g.fb[:] = F.reshape(shape), fb.reshape(shape)
    g.un[:], g.a[:] = un.reshape((3,) + shape), a.reshape((3,) + shape)
    if not os.path.exists(REF_ASM_SO):
        pytest.skip("flang build of the reference (oracle/_ref) not present")
    lib = ctypes.CDLL(REF_ASM_SO, mode=os.RTLD_LAZY)
    g.xf[:] = 0.0
    g.dx[:] = 0.0
    for l, n in enumerate((nx, ny, nz)):
        x = np.ascontiguousarray(r["xo"][l, :n + 4])
        dx = np.zeros_like(x)
        lib.ref_set_dx(len(x), ptr(x), ptr(dx))
        m = min(n + 3, g.stride)  # box cells 0 .. n+1, then the last face where it fits
        g.xf[l, :m], g.dx[l, :m] = x[1:m + 1], dx[1:m + 1]
    g.h = tuple(float(g.dx[l, 1]) for l in range(3))
    if r["hdr"][3] > 0:
        g.dt = float(r["hdr"][3])
    return g, P.reshape(shape)


def assert_same_case(g, p, gr, pr):
    """Every field the assembly and the pressure loop read, bit for bit."""
    assert np.array_equal(g.solid, gr.solid)
    for name in ("F", "fb", "a", "un", "xf", "dx"):
        assert np.array_equal(getattr(g, name), getattr(gr, name)), name
    assert np.array_equal(p, pr)
    assert g.h == gr.h and g.dt == gr.dt


@pytest.mark.parametrize("name,sparse", CASES)
def test_object_grid_files_match_reference(tmp_path, name, sparse):
    """A grid with object cells of every kind (configs.config_objects:
    pressure-boundary cells, inflow boundary objects, appended obstacle
    cells) written to case files: the product's reader and the reference's
    reader + set_dx give the same box grid, pressure and spacing."""
    from clowns3d import assembly, configs, decks
    g0 = configs.config_objects(40, 24, 16, seed=11)
    c = assembly.classify(g0)
    p0 = np.where(g0.solid.ravel() == 1, 0.0, np.random.default_rng(7).standard_normal(g0.ncell) * 50)
    paths = write_grid_case(tmp_path, g0, c["nf"], c["nb"], p0, sparse=sparse)
    g1, p1, d = decks.read_case(*paths)
    assert d.inne2 > d.inne
    gr, pr = reference_grid(paths, g1)
    assert_same_case(g1, p1, gr, pr)
    codes = set(np.unique(gr.solid).tolist())
    assert {4, 5, 7, -2, -30, -1, -128} <= codes


def test_object_codes(tmp_path):
    """Object cells read into their codes: boundary objects (nfb 4, 5, 7 kept,
    1 -> -2, 150 -> -26), appended cells (300 -> -1, 200 -> -128) with U and a
    but no F, P or fb, nfb = -1 cells -> -127 with zero fields."""
    for sparse in (False, True):
        sub = tmp_path / str(sparse)
        sub.mkdir()
        case = make_case(sub, 6, 5, 4, seed=9, sparse=sparse, objects=True)
        g, p, d = ours(case)
        sol = g.solid.ravel()
        assert set(np.unique(sol).tolist()) >= {0, 1, 4, 5, 7, -2, -26, -1, -128, -127}
        app = case["cells"][d.inne:]
        assert len(app) > 0 and set(np.unique(sol[app]).tolist()) <= {-1, -128}
        assert np.all(g.F.ravel()[app] == 0) and np.all(p.ravel()[app] == 0) and np.all(g.fb.ravel()[app] == 0)
        assert np.abs(g.un.reshape(3, -1)[:, app]).sum() > 0
        s127 = sol == -127
        assert s127.any() and np.all(g.fb.ravel()[s127] == 0) and np.all(g.a.reshape(3, -1)[:, s127] == 0)


def test_truncated_and_missing_files(tmp_path):
    from clowns3d import decks
    from clowns3d._lib import CLOWNS_EDECK, ClownsError
    case = make_case(tmp_path, 5, 4, 3, seed=10, sparse=False)
    raw = case["data"].read_bytes()
    cut = tmp_path / "cut.data"
    cut.write_bytes(raw[:len(raw) // 2])
    for args in ((case["grid"], cut, case["fdata"]), (tmp_path / "none.xyz", case["data"], None),
                 (case["grid"], case["data"], tmp_path / "none.fdata0")):
        with pytest.raises(ClownsError) as e:
            decks.read_case(*args)
        assert e.value.code == CLOWNS_EDECK
    short = tmp_path / "short.xyz"
    shutil.copy(case["grid"], short)
    with open(short, "r+b") as f:
        f.truncate(40)
    with pytest.raises(ClownsError):
        decks.read_case(short, case["data"], None)


@pytest.mark.parametrize("sparse", [False, True])
def test_grid_through_files(tmp_path, sparse):
    """A configs.Grid written as case files reads back as the same Grid:
    every array the classification and the assembly take, with dx from
    set_dx of the faces equal to the uniform cell sizes."""
    from clowns3d import assembly
    from _deckfiles import write_grid_case
    from test_oracle_assembly import GRIDS
    g0 = GRIDS["wavy32x24x16"]()
    c = assembly.classify(g0)
    p0 = np.random.default_rng(3).standard_normal(g0.ncell) * (g0.solid.ravel() == 0)
    paths = write_grid_case(tmp_path, g0, c["nf"], c["nb"], p0, sparse=sparse, be=True)
    g1, p1, d = decks_read(*paths)
    for k in ("solid", "F", "a", "fb"):
        assert np.array_equal(getattr(g1, k), getattr(g0, k)), k
    # velocities live on numbered cells only; solid ones read back as zero
    open_ = g0.solid == 0
    assert np.array_equal(g1.un, np.where(open_, g0.un, 0.0))
    for l, n in enumerate((g0.nx, g0.ny, g0.nz)):
        assert np.array_equal(g1.xf[l, :n + 2], g0.xf[l, :n + 2])
        assert np.array_equal(g1.dx[l, :n + 2], g0.dx[l, :n + 2])
    assert g1.dt == g0.dt and g1.h == g0.h
    assert np.array_equal(p1.ravel(), p0)
    c1 = assembly.classify(g1)
    for k in ("nf", "nb", "mno", "row2cell", "sp_row"):
        assert np.array_equal(c1[k], c[k]), k


def decks_read(grid, data, fdata):
    from clowns3d import decks
    return decks.read_case(grid, data, fdata)


@pytest.mark.parametrize("be", [False, True])
def test_grid_writer_reproduces_sparse_grid(tmp_path, be):
    """Read a .xyzn grid, write it back: the same bytes."""
    from clowns3d import decks
    case = make_case(tmp_path, 9, 6, 5, seed=41, sparse=True, be=be)
    g, p, d = ours(case)
    fa = decks.faces(case["grid"], d)
    for l, n in enumerate((9, 6, 5)):
        assert np.array_equal(fa[l, :n + 4], case["coords"][l]) and not fa[l, n + 4:].any()
    out = tmp_path / "out.xyzn"
    decks.write_grid(out, d, g.solid, fa, big_endian=be)
    assert out.read_bytes() == case["grid"].read_bytes()


def test_grid_writer_matches_reference(tmp_path):
    """A dense grid read at t0 = 0 makes the reference write its .xyzn form
    (data_io.F90:137-165, output_xyzn's format); ours from the same case is
    byte-identical, and reads back as the same box."""
    from clowns3d import decks
    case = make_case(tmp_path, 8, 7, 4, seed=42, sparse=False, t0=0.0)
    ref_read(case)
    theirs = tmp_path / "case.data.xyzn"
    assert theirs.exists()
    g, p, d = ours(case)
    mine = tmp_path / "mine.xyzn"
    decks.write_grid(mine, d, g.solid, decks.faces(case["grid"], d), big_endian=False)
    assert mine.read_bytes() == theirs.read_bytes()
    d2 = decks.DeckDesc()
    from clowns3d import _lib as L
    L.check(L.lib().clowns_deck_open(str(mine).encode(), str(case["data"]).encode(), ctypes.byref(d2)), "open")
    assert d2.sparse == 1 and d2.inne == d.inne


def test_corrupt_record_length_and_bad_descriptor(tmp_path):
    """A record length larger than the file is an unreadable file, not an
    allocation; a descriptor whose extents disagree with its sizes is refused
    before any array is touched."""
    import struct
    from clowns3d import decks
    from clowns3d._lib import CLOWNS_EBADARG, CLOWNS_EDECK, ClownsError
    case = make_case(tmp_path, 5, 4, 3, seed=12, sparse=False)
    bad = tmp_path / "bad.data"
    raw = bytearray(case["data"].read_bytes())
    raw[32:36] = struct.pack("<i", 0x7FFFFFF0)  # the nf record's length
    bad.write_bytes(bytes(raw))
    with pytest.raises(ClownsError) as e:
        decks.read_case(case["grid"], bad, None)
    assert e.value.code == CLOWNS_EDECK
    g, p, d = ours(case)
    d.stride = 3  # too small for nx = 5
    with pytest.raises(ClownsError) as e:
        decks.write_data(tmp_path / "x.data", d, case["nf"], case["nb"], g.un.reshape(3, -1), g.F, p, 1.0, 0.1)
    assert e.value.code == CLOWNS_EBADARG
