// decks.cpp -- the reference's on-disk formats for the pressure path: read a
// 2CLOWNS-3D case (grid file, data file, open-ratio file) into the box layout
// the assembly and the solver take, and write the data file back with the
// solved pressure.  Host code, start-up and output only.
//
//   grid file   read_data, src2LOWNS/data_io.F90:98-128: extents, face
//               coordinates per axis, and (".xyzn") the list of cells with
//               nf = -1 (output_xyzn, data_io.F90:1531-1560)
//   data file   read_data, data_io.F90:130-156 (dense nf/nfb) and :219-300
//               (fields, optional dt0); output_data, data_io.F90:708-783
//   .xyzn grid  output_xyzn, data_io.F90:1505-1560 (and its copy at
//               read_data :137-165 for a dense grid at t0 = 0)
//   open ratios read_fdata, data_io.F90:304-364 (a0, fb0)
//
// The files are Fortran unformatted sequential: each record is a 4-byte
// length, the payload, and the length again.  The reference build writes
// them big-endian (-convert big_endian, src2LOWNS/Makefile:5); the reader
// takes either byte order (from the first record's length), the writer
// writes the one asked for.  Default INTEGER is 4 bytes and REAL 8 (-r8).
//
// Cells with nf = -1 (data_io.F90:168-199): nfb = 0 cells are not numbered
// (mn = 0: the box model's solid value 1); boundary objects (1 <= nfb <
// 200) are numbered with the flow cells and appended objects (nfb >= 200)
// after them, and become solid cells with their object code (see
// object_code); nfb = -1 cells map to the reference's second sentinel cell
// inne2+1 (flag -1, b = -1; code -127, read with zero fields like a cell
// outside the case).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "clowns_pcg.h"

namespace {

uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
uint64_t bswap64(uint64_t v) {
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

// One Fortran unformatted sequential file, read record by record.
class RecordReader {
  public:
    explicit RecordReader(const char *path) : f_(path ? std::fopen(path, "rb") : nullptr) {
        if (f_) {
            std::fseek(f_, 0, SEEK_END);
            size_ = std::ftell(f_);
            std::fseek(f_, 0, SEEK_SET);
        }
    }
    ~RecordReader() {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr; }
    // Reads the next record; on the first one the byte order is set from the
    // length marker (the one of the two readings that fits the file).
    bool next() {
        uint32_t m;
        const long pos = std::ftell(f_);
        if (std::fread(&m, 4, 1, f_) != 1) return false;
        if (!order_set_) {
            const uint32_t sw = bswap32(m);
            swap_ = !((long)m + 8 <= size_) && ((long)sw + 8 <= size_);
            order_set_ = true;
        }
        const uint32_t n = swap_ ? bswap32(m) : m;
        if (pos < 0 || (long)n + 8 > size_ - pos) return false;  // a length the file cannot hold
        rec_.resize(n);
        if (n && std::fread(rec_.data(), 1, n, f_) != n) return false;
        uint32_t t;
        if (std::fread(&t, 4, 1, f_) != 1) return false;
        if ((swap_ ? bswap32(t) : t) != n) return false;
        pos_ = 0;
        return true;
    }
    size_t size() const { return rec_.size(); }
    bool big_endian() const { return swap_; }  // the host is little-endian
    bool i32(int32_t &v) {
        if (pos_ + 4 > rec_.size()) return false;
        uint32_t u;
        std::memcpy(&u, rec_.data() + pos_, 4);
        pos_ += 4;
        u = swap_ ? bswap32(u) : u;
        std::memcpy(&v, &u, 4);
        return true;
    }
    bool f64(double &v) {
        if (pos_ + 8 > rec_.size()) return false;
        uint64_t u;
        std::memcpy(&u, rec_.data() + pos_, 8);
        pos_ += 8;
        u = swap_ ? bswap64(u) : u;
        std::memcpy(&v, &u, 8);
        return true;
    }

  private:
    FILE *f_;
    long size_ = 0;
    std::vector<char> rec_;
    size_t pos_ = 0;
    bool swap_ = false, order_set_ = false;
};

// The descriptor a read or write relies on for the caller's array sizes.
bool desc_ok(const clowns_deck_desc *d) {
    return d->nx >= 1 && d->ny >= 1 && d->nz >= 1 && d->ie - d->is == d->nx && d->je - d->js == d->ny &&
           d->ke - d->ks == d->nz && d->stride >= std::max(d->nx, std::max(d->ny, d->nz)) + 2;
}

class RecordWriter {
  public:
    RecordWriter(const char *path, bool big_endian)
        : f_(path ? std::fopen(path, "wb") : nullptr), swap_(big_endian) {}
    ~RecordWriter() {
        if (f_) std::fclose(f_);
    }
    bool ok() const { return f_ != nullptr && good_; }
    void i32(int32_t v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u = swap_ ? bswap32(u) : u;
        put(&u, 4);
    }
    void f64(double v) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        u = swap_ ? bswap64(u) : u;
        put(&u, 8);
    }
    void end() {  // closes the current record
        uint32_t n = (uint32_t)buf_.size();
        uint32_t m = swap_ ? bswap32(n) : n;
        good_ &= std::fwrite(&m, 4, 1, f_) == 1;
        if (n) good_ &= std::fwrite(buf_.data(), 1, n, f_) == n;
        good_ &= std::fwrite(&m, 4, 1, f_) == 1;
        buf_.clear();
    }

  private:
    void put(const void *p, size_t n) {
        const char *c = (const char *)p;
        buf_.insert(buf_.end(), c, c + n);
    }
    FILE *f_;
    bool swap_;
    bool good_ = true;
    std::vector<char> buf_;
};

// The object code of an nf = -1 cell from its nfb (the box grids' solid
// value; 1: a cell the case does not number).  A boundary object with
// nfb = b (1 <= b < 200) keeps b for 2..126, takes -2 for b = 1 and 124 - b
// (-3..-75) for b >= 127; appended cells (nfb >= 200) take -1 when
// nfb = 300 (their b becomes -1) and -128 otherwise (b = 0); nfb = -1 cells
// take -127.
int8_t object_code(int32_t nfb) {
    if (nfb == -1) return -127;
    if (nfb <= 0) return 1;
    if (nfb >= 200) return nfb == 300 ? -1 : -128;
    if (nfb == 1) return -2;
    if (nfb >= 127) return (int8_t)(124 - nfb);
    return (int8_t)nfb;
}
// The nfb a file holds for an object code (its inverse).
int32_t file_nfb(int8_t code) {
    if (code == 1 || code == 0) return 0;
    if (code == -127) return -1;
    if (code == -1) return 300;
    if (code == -128) return 200;
    if (code == -2) return 1;
    if (code < 0) return 124 - code;
    return code;
}
// nff%b the reference keeps for a cell of that code after read_data
// (appended cells' nfb is reset, data_io.F90:190-194).
int32_t kept_b(int8_t code) {
    if (code == -1) return -1;
    if (code == -128) return 0;
    return file_nfb(code);
}
bool in_main(int32_t nf, int32_t nfb) { return nf >= 0 || (nfb >= 1 && nfb < 200); }
bool appended(int32_t nf, int32_t nfb) { return nf < 0 && nfb >= 200; }

// The cell numbering of read_data (data_io.F90:168-199): j outer, k, i
// inner over flow cells and boundary objects (1..inne), then the same sweep
// over appended object cells (inne+1..inne2).  Returns the numbered box
// cells in order; *inne the first part's count.
std::vector<long> numbering(const clowns_deck_desc *d, const std::vector<int32_t> &nf,
                            const std::vector<int32_t> &nfb, long *inne) {
    const long X = d->nx + 2, Y = d->ny + 2, Z = d->nz + 2;
    std::vector<long> cells;
    for (int pass = 0; pass < 2; ++pass) {
        for (long j = 0; j < Y; ++j)
            for (long k = 0; k < Z; ++k)
                for (long i = 0; i < X; ++i) {
                    const long c = i + X * (j + Y * k);
                    if (pass == 0 ? in_main(nf[c], nfb[c]) : appended(nf[c], nfb[c])) cells.push_back(c);
                }
        if (pass == 0) *inne = (long)cells.size();
    }
    return cells;
}

bool ends_with_z(const char *path) {
    const size_t n = std::strlen(path);
    return n > 0 && path[n - 1] == 'z';
}

// nf and nfb of every box cell from the grid file (sparse) or the data file
// (dense); the data reader is left after the nf/nfb records.
int read_flags(const clowns_deck_desc *d, RecordReader &grid, RecordReader &data, std::vector<int32_t> &nf,
               std::vector<int32_t> &nfb) {
    const long X = d->nx + 2, Y = d->ny + 2, Z = d->nz + 2, nc = X * Y * Z;
    nf.assign(nc, 0);
    nfb.assign(nc, 0);
    auto cell = [&](int32_t i, int32_t k, int32_t j) -> long {
        const long ii = i - (d->is - 1), jj = j - (d->js - 1), kk = k - (d->ks - 1);
        if (ii < 0 || ii >= X || jj < 0 || jj >= Y || kk < 0 || kk >= Z) return -1;
        return ii + X * (jj + Y * kk);
    };
    if (d->sparse) {
        // grid file (the header record already read): three coordinate
        // records, icc, icc cell records
        for (int r = 0; r < 3; ++r)
            if (!grid.next()) return CLOWNS_EDECK;
        int32_t icc;
        if (!grid.next() || !grid.i32(icc) || icc < 0) return CLOWNS_EDECK;
        for (int32_t n = 0; n < icc; ++n) {
            int32_t i, k, j, a, b;
            if (!grid.next() || !grid.i32(i) || !grid.i32(k) || !grid.i32(j) || !grid.i32(a) || !grid.i32(b))
                return CLOWNS_EDECK;
            const long c = cell(i, k, j);
            if (c < 0) return CLOWNS_EDECK;
            nf[c] = a;
            nfb[c] = b;
        }
        return CLOWNS_OK;
    }
    // data file: (((NF3(I,k,j),K=ks-1,KE),J=js-1,je),I=is-1,IE) then NFB3
    for (int which = 0; which < 2; ++which) {
        std::vector<int32_t> &dst = which == 0 ? nf : nfb;
        if (!data.next() || (long)data.size() != 4 * nc) return CLOWNS_EDECK;
        for (long i = 0; i < X; ++i)
            for (long j = 0; j < Y; ++j)
                for (long k = 0; k < Z; ++k) {
                    int32_t v;
                    if (!data.i32(v)) return CLOWNS_EDECK;
                    dst[i + X * (j + Y * k)] = v;
                }
    }
    return CLOWNS_OK;
}

} // namespace

static int clowns_deck_open_impl(const char *grid_path, const char *data_path, clowns_deck_desc *d) {
    if (!grid_path || !data_path || !d) return CLOWNS_EBADARG;
    std::memset(d, 0, sizeof(*d));
    RecordReader grid(grid_path);
    if (!grid.ok() || !grid.next()) return CLOWNS_EDECK;
    int32_t e[6];
    for (int k = 0; k < 6; ++k)
        if (!grid.i32(e[k])) return CLOWNS_EDECK;
    d->is = e[0];
    d->js = e[1];
    d->ks = e[2];
    d->ie = e[3];
    d->je = e[4];
    d->ke = e[5];
    d->nx = d->ie - d->is;
    d->ny = d->je - d->js;
    d->nz = d->ke - d->ks;
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return CLOWNS_EDECK;
    d->stride = std::max(d->nx, std::max(d->ny, d->nz)) + 2;
    d->sparse = ends_with_z(grid_path) ? 0 : 1;  // read_data, data_io.F90:100-105
    d->big_endian = grid.big_endian() ? 1 : 0;
    RecordReader data(data_path);
    double h[3];
    if (!data.ok() || !data.next() || !data.f64(h[0]) || !data.f64(h[1]) || !data.f64(h[2])) return CLOWNS_EDECK;
    d->t0 = h[0];
    d->g = h[1];
    d->nu = h[2];
    std::vector<int32_t> nf, nfb;
    int rc = read_flags(d, grid, data, nf, nfb);
    if (rc) return rc;
    long inne = 0;
    d->inne2 = (int32_t)numbering(d, nf, nfb, &inne).size();
    d->inne = (int32_t)inne;
    return CLOWNS_OK;
}

static int clowns_deck_read_impl(const char *grid_path, const char *data_path, const char *fdata_path,
                                clowns_deck_desc *d, int8_t *solid, double *F, double *un, double *p,
                                double *a, double *fb, double *x, double *dx) {
    if (!grid_path || !data_path || !d || !solid || !F || !un || !p || !a || !fb || !x || !dx)
        return CLOWNS_EBADARG;
    if (!desc_ok(d)) return CLOWNS_EBADARG;
    const long X = d->nx + 2, Y = d->ny + 2, Z = d->nz + 2, nc = X * Y * Z;
    const int S = d->stride;
    RecordReader grid(grid_path), data(data_path);
    if (!grid.ok() || !data.ok() || !grid.next() || !data.next()) return CLOWNS_EDECK;
    // coordinates X(l, ls-2 .. le+1) (data_io.F90:113-116) and dx by set_dx
    // on that range; the box cell c is the reference's cell ls-1+c, so
    // x[l*S + c] = X(l, ls-1+c) and dx[l*S + c] = dx(l, ls-1+c), c = 0 .. n+1
    const int lo[3] = {d->is, d->js, d->ks}, hi[3] = {d->ie, d->je, d->ke};
    for (int l = 0; l < 3; ++l) {
        const int m = hi[l] - lo[l] + 4;
        std::vector<double> xs(m), dxs(m);
        if (!grid.next() || (long)grid.size() != 8L * m) return CLOWNS_EDECK;
        for (int q = 0; q < m; ++q)
            if (!grid.f64(xs[q])) return CLOWNS_EDECK;
        int rc = clowns_set_dx(xs.data(), m, dxs.data());
        if (rc) return rc;
        for (int c = 0; c < S; ++c) {
            x[l * S + c] = c + 1 < m ? xs[c + 1] : 0.0;
            dx[l * S + c] = c + 1 < m ? dxs[c + 1] : 0.0;
        }
    }
    // restart the grid reader for read_flags (it skips the four records)
    RecordReader grid2(grid_path);
    if (!grid2.next()) return CLOWNS_EDECK;
    std::vector<int32_t> nf, nfb;
    int rc = read_flags(d, grid2, data, nf, nfb);
    if (rc) return rc;
    long inne = 0;
    const std::vector<long> cells = numbering(d, nf, nfb, &inne);
    const long inne2 = (long)cells.size();
    // without an open-ratio file the cells 1..inne are fully open and the
    // appended ones closed (read_fdata, data_io.F90:359-362 and :314-316)
    for (long c = 0; c < nc; ++c) {
        solid[c] = nf[c] < 0 ? object_code(nfb[c]) : 0;
        const bool open = in_main(nf[c], nfb[c]);
        F[c] = 0.0;
        p[c] = 0.0;
        fb[c] = open ? 1.0 : 0.0;
        for (int l = 0; l < 3; ++l) {
            un[l * nc + c] = 0.0;
            a[l * nc + c] = open ? 1.0 : 0.0;
        }
    }
    // U of every numbered cell; F and P of the cells 1..inne
    auto put = [&](long nn, const double *u, const double *f, const double *pp) -> bool {
        if (nn < 1 || nn > inne2 || (nn > inne && (f || pp))) return false;
        const long c = cells[nn - 1];
        for (int l = 0; l < 3; ++l) un[l * nc + c] = u[l];
        if (f) F[c] = *f;
        if (pp) p[c] = *pp;
        return true;
    };
    // fields (data_io.F90:219-253); t0 = 0 with a dense grid still carries them
    if (!d->sparse) {
        for (long nn = 1; nn <= inne2; ++nn) {
            double u[3], f, pp;
            if (nn <= inne) {
                if (!data.next() || !data.f64(u[0]) || !data.f64(u[1]) || !data.f64(u[2]) || !data.f64(f) ||
                    !data.f64(pp) || !put(nn, u, &f, &pp))
                    return CLOWNS_EDECK;
            } else if (!data.next() || !data.f64(u[0]) || !data.f64(u[1]) || !data.f64(u[2]) ||
                       !put(nn, u, nullptr, nullptr)) {
                return CLOWNS_EDECK;
            }
        }
    } else {
        int32_t icc, icc2;
        if (!data.next() || !data.i32(icc) || !data.i32(icc2) || icc < 0 || icc2 < 0) return CLOWNS_EDECK;
        for (int32_t q = 0; q < icc; ++q) {
            int32_t nn, a1, b1;
            double u[3], f, pp;
            if (!data.next() || !data.i32(nn) || !data.f64(u[0]) || !data.f64(u[1]) || !data.f64(u[2]) ||
                !data.f64(f) || !data.f64(pp) || !data.i32(a1) || !data.i32(b1) || nn > inne ||
                !put(nn, u, &f, &pp))
                return CLOWNS_EDECK;
        }
        for (int32_t q = 0; q < icc2; ++q) {
            int32_t nn;
            double u[3];
            if (!data.next() || !data.i32(nn) || !data.f64(u[0]) || !data.f64(u[1]) || !data.f64(u[2]) ||
                !put(nn, u, nullptr, nullptr))
                return CLOWNS_EDECK;
        }
    }
    d->has_dt0 = 0;
    d->dt0 = 0.0;
    // optional, and only after a restart time t0 > 0 (data_io.F90:283-287)
    if (d->t0 > 0.0 && data.next() && data.f64(d->dt0)) d->has_dt0 = 1;
    // open ratios (read_fdata, data_io.F90:310-330): a0(0:2,nn), fb0(nn)
    if (fdata_path && fdata_path[0]) {
        RecordReader fd(fdata_path);
        if (!fd.ok()) return CLOWNS_EDECK;
        // a0 of every numbered cell, fb0 of the cells 1..inne
        auto put_a = [&](long nn, const double *av, const double *fbv) -> bool {
            if (nn < 1 || nn > inne2 || (nn > inne && fbv)) return false;
            const long c = cells[nn - 1];
            for (int l = 0; l < 3; ++l) a[l * nc + c] = av[l];
            if (fbv) fb[c] = *fbv;
            return true;
        };
        if (d->sparse) {
            int32_t icc, icc1;
            if (!fd.next() || !fd.i32(icc) || !fd.i32(icc1) || icc < 0 || icc1 < 0) return CLOWNS_EDECK;
            for (int32_t q = 0; q < icc; ++q) {
                int32_t nn;
                double av[3], fbv;
                if (!fd.next() || !fd.i32(nn) || !fd.f64(av[0]) || !fd.f64(av[1]) || !fd.f64(av[2]) ||
                    !fd.f64(fbv) || nn > inne || !put_a(nn, av, &fbv))
                    return CLOWNS_EDECK;
            }
            for (int32_t q = 0; q < icc1; ++q) {
                int32_t nn;
                double av[3];
                if (!fd.next() || !fd.i32(nn) || !fd.f64(av[0]) || !fd.f64(av[1]) || !fd.f64(av[2]) ||
                    !put_a(nn, av, nullptr))
                    return CLOWNS_EDECK;
            }
        } else {
            for (long nn = 1; nn <= inne2; ++nn) {
                double av[3], fbv;
                if (nn <= inne) {
                    if (!fd.next() || !fd.f64(av[0]) || !fd.f64(av[1]) || !fd.f64(av[2]) || !fd.f64(fbv) ||
                        !put_a(nn, av, &fbv))
                        return CLOWNS_EDECK;
                } else if (!fd.next() || !fd.f64(av[0]) || !fd.f64(av[1]) || !fd.f64(av[2]) ||
                           !put_a(nn, av, nullptr)) {
                    return CLOWNS_EDECK;
                }
            }
        }
    }
    d->inne = (int32_t)inne;
    d->inne2 = (int32_t)inne2;
    return CLOWNS_OK;
}

static int clowns_deck_faces_impl(const char *grid_path, const clowns_deck_desc *d, double *faces) {
    if (!grid_path || !d || !faces) return CLOWNS_EBADARG;
    if (!desc_ok(d)) return CLOWNS_EBADARG;
    const int S2 = d->stride + 2;
    RecordReader grid(grid_path);
    if (!grid.ok() || !grid.next()) return CLOWNS_EDECK;
    const int lo[3] = {d->is, d->js, d->ks}, hi[3] = {d->ie, d->je, d->ke};
    for (int l = 0; l < 3; ++l) {
        const int m = hi[l] - lo[l] + 4;
        if (m > S2 || !grid.next() || (long)grid.size() != 8L * m) return CLOWNS_EDECK;
        for (int q = 0; q < S2; ++q) {
            faces[l * S2 + q] = 0.0;
            if (q < m && !grid.f64(faces[l * S2 + q])) return CLOWNS_EDECK;
        }
    }
    return CLOWNS_OK;
}

static int clowns_deck_write_grid_impl(const char *path, const clowns_deck_desc *d, const int8_t *solid,
                                      const double *faces, int32_t big_endian) {
    if (!path || !d || !solid || !faces) return CLOWNS_EBADARG;
    if (!desc_ok(d)) return CLOWNS_EBADARG;
    const long X = d->nx + 2, Y = d->ny + 2, Z = d->nz + 2;
    const int S2 = d->stride + 2;
    RecordWriter w(path, big_endian != 0);
    if (!w.ok()) return CLOWNS_EDECK;
    // output_xyzn, data_io.F90:1531-1560: extents, faces per axis, then the
    // nf = -1 cells with I outer, J, K inner as (i, k, j, nf, nfb)
    const int32_t e[6] = {d->is, d->js, d->ks, d->ie, d->je, d->ke};
    for (int k = 0; k < 6; ++k) w.i32(e[k]);
    w.end();
    const int n[3] = {d->nx, d->ny, d->nz};
    for (int l = 0; l < 3; ++l) {
        for (int q = 0; q < n[l] + 4; ++q) w.f64(faces[l * S2 + q]);
        w.end();
    }
    int32_t icc = 0;
    for (long c = 0; c < X * Y * Z; ++c) icc += solid[c] ? 1 : 0;
    w.i32(icc);
    w.end();
    for (long i = 0; i < X; ++i)
        for (long j = 0; j < Y; ++j)
            for (long k = 0; k < Z; ++k) {
                const int8_t sv = solid[i + X * (j + Y * k)];
                if (!sv) continue;
                w.i32((int32_t)(i + d->is - 1));
                w.i32((int32_t)(k + d->ks - 1));
                w.i32((int32_t)(j + d->js - 1));
                w.i32(-1);
                w.i32(file_nfb(sv));
                w.end();
            }
    return w.ok() ? CLOWNS_OK : CLOWNS_EDECK;
}

static int clowns_deck_write_data_impl(const char *path, const clowns_deck_desc *d, const int8_t *nf,
                                      const int8_t *nb, const double *un, const double *F, const double *p,
                                      double t, double dt, int32_t big_endian) {
    if (!path || !d || !nf || !nb || !un || !F || !p) return CLOWNS_EBADARG;
    if (!desc_ok(d)) return CLOWNS_EBADARG;
    const long X = d->nx + 2, Y = d->ny + 2, Z = d->nz + 2, nc = X * Y * Z;
    // the numbering from the flags and, for nf = -1 cells, the object codes
    std::vector<int32_t> nf32(nc), nfb32(nc);
    for (long c = 0; c < nc; ++c) {
        nf32[c] = nf[c];
        nfb32[c] = nf[c] < 0 ? file_nfb(nb[c]) : 0;
    }
    long inne = 0;
    const std::vector<long> cells = numbering(d, nf32, nfb32, &inne);
    RecordWriter w(path, big_endian != 0);
    if (!w.ok()) return CLOWNS_EDECK;
    w.f64(t);  // output_data, data_io.F90:717
    w.f64(d->g);
    w.f64(d->nu);
    w.end();
    if (!d->sparse) {
        // (((NFF(mn(I,k,j))%f,K=ks-1,KE),J=js-1,JE),I=is-1,IE), then %b with
        // the appended cells' 300 / 200 put back (data_io.F90:719-737); a
        // cell outside the case reads as the sentinel cell 0 (f = -1, b = 0)
        for (int which = 0; which < 2; ++which) {
            for (long i = 0; i < X; ++i)
                for (long j = 0; j < Y; ++j)
                    for (long k = 0; k < Z; ++k) {
                        const long c = i + X * (j + Y * k);
                        int32_t v = nf[c];
                        if (which == 1) v = nf[c] < 0 ? file_nfb(nb[c]) : nb[c];
                        w.i32(v);
                    }
            w.end();
        }
        for (long q = 0; q < (long)cells.size(); ++q) {
            const long c = cells[q];
            w.f64(un[c]);
            w.f64(un[nc + c]);
            w.f64(un[2 * nc + c]);
            if (q < inne) {
                w.f64(F[c]);
                w.f64(p[c]);
            }
            w.end();
        }
    } else {
        // only cells with state: not (f = 0, U = 0, F < 1e-13) (data_io.F90:745-759)
        auto keep = [&](long c) {
            const double s = un[c] * un[c] + un[nc + c] * un[nc + c] + un[2 * nc + c] * un[2 * nc + c];
            return !(nf[c] == 0 && s == 0.0 && F[c] < 1.0e-13);
        };
        int32_t icc = 0, icc2 = 0;
        for (long q = 0; q < (long)cells.size(); ++q)
            if (keep(cells[q])) (q < inne ? icc : icc2) += 1;
        w.i32(icc);
        w.i32(icc2);
        w.end();
        for (long q = 0; q < (long)cells.size(); ++q) {
            const long c = cells[q];
            if (!keep(c)) continue;
            w.i32((int32_t)(q + 1));
            w.f64(un[c]);
            w.f64(un[nc + c]);
            w.f64(un[2 * nc + c]);
            if (q < inne) {
                w.f64(F[c]);
                w.f64(p[c]);
                w.i32(nf[c]);
                w.i32(nf[c] < 0 ? kept_b(nb[c]) : nb[c]);
            }
            w.end();
        }
    }
    w.f64(dt);
    w.end();
    return w.ok() ? CLOWNS_OK : CLOWNS_EDECK;
}

// The C-ABI: the readers and writers allocate host vectors; an allocation
// failure is reported as CLOWNS_ENOMEM rather than crossing the boundary.
template <typename F> static int guarded(F f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return CLOWNS_ENOMEM;
    } catch (...) {
        return CLOWNS_EDECK;
    }
}

extern "C" int clowns_deck_open(const char *grid_path, const char *data_path, clowns_deck_desc *d) {
    return guarded([&] { return clowns_deck_open_impl(grid_path, data_path, d); });
}

extern "C" int clowns_deck_read(const char *grid_path, const char *data_path, const char *fdata_path,
                                clowns_deck_desc *d, int8_t *solid, double *F, double *un, double *p,
                                double *a, double *fb, double *x, double *dx) {
    return guarded([&] {
        return clowns_deck_read_impl(grid_path, data_path, fdata_path, d, solid, F, un, p, a, fb, x, dx);
    });
}

extern "C" int clowns_deck_faces(const char *grid_path, const clowns_deck_desc *d, double *faces) {
    return guarded([&] { return clowns_deck_faces_impl(grid_path, d, faces); });
}

extern "C" int clowns_deck_write_grid(const char *path, const clowns_deck_desc *d, const int8_t *solid,
                                      const double *faces, int32_t big_endian) {
    return guarded([&] { return clowns_deck_write_grid_impl(path, d, solid, faces, big_endian); });
}

extern "C" int clowns_deck_write_data(const char *path, const clowns_deck_desc *d, const int8_t *nf,
                                      const int8_t *nb, const double *un, const double *F, const double *p,
                                      double t, double dt, int32_t big_endian) {
    return guarded([&] { return clowns_deck_write_data_impl(path, d, nf, nb, un, F, p, t, dt, big_endian); });
}
