// geometry.cpp -- host side of the pressure-system assembly: cell
// classification, row numbering and the surface position the surface rows
// need.  These are serial, order-dependent algorithms in the reference (the
// OpenMP loops of set_nflag race, so only the one-thread order is defined),
// so they run on the host once per time step; the per-row coefficient and
// right-hand-side work runs on the GPU (assembly kernels in pcg_solver.hip).
//
//   classify        set_nflag, src2LOWNS/nfd2015.F90:7-306 (fxx1 :308-332),
//                   shipped flags -DNORMAL -DOBJECT, objno = 0
//   number_rows     src2LOWNS/1sub-momentum.F90:278-296 (ik == 0 block)
//   surface_pos     set_initial_surface_position,
//                   src2LOWNS/1sub-normal-ini.F90:23-40 (normal component)
//   clowns_set_dx   set_dx, src2LOWNS/1sub-main.F90:53-74, with round,
//                   lib/src/xy_calc.f90:101-112 (start-up, per axis)
//
// Box convention: cells i=0..nx+1, j=0..ny+1, k=0..nz+1 with a solid outer
// ring, index c = i + X*(j + Y*k), X = nx+2, Y = ny+2 (k-major: the row
// order of the reference's fln list).  solid[c] != 0: a cell of flag -1.
// solid[c] = 1: a cell outside the case, which reads as the reference's
// sentinel cell 0 (mn = 0): flag -1, F = fb = 0, closed faces.  Any other
// value: an object cell of the case (numbered by the case reader,
// src2LOWNS/data_io.F90:171-199), read with its own F, fb and a, whose nb
// becomes that value (its object code: 4 a pressure-boundary cell).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <new>
#include <vector>

#include "clowns_pcg.h"

namespace {

class Classifier {
  public:
    Classifier(const clowns_grid_desc &g, const int8_t *solid, double *F, const double *fb,
               const double *a, const double *dx, const double *xf, int8_t *nf, int8_t *nb)
        : nx_(g.nx), ny_(g.ny), nz_(g.nz), X_(g.nx + 2), Y_(g.ny + 2), Z_(g.nz + 2),
          stride_(g.stride), solid_(solid), F_(F), fb_(fb), dx_(dx), xf_(xf), nf_(nf), nb_(nb) {
        ncell_ = (long)X_ * Y_ * Z_;
        for (int l = 0; l < 3; ++l) a_[l] = a + (size_t)l * ncell_;
        step_[0] = 1;
        step_[1] = X_;
        step_[2] = (long)X_ * Y_;
    }

    long ncell() const { return ncell_; }
    long step(int l) const { return step_[l]; }
    void coords(long c, int p[3]) const {
        p[0] = (int)(c % X_);
        p[1] = (int)((c / X_) % Y_);
        p[2] = (int)(c / step_[2]);
    }
    // Every cell touched below is a flow cell (i in 1..n) or its +-1
    // neighbour, so c +- step stays inside the box.
    int flag(long c) const { return solid_[c] ? -1 : nf_[c]; }
    double fill(long c) const { return solid_[c] == 1 ? 0.0 : F_[c]; }
    double open(long c) const { return solid_[c] == 1 ? 0.0 : fb_[c]; }
    double face(int l, long c) const { return solid_[c] == 1 ? 0.0 : a_[l][c]; }
    double h(int l, int i) const { return dx_[(size_t)l * stride_ + i]; }
    double xlo(int l, int i) const { return xf_[(size_t)l * stride_ + i]; }

    void run() {
        initial_flags();
        for (;;) {
            build_lists();
            if (drop_unsupported()) continue;  // nfd2015.F90:128-131
            if (drop_one_sided()) continue;    // :155-158
            if (orient()) continue;            // :246 (igai)
            break;
        }
    }

    const std::vector<long> &fln() const { return fln_; }

  private:
    // nfd2015.F90:36-59
    void initial_flags() {
        imax_ = 0;
        kmax_ = 0;
        for (long c = 0; c < ncell_; ++c) {
            if (solid_[c]) {
                nf_[c] = -1;
                nb_[c] = solid_[c] == 1 ? 0 : solid_[c];
                continue;
            }
            int p[3];
            coords(c, p);
            const double f = F_[c];
            const bool air = f <= 1e-12 ||
                             (f * fb_[c] < 1.0e-3 && flag(c - step_[2]) == -1 && fill(c + step_[2]) == 0.0);
            nb_[c] = 0;
            if (air) {
                nf_[c] = 0;
                F_[c] = f > 0.0 ? f : 0.0;
            } else {
                nf_[c] = 1;
                imax_ = std::max(imax_, p[0]);
                kmax_ = std::max(kmax_, p[2]);
            }
        }
        imax_ = std::min(nx_, imax_ + 1);
        kmax_ = std::min(nz_, kmax_ + 1);
    }

    // nfd2015.F90:66-102: fln in (k, j, i) order; a flow cell with an air
    // neighbour across an open face becomes a surface cell (b = -3 for now).
    void build_lists() {
        fln_.clear();
        sfn_.clear();
        for (int k = 1; k <= kmax_; ++k)
            for (int j = 1; j <= ny_; ++j)
                for (int i = 1; i <= imax_; ++i) {
                    const long c = i + (long)X_ * (j + (long)Y_ * k);
                    if (flag(c) < 1) continue;
                    fln_.push_back(c);
                    bool surface = false;
                    for (int l = 0; l < 3 && !surface; ++l) {
                        const long up = c + step_[l], dn = c - step_[l];
                        surface = (flag(up) == 0 && face(l, up) > 0.0) || (flag(dn) == 0 && a_[l][c] > 0.0);
                    }
                    if (surface) {
                        nf_[c] = 2;
                        nb_[c] = -3;
                        sfn_.push_back(c);
                    } else {
                        nf_[c] = 1;
                    }
                }
    }

    // nfd2015.F90:110-126: a surface cell with no fluid neighbour (and no
    // object below) becomes air.  Descending sfn order.
    bool drop_unsupported() {
        bool changed = false;
        for (size_t s = sfn_.size(); s-- > 0;) {
            const long c = sfn_[s];
            if (flag(c - step_[2]) == -1) continue;
            bool wet = false;
            for (int l = 0; l < 3 && !wet; ++l) wet = flag(c + step_[l]) == 1 || flag(c - step_[l]) == 1;
            if (wet) continue;
            nf_[c] = 0;
            nb_[c] = 0;
            changed = true;
        }
        return changed;
    }

    // nfd2015.F90:135-152: keep a surface cell only if some axis has fluid on
    // one side and air on the other.  Order dependent: descending sfn.
    bool drop_one_sided() {
        bool changed = false;
        for (size_t s = sfn_.size(); s-- > 0;) {
            const long c = sfn_[s];
            if (flag(c - step_[2]) == -1 && flag(c + step_[2]) == 0) continue;
            bool keep = false;
            for (int l = 0; l < 3 && !keep; ++l) {
                const int fp = flag(c + step_[l]), fm = flag(c - step_[l]);
                keep = (fp == 1 && fm == 0) || (fm == 1 && fp == 0);
            }
            if (keep) continue;
            nf_[c] = 0;
            nb_[c] = 0;
            changed = true;
        }
        return changed;
    }

    // fxx1 (nfd2015.F90:308-332): fill-weighted area on the air side; the
    // running sum is multiplied by the third axis' size inside the loop.
    double fxx1(int ix, int ll, const int lg[3]) const {
        double s = 0.0;
        for (int l = 0; l < 3; ++l) {
            if (l == ll) continue;
            const long base = (lg[0] + (long)X_ * (lg[1] + (long)Y_ * lg[2])) + ix * step_[ll];
            for (int ii = -1; ii <= 1; ++ii) {
                const long n = base + ii * step_[l];
                s = s + h(l, lg[l] + ii) * fill(n) * open(n);
            }
            for (int m = 0; m < 3; ++m)
                if (m != l && m != ll) s = s * h(m, lg[m]);
        }
        return s * h(ll, lg[ll] + ix);
    }

    // nfd2015.F90:167-243: orientation b = +-(axis+1) of each surface cell.
    // Returns true when a cell had no orientation (it became air: restart).
    bool orient() {
        int igai = 0;
        for (size_t s = sfn_.size(); s-- > 0;) {
            const long c = sfn_[s];
            nb_[c] = 0;
            int ix[3] = {0, 0, 0};
            for (int l = 0; l < 3; ++l) {
                const long up = c + step_[l], dn = c - step_[l];
                if (flag(dn) == 1 && flag(up) == 0 && face(l, up) > 0.0)
                    ix[l] = -1;
                else if (flag(dn) == 0 && flag(up) == 1 && a_[l][c] > 0.0)
                    ix[l] = 1;
            }
            if (std::abs(ix[0]) + std::abs(ix[1]) + std::abs(ix[2]) == 0 &&
                flag(c - step_[2]) == -1 && flag(c + step_[2]) == 0) {
                // surface cell resting on an object (NORMAL + OBJECT, :192-212)
                if (fb_[c] != 0.0) {
                    ix[2] = -1;
                } else {
                    for (int l = 0; l < 2; ++l) {
                        const long up = c + step_[l], dn = c - step_[l];
                        if (flag(dn) > 0 && flag(up) == 0 && face(l, up) > 0.0)
                            ix[l] = -1;
                        else if (flag(dn) == 0 && flag(up) > 0 && a_[l][c] > 0.0)
                            ix[l] = 1;
                    }
                    if (ix[0] == 0 && ix[1] == 0) ix[2] = -1;
                }
            }
            const int n = std::abs(ix[0]) + std::abs(ix[1]) + std::abs(ix[2]);
            int dir = 0;
            if (n == 1) {
                for (int l = 1; l < 3; ++l)
                    if (std::abs(ix[l]) > std::abs(ix[dir])) dir = l;
                nb_[c] = (int8_t)((dir + 1) * ix[dir]);
            } else if (n == 0) {
                ++igai;
                nf_[c] = 0;
            } else {
                int p[3];
                coords(c, p);
                double w[3];
                for (int l = 0; l < 3; ++l) w[l] = ix[l] == 0 ? 0.0 : fxx1(ix[l], l, p);
                for (int l = 1; l < 3; ++l)
                    if (w[l] > w[dir]) dir = l;
                nb_[c] = (int8_t)((dir + 1) * ix[dir]);
            }
        }
        return igai != 0;
    }

    int nx_, ny_, nz_, X_, Y_, Z_, stride_;
    long ncell_;
    long step_[3];
    const int8_t *solid_;
    double *F_;
    const double *fb_;
    const double *a_[3];
    const double *dx_, *xf_;
    int8_t *nf_, *nb_;
    int imax_ = 0, kmax_ = 0;
    std::vector<long> fln_, sfn_;
};

} // namespace

static int classify_impl(const clowns_grid_desc *g, const int8_t *solid, double *F, const double *fb,
                         const double *a, const double *dx, const double *xf, int8_t *nf, int8_t *nb,
                         int32_t *mno, int32_t *row2cell, double *sp_row, int32_t cap, int32_t *nrows,
                         double *mp) {
    if (!g || !solid || !F || !fb || !a || !dx || !xf || !nf || !nb || !mno || !row2cell ||
        !nrows || !mp)
        return CLOWNS_EBADARG;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->stride < std::max(g->nx, std::max(g->ny, g->nz)) + 2)
        return CLOWNS_EBADARG;
    Classifier C(*g, solid, F, fb, a, dx, xf, nf, nb);
    // the outer ring must be solid (closed walls): every neighbour access
    // of a flow cell then stays inside the box
    const int X = g->nx + 2, Y = g->ny + 2, Z = g->nz + 2;
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < Y; ++j)
            for (int i = 0; i < X; ++i)
                if ((i == 0 || j == 0 || k == 0 || i == X - 1 || j == Y - 1 || k == Z - 1) &&
                    !solid[i + (long)X * (j + (long)Y * k)])
                    return CLOWNS_EBADARG;
    C.run();

    // mp = (smallest cell size)^2 (1sub-matrix.F90:17)
    double h = 1e300;
    const int n3[3] = {g->nx, g->ny, g->nz};
    for (int l = 0; l < 3; ++l)
        for (int i = 1; i <= n3[l]; ++i) {
            const double v = C.h(l, i);
            if (v > 0.0 && v < h) h = v;
        }
    *mp = h * h;

    // Row numbering over fln (1sub-momentum.F90:278-296).
    for (long c = 0; c < C.ncell(); ++c) mno[c] = 0;
    int32_t N = 0;
    for (long c : C.fln()) {
        const int f = nf[c];
        if (f == 0 || f == -1 || fb[c] == 0.0) continue;
        if (f == 2 && nb[c] != 0) {
            const int l = std::abs(nb[c]) - 1, sg = nb[c] > 0 ? 1 : -1;
            if (C.flag(c + sg * C.step(l)) == -1) continue;
        }
        if (N >= cap) return CLOWNS_EBADARG;
        mno[c] = ++N;
        row2cell[N - 1] = (int32_t)c;
        if (sp_row) {
            double s = 0.0;
            if (f == 2 && nb[c] != 0) {
                // surface position along the normal (1sub-normal-ini.F90:27-38)
                int p[3];
                C.coords(c, p);
                const int b = nb[c], l = std::abs(b) - 1;
                double fff;
                if (b == -3 && fb[c] < 1.0)
                    fff = 1.0 - fb[c] * (1.0 - F[c]);
                else if (b < 0)
                    fff = F[c];
                else
                    fff = 1.0 - F[c];
                s = C.xlo(l, p[l]) + C.h(l, p[l]) * fff;
                if (std::fabs(s) < 1e-15) s = 0.0;
            }
            sp_row[N - 1] = s;
        }
    }
    *nrows = N;
    return CLOWNS_OK;
}

// The classifier keeps host lists (fln, sfn): an allocation failure is
// reported as CLOWNS_ENOMEM rather than thrown across the C-ABI.
extern "C" int clowns_classify(const clowns_grid_desc *g, const int8_t *solid, double *F,
                               const double *fb, const double *a, const double *dx,
                               const double *xf, int8_t *nf, int8_t *nb, int32_t *mno,
                               int32_t *row2cell, double *sp_row, int32_t cap, int32_t *nrows,
                               double *mp) {
    try {
        return classify_impl(g, solid, F, fb, a, dx, xf, nf, nb, mno, row2cell, sp_row, cap, nrows, mp);
    } catch (const std::bad_alloc &) {
        return CLOWNS_ENOMEM;
    }
}

namespace {
// Fortran NINT: nearest integer, halves away from zero (v - trunc(v) is
// exact in binary floating point, so the half test is too).
double nint_away(double v) {
    const double t = std::trunc(v);
    return std::fabs(v - t) >= 0.5 ? t + (v > 0 ? 1.0 : -1.0) : t;
}
} // namespace

extern "C" int clowns_set_dx(double *x, int32_t m, double *dx) {
    if (!x || !dx || m < 4) return CLOWNS_EBADARG;
    // positions in x[]/dx[] of the reference's indices ls and le
    const int ls = 2, le = m - 2;
    if (x[le + 1] == 0.0) x[le + 1] = x[le] + x[le] - x[le - 1];
    // decimals guessed from the first cell size
    int dec = 0;
    while (dec <= 10 && !((x[ls + 1] - x[ls]) / std::pow(10.0, 4 - dec) > 1.0)) ++dec;
    const double A = std::pow(10.0, dec);
    for (int i = 0; i <= le; ++i) dx[i] = nint_away((x[i + 1] - x[i]) * A) / A;
    dx[le + 1] = dx[le];
    return CLOWNS_OK;
}

