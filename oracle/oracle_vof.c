/*
 * oracle_vof.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatement of the VOF update cal_vof (src2LOWNS/1sub-vof.F90:10-38)
 * of the shipped build (-DNORMAL -DSLIDE0, no DRI/HENDO/DAKU), serial:
 *   face_flux()     cal_quvwf + QFF            1sub-vof.F90:40-116
 *   surface stage   quvwf_kaidan               1sub-vof.F90:157-183 (builds
 *                   without NORMAL), or, for NORMAL builds, the PLIC face
 *                   fluxes of quvwf_normal (1sub-vof.F90:119-154) given by
 *                   the caller as (value, mask) per face
 *   modify()        quvwf_modify               1sub-vof.F90:186-213
 *   new_f()         cal_new_f, with its dt retry  1sub-vof.F90:215-269
 *   slide()         simple_slide(0), cal_fslide_m0  1sub-vof.F90:302-414
 *   df_vof()        df_vof + renzoku_fv        1sub-vof.F90:271-300,
 *                                              lib/src/math_func.F90:51-54
 * set_nflag, which cal_vof calls last, is the classification restated in
 * oracle_assembly.c.
 *
 * Box layout as in oracle_assembly.c: c = i + X*(j + Y*k); a solid cell is
 * the reference's mn = 0 sentinel (flags f = -1, b = 0; a(:,0) = 0; its
 * face flux reads as 0).  un[l][c], a[l][c], qf[l][c] belong to the low face
 * of c in direction l.  The surface list sfn(1:isfn) is, after set_nflag,
 * the nf = 2 cells in k, j, i scan order (nfd2015.F90:67-95; a demotion
 * restarts the scan), i.e. ascending c; quvwf_modify walks it in that order.
 *
 * Pinning (tests/test_oracle_vof.py): bit-exact against the flang build of
 * those routines (oracle/_ref/libref_vof.so) on synthetic grids.
 * plic_flux (quvwf_normal from PLIC segment volumes): parity unpinned -- the
 * reference's loop needs its PLIC state, which is out of scope; the test
 * only restates the same formula.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VOF_EVOF 4008      /* the reference STOPs (F above F_lim on a surface cell, dt below DT_lim) */
#define VOF_ENONFINITE 4001 /* a correction aimed at a solid cell (the reference's sentinel) */

static const double ONE = 1.0, ZERO = 0.0, TINY = 1e-15, F_LIM = 1.0 + 1e-5, DT_LIM = 1e-6;

typedef struct {
    int X, Y, Z, S;
    long nc, st[3];
    const int8_t *solid, *nf, *nb;
    const double *F, *Fp, *fb, *a, *un, *dx;
    double *qf, *Fn, dt;
    int n_modified, n_slid;
} V;

static inline double dxc(const V *v, int l, long c) {
    int i = (int)(c % v->X), j = (int)((c / v->X) % v->Y), k = (int)(c / ((long)v->X * v->Y));
    int q = l == 0 ? i : (l == 1 ? j : k);
    return v->dx[l * v->S + q];
}
/* a(l, mn(c)) and qf(l, mn(c)): the sentinel for a solid cell */
static inline double A_(const V *v, int l, long c) { return v->solid[c] ? ZERO : v->a[l * v->nc + c]; }
static inline double Q_(const V *v, int l, long c) { return v->solid[c] ? ZERO : v->qf[l * v->nc + c]; }

/* df_vof(nn) = fp - dt * renzoku_fv(...) (1sub-vof.F90:271-300) */
static double df_vof(const V *v, long c) {
    const long cx = c + v->st[0], cy = c + v->st[1], cz = c + v->st[2];
    const double up = Q_(v, 0, cx), um = Q_(v, 0, c), vp = Q_(v, 1, cy), vm = Q_(v, 1, c);
    const double wp = Q_(v, 2, cz), wm = Q_(v, 2, c);
    const double axp = A_(v, 0, cx), axm = A_(v, 0, c), ayp = A_(v, 1, cy), aym = A_(v, 1, c);
    const double azp = A_(v, 2, cz), azm = A_(v, 2, c);
    const double ren = ((up * axp - um * axm) / dxc(v, 0, c) + (vp * ayp - vm * aym) / dxc(v, 1, c) +
                        (wp * azp - wm * azm) / dxc(v, 2, c)) / v->fb[c];
    return v->Fp[c] - v->dt * ren;
}

static inline double fsign(double a, double b) { return signbit(b) ? -fabs(a) : fabs(a); }

/* QFF (1sub-vof.F90:61-116) for the box model: donor and acceptor are
 * numbered cells, so the donor flag is 0, 1 or 2 */
static double qff(const V *v, double qx, double dxx, long nnd, long nna, int nde) {
    if (v->solid[nnd] || v->solid[nna] || qx == ZERO) return ZERO;
    const int f = v->nf[nnd], b = v->nb[nnd];
    const double fd = v->F[nnd];
    if (f == 2 && abs(b) == nde) {
        if ((double)b * qx < ZERO) {
            double t = fabs(qx) - (ONE - fd) * dxx / v->dt;
            return fsign(ONE, qx) * (t > ZERO ? t : ZERO);
        }
        double t = fd * dxx / v->dt;
        double m = fabs(qx);
        return fsign(ONE, qx) * (m < t ? m : t);
    }
    return qx * fd;
}

/* cal_quvwf (1sub-vof.F90:40-59) */
static void face_flux(V *v) {
    for (long c = 0; c < v->nc; ++c) {
        for (int l = 0; l < 3; ++l) v->qf[l * v->nc + c] = ZERO;
        if (v->solid[c]) continue;
        for (int l = 0; l < 3; ++l) {
            const double u = v->un[l * v->nc + c];
            const int iu = (int)((ONE - fsign(ONE, u)) / 2);
            const long nnd = c + (iu - 1) * v->st[l];
            const long nna = c - iu * v->st[l];
            /* dx(l, lg(l)-1+iu): the donor's size along l */
            const double dxx = dxc(v, l, iu ? c : c - v->st[l]);
            v->qf[l * v->nc + c] = qff(v, u, dxx, nnd, nna, l + 1);
        }
    }
}

/* quvwf_kaidan (1sub-vof.F90:157-183) */
static void kaidan(V *v) {
    for (long c = 0; c < v->nc; ++c) {
        if (v->solid[c] || v->nf[c] != 2) continue;
        const double fb = v->fb[c], fl = v->F[c];
        for (int l = 0; l < 2; ++l) {
            const double u = v->un[l * v->nc + c], a = v->a[l * v->nc + c];
            if (u < ZERO && a > ZERO && a < ONE && a != fb) {
                double t = a - fb + fl * fb;
                v->qf[l * v->nc + c] = u * (t > ZERO ? t : ZERO) / a;
            }
        }
        for (int l = 0; l < 2; ++l) {
            const long p = c + v->st[l];
            if (v->solid[p]) continue; /* un(:,0) = 0 */
            const double u = v->un[l * v->nc + p], a = v->a[l * v->nc + p];
            if (u > ZERO && a > ZERO && a < ONE && a != fb) {
                double t = a - fb + fl * fb;
                v->qf[l * v->nc + p] = u * (t > ZERO ? t : ZERO) / a;
            }
        }
    }
}

/* quvwf_normal (1sub-vof.F90:119-154) with cal_idou2's flux
 * (1sub-normal.F90:29-99) from given PLIC segment volumes: for each listed
 * surface cell and each of its faces (l, ll = 0 low / 1 high) through which
 * it drains (un * (1 - 2 ll) < 0, a != 0), qf = sign(u) * V / area / dt / a,
 * 0 where |u| dt < SMALL or V = 0.  vol[6*m + 2*l + ll] belongs to cells[m]. */
static void plic_flux(V *v, int n, const int32_t *cells, const double *vol) {
    const double SMALL = 1e-9;
    for (int m = 0; m < n; ++m) {
        const long c = cells[m];
        const double dxx = dxc(v, 0, c), dyy = dxc(v, 1, c), dzz = dxc(v, 2, c);
        for (int l = 0; l < 3; ++l)
            for (int ll = 0; ll < 2; ++ll) {
                const long nn = c + ll * v->st[l];
                const double lsn = 1.0 - 2.0 * ll;
                const double u = v->solid[nn] ? ZERO : v->un[l * v->nc + nn];
                const double a = A_(v, l, nn);
                if (u * lsn >= ZERO || a == ZERO) continue;
                double idou;
                if (fabs(u) * v->dt < SMALL) {
                    idou = ZERO;
                } else {
                    const double area = l == 2 ? dxx * dyy : (l == 1 ? dxx * dzz : dzz * dyy);
                    const double vs = vol[6 * m + 2 * l + ll];
                    idou = vs == ZERO ? ZERO : fsign(ONE, u) * vs / area / v->dt;
                }
                v->qf[l * v->nc + nn] = idou / a;
            }
    }
}

/* quvwf_modify (1sub-vof.F90:186-213), sfn order = ascending c */
static int modify(V *v) {
    for (long c = 0; c < v->nc; ++c) {
        if (v->solid[c] || v->nf[c] != 2) continue;
        const double d = df_vof(v, c);
        if (d < F_LIM) continue;
        const int b = v->nb[c], l = abs(b) - 1;
        const long n1 = b > 0 ? c : c + v->st[l];
        if (v->solid[n1]) return VOF_ENONFINITE;
        const double s = b > 0 ? ONE : -ONE;
        double *q = &v->qf[l * v->nc + n1];
        *q = *q - dxc(v, l, c) * (d - ONE) / v->dt * v->fb[c] / v->a[l * v->nc + n1] * s;
        ++v->n_modified;
    }
    return 0;
}

/* cal_new_f (1sub-vof.F90:215-269): returns the number of dt reductions or
 * a negative error */
static int new_f(V *v) {
    int retries = 0;
    for (;;) {
        int icon = 0;
        for (long c = 0; c < v->nc; ++c) {
            if (v->solid[c]) continue;
            double d = df_vof(v, c);
            if (-d * v->dt * v->fb[c] > 0.4) {
                icon = 30;
            } else if (fabs(d * v->dt) < TINY && d != ZERO) {
                d = ZERO;
            }
            if (d > F_LIM && v->nf[c] == 2) return -VOF_EVOF;
            v->Fn[c] = d;
        }
        if (icon != 30) return retries;
        v->dt = v->dt / 1.1;
        if (v->dt < DT_LIM) return -VOF_EVOF;
        ++retries;
    }
}

/* simple_slide(0) and cal_fslide_m0(nn, nnm, -3) (1sub-vof.F90:302-414) */
static void slide(V *v) {
    const int nx = v->X - 2, ny = v->Y - 2, nz = v->Z - 2;
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
            for (int k = nz; k >= 1; --k) {
                const long c = i + (long)v->X * (j + (long)v->Y * k);
                const long m = c - v->st[2];
                if (v->solid[c]) break;
                if (v->nf[c] == 1) continue;
                if (v->solid[m]) continue;
                const double dx0 = v->dx[0 * v->S + i], dx1 = v->dx[1 * v->S + j];
                const double vmesh = dx0 * dx1 * v->dx[2 * v->S + k];
                const double dvc = vmesh * v->fb[c];
                const double dvm = dx0 * dx1 * v->dx[2 * v->S + k - 1] * v->fb[m];
                if (dvc * dvm == ZERO) continue;
                v->Fn[c] = df_vof(v, c);
                v->Fn[m] = df_vof(v, m);
                const double t1 = dvm * (ONE - v->Fn[m]), t2 = dvc * v->Fn[c];
                const double f2 = t1 < t2 ? t1 : t2;
                if (fabs(f2) == ZERO) continue;
                double *q = &v->qf[2 * v->nc + c];
                *q = *q + fsign(ONE, -3.0) * f2 / vmesh / v->dt * v->dx[2 * v->S + k] / v->a[2 * v->nc + c];
                ++v->n_slid;
                const double ft = df_vof(v, c);
                v->Fn[c] = fabs(ft) < TINY ? ZERO : ft;
                v->Fn[m] = df_vof(v, m);
            }
}

/* One cal_vof without its closing set_nflag.  mode 0: NORMAL build, with the
 * PLIC face fluxes qs where qs_mask != 0 (either may be NULL: none);
 * mode 1: quvwf_kaidan; mode 2: NORMAL build, quvwf_normal on the n_plic
 * listed cells with their PLIC segment volumes (plic_flux above).  F: current F (after set_nflag), Fp: F of the
 * previous step, nf/nb: the flags.  Out: qf (3*ncell), Fn (ncell; solid
 * cells keep F), dt (reduced by cal_new_f's retries), *retries, and
 * counts[0..1] (may be NULL): surface cells corrected by quvwf_modify, cells
 * that moved volume down in simple_slide.
 * Returns 0, VOF_EVOF or VOF_ENONFINITE. */
int oracle_vof(int nx, int ny, int nz, int stride, const int8_t *solid, const int8_t *nf, const int8_t *nb,
               const double *F, const double *Fp, const double *fb, const double *a, const double *un,
               const double *dx, double *dt, int mode, const double *qs, const int8_t *qs_mask, int n_plic,
               const int32_t *plic_cells, const double *plic_vol, double *qf, double *Fn, int *retries,
               int *counts) {
    V v;
    v.n_modified = v.n_slid = 0;
    v.X = nx + 2;
    v.Y = ny + 2;
    v.Z = nz + 2;
    v.S = stride;
    v.nc = (long)v.X * v.Y * v.Z;
    v.st[0] = 1;
    v.st[1] = v.X;
    v.st[2] = (long)v.X * v.Y;
    v.solid = solid;
    v.nf = nf;
    v.nb = nb;
    v.F = F;
    v.Fp = Fp;
    v.fb = fb;
    v.a = a;
    v.un = un;
    v.dx = dx;
    v.qf = qf;
    v.Fn = Fn;
    v.dt = *dt;
    memcpy(Fn, F, sizeof(double) * v.nc);
    face_flux(&v);
    if (mode == 1) {
        kaidan(&v);
    } else if (mode == 2) {
        plic_flux(&v, n_plic, plic_cells, plic_vol);
    } else if (qs && qs_mask) {
        for (long c = 0; c < v.nc; ++c)
            for (int l = 0; l < 3; ++l)
                if (!solid[c] && qs_mask[l * v.nc + c]) qf[l * v.nc + c] = qs[l * v.nc + c];
    }
    int rc = modify(&v);
    if (rc) return rc;
    int r = new_f(&v);
    if (r < 0) return -r;
    *retries = r;
    slide(&v);
    *dt = v.dt;
    if (counts) {
        counts[0] = v.n_modified;
        counts[1] = v.n_slid;
    }
    return 0;
}
