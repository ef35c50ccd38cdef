/*
 * oracle_assembly.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatement of how 2CLOWNS-3D builds the pressure system that
 * PCGPME solves, for one pressure-correction pass (ik = 0):
 *   classify()      set_nflag            src2LOWNS/nfd2015.F90:7-306 (serial,
 *                                        descending-sfn semantics), fxx1 :308-332
 *   surface_pos()   set_initial_surface_position  src2LOWNS/1sub-normal-ini.F90:6-47
 *   init_matrix()   init_matrix, cal_alu_f        src2LOWNS/1sub-matrix.F90:3-66,82-87
 *   number_rows()   cal_err_of_continuity, ik==0  src2LOWNS/1sub-momentum.F90:273-296
 *   rhs()           cal_err_of_continuity RHS     src2LOWNS/1sub-momentum.F90:313-346
 *                   renzoku_n / renzoku_fv        src2LOWNS/1sub-main.F90:579-602,
 *                                                 lib/src/math_func.F90:51-54
 *   cal_matrix()    cal_matrix_element, cal_alu   src2LOWNS/1sub-momentum.F90:465-551,
 *                                                 src2LOWNS/1sub-matrix.F90:75-80
 * with the shipped build's flags (-DNORMAL -DOBJECT, no DAKU/DRI/HENDO).
 *
 * Grid: cells i=1..nx, j=1..ny, k=1..nz (is=js=ks=1, ie=nx+1, je=ny+1,
 * ke=nz+1) plus the boundary ring i=0, nx+1 (etc.), stored on the box
 * (nx+2)*(ny+2)*(nz+2) with index c = i + X*(j + Y*k), X=nx+2, Y=ny+2.
 * solid[c] != 0 marks a cell of flag -1 (nff%f = iob).  solid[c] = 1: a cell
 * outside the case, which the reference gives mn = 0 (src2LOWNS/
 * data_io.F90:168-186), i.e. the sentinel nff(0) = (f=-1, b=0) with
 * fb(0) = 0, a(:,0) = 0 and un(:,0) = 0 (data_io.F90:274-277,
 * main_2CLOWNS.F90:73); that is what reading such a neighbour yields here.
 * Any other value is an object cell of the case (numbered by the case
 * reader, data_io.F90:171-199, so its f, fb, a and un are read as they are)
 * and is its object code, which becomes its nb: for a boundary object with
 * nfb = b (1 <= b < 200) b itself when 2 <= b <= 126 (4: a pressure-boundary
 * cell), -2 when b = 1, 124 - b when b >= 127; -1 and -128 for the object
 * cells appended to the numbering with nfb 300 and 200 (b = -1 and 0).  a[l][c] is the open ratio of the low face of c in
 * direction l (between c - e_l and c); un[l][c] the velocity on that face.
 * dx[l][i] (i = 0..n_l+1) the cell sizes after set_dx rounding, xf[l][i]
 * (i = 0..n_l+1) the low-face coordinates x(l,i).
 *
 * Pinning (tests/test_oracle_assembly.py): bit-exact against the flang build
 * of those reference routines (oracle/_ref/libref_assembly.so) on the
 * synthetic configurations; the uniform-grid coefficients in closed form.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz, X, Y, Z;
    long ncell;
    const int8_t *solid;
    double *F;          /* F is clipped to >= 0 for air cells, as the reference does */
    const double *fb, *a[3], *un[3], *dx[3], *xf[3];
    int8_t *nf, *nb;    /* out: flags f and b per cell */
    int *fln, *sfn;     /* lists of cell indices */
    int ifln, isfn;
} G;

static inline long cid(const G *g, int i, int j, int k) {
    return (long)i + (long)g->X * ((long)j + (long)g->Y * (long)k);
}
static inline int inside(const G *g, int i, int j, int k) {
    return i >= 0 && i < g->X && j >= 0 && j < g->Y && k >= 0 && k < g->Z;
}
/* nff(mn(...))%f of a cell: solid / out of box -> the sentinel -1 */
static inline int fflag(const G *g, int i, int j, int k) {
    if (!inside(g, i, j, k)) return -1;
    long c = cid(g, i, j, k);
    return g->solid[c] ? -1 : g->nf[c];
}
/* nff(mn(...))%b of a solid cell: its object code (0 outside the case) */
static inline int bflag(const G *g, int i, int j, int k) {
    if (!inside(g, i, j, k)) return 0;
    long c = cid(g, i, j, k);
    return g->solid[c] ? (g->solid[c] == 1 ? 0 : g->solid[c]) : g->nb[c];
}
/* f, fb, a, un of mn(...): a cell outside the case reads the sentinel's 0 */
static inline double fval(const G *g, int i, int j, int k) {
    if (!inside(g, i, j, k)) return 0.0;
    long c = cid(g, i, j, k);
    return g->solid[c] == 1 ? 0.0 : g->F[c];
}
static inline double fbval(const G *g, int i, int j, int k) {
    if (!inside(g, i, j, k)) return 0.0;
    long c = cid(g, i, j, k);
    return g->solid[c] == 1 ? 0.0 : g->fb[c];
}
/* a(l, mn(i,j,k)) */
static inline double aval(const G *g, int l, int i, int j, int k) {
    if (!inside(g, i, j, k)) return 0.0;
    long c = cid(g, i, j, k);
    return g->solid[c] == 1 ? 0.0 : g->a[l][c];
}
static inline double unval(const G *g, int l, int i, int j, int k) {
    if (!inside(g, i, j, k)) return 0.0;
    long c = cid(g, i, j, k);
    return g->solid[c] == 1 ? 0.0 : g->un[l][c];
}

static const int E[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; /* lf (main_2CLOWNS.F90:33) */

/* fxx1 (nfd2015.F90:308-332), literally: the running sum is multiplied by the
 * third direction's size inside the l loop. */
static double fxx1(const G *g, int ix, int ll, const int lg[3]) {
    double s = 0.0;
    for (int l = 0; l < 3; ++l) {
        if (l == ll) continue;
        for (int ii = -1; ii <= 1; ++ii) {
            int p[3];
            for (int d = 0; d < 3; ++d) p[d] = lg[d] + ix * E[ll][d] + ii * E[l][d];
            s = s + g->dx[l][lg[l] + ii] * fval(g, p[0], p[1], p[2]) * fbval(g, p[0], p[1], p[2]);
        }
        for (int lll = 0; lll < 3; ++lll) {
            if (lll == l || ll == lll) continue;
            s = s * g->dx[lll][lg[lll]];
        }
    }
    s = s * g->dx[ll][lg[ll] + ix];
    return s;
}

/* set_nflag (nfd2015.F90:7-306) for a 3D grid, serial. */
static void classify(G *g) {
    const double VSMALL = 1e-12;
    int imax = 0, kmax = 0;
    /* :36-56 */
    for (int k = 0; k < g->Z; ++k)
        for (int j = 0; j < g->Y; ++j)
            for (int i = 0; i < g->X; ++i) {
                long c = cid(g, i, j, k);
                if (g->solid[c]) { g->nf[c] = -1; g->nb[c] = g->solid[c] == 1 ? 0 : g->solid[c]; continue; }
                double f = g->F[c];
                if (f <= VSMALL || (f * g->fb[c] < 1.0e-3 && fflag(g, i, j, k - 1) == -1 &&
                                    fval(g, i, j, k + 1) == 0.0)) {
                    g->nf[c] = 0;
                    g->nb[c] = 0;
                    g->F[c] = f > 0.0 ? f : 0.0;
                } else {
                    g->nf[c] = 1;
                    g->nb[c] = 0;
                    if (i > imax) imax = i;
                    if (k > kmax) kmax = k;
                }
            }
    const int ie = g->nx + 1, ke = g->nz + 1;
    imax = (ie - 1 < imax + 1) ? ie - 1 : imax + 1;  /* :58-59 */
    kmax = (ke - 1 < kmax + 1) ? ke - 1 : kmax + 1;
restart:  /* label 3 / 4 */
    for (;;) {
        g->isfn = 0;
        g->ifln = 0;
        /* :69-102 */
        for (int k = 1; k <= kmax; ++k)
            for (int j = 1; j <= g->ny; ++j)
                for (int i = 1; i <= imax; ++i) {
                    long c = cid(g, i, j, k);
                    int f = fflag(g, i, j, k);
                    if (f == -1) continue;
                    if (f >= 1) {
                        g->fln[g->ifln++] = (int)c;
                        int surf = 0;
                        for (int l = 0; l < 3 && !surf; ++l) {
                            int pi = i + E[l][0], pj = j + E[l][1], pk = k + E[l][2];
                            int mi = i - E[l][0], mj = j - E[l][1], mk = k - E[l][2];
                            if (fflag(g, pi, pj, pk) == 0 && aval(g, l, pi, pj, pk) > 0.0) surf = 1;
                            else if (fflag(g, mi, mj, mk) == 0 && g->a[l][c] > 0.0) surf = 1;
                        }
                        if (!surf) {
                            g->nf[c] = 1;
                        } else {
                            g->nf[c] = 2;
                            g->sfn[g->isfn++] = (int)c;
                            g->nb[c] = -3;
                        }
                    }
                }
        /* :108-131, descending */
        int n2 = 0;
        for (int s = g->isfn - 1; s >= 0; --s) {
            long c = g->sfn[s];
            int i = (int)(c % g->X), j = (int)((c / g->X) % g->Y), k = (int)(c / ((long)g->X * g->Y));
            if (fflag(g, i, j, k - 1) == -1) continue;
            int keep = 0;
            for (int l = 0; l < 3 && !keep; ++l) {
                if (fflag(g, i + E[l][0], j + E[l][1], k + E[l][2]) == 1) keep = 1;
                else if (fflag(g, i - E[l][0], j - E[l][1], k - E[l][2]) == 1) keep = 1;
            }
            if (keep) continue;
            n2 = 1;
            g->nf[c] = 0;
            g->nb[c] = 0;
        }
        if (n2) continue;
        /* :133-158, descending */
        for (int s = g->isfn - 1; s >= 0; --s) {
            long c = g->sfn[s];
            int i = (int)(c % g->X), j = (int)((c / g->X) % g->Y), k = (int)(c / ((long)g->X * g->Y));
            if (fflag(g, i, j, k - 1) == -1 && fflag(g, i, j, k + 1) == 0) continue;
            int keep = 0;
            for (int l = 0; l < 3 && !keep; ++l) {
                int fp = fflag(g, i + E[l][0], j + E[l][1], k + E[l][2]);
                int fm = fflag(g, i - E[l][0], j - E[l][1], k - E[l][2]);
                if ((fp == 1 && fm == 0) || (fm == 1 && fp == 0)) keep = 1;
            }
            if (keep) continue;
            n2 = 1;
            g->nf[c] = 0;
            g->nb[c] = 0;
        }
        if (n2) continue;
        /* :163-246, descending */
        int igai = 0;
        for (int s = g->isfn - 1; s >= 0; --s) {
            long c = g->sfn[s];
            int i = (int)(c % g->X), j = (int)((c / g->X) % g->Y), k = (int)(c / ((long)g->X * g->Y));
            g->nb[c] = 0;
            int ix[3] = {0, 0, 0};
            for (int l = 0; l < 3; ++l) {
                int pi = i + E[l][0], pj = j + E[l][1], pk = k + E[l][2];
                int mi = i - E[l][0], mj = j - E[l][1], mk = k - E[l][2];
                if (fflag(g, mi, mj, mk) == 1 && fflag(g, pi, pj, pk) == 0 && aval(g, l, pi, pj, pk) > 0.0)
                    ix[l] = -1;
                else if (fflag(g, mi, mj, mk) == 0 && fflag(g, pi, pj, pk) == 1 && g->a[l][c] > 0.0)
                    ix[l] = 1;
            }
            int sabs = abs(ix[0]) + abs(ix[1]) + abs(ix[2]);
            if (sabs == 0) {
                if (fflag(g, i, j, k - 1) == -1 && fflag(g, i, j, k + 1) == 0) {
                    /* NORMAL + OBJECT (:192-212); objno = 0 for these grids */
                    if (g->fb[c] != 0.0) {
                        ix[2] = -1;
                    } else {
                        for (int l = 0; l < 2; ++l) {
                            int pi = i + E[l][0], pj = j + E[l][1], pk = k + E[l][2];
                            int mi = i - E[l][0], mj = j - E[l][1], mk = k - E[l][2];
                            if (fflag(g, mi, mj, mk) > 0 && fflag(g, pi, pj, pk) == 0 &&
                                aval(g, l, pi, pj, pk) > 0.0)
                                ix[l] = -1;
                            else if (fflag(g, mi, mj, mk) == 0 && fflag(g, pi, pj, pk) > 0 &&
                                     g->a[l][c] > 0.0)
                                ix[l] = 1;
                        }
                        if (abs(ix[0]) + abs(ix[1]) == 0) ix[2] = -1;
                    }
                }
            }
            sabs = abs(ix[0]) + abs(ix[1]) + abs(ix[2]);
            if (sabs == 1) {
                int dir = 0; /* maxloc(abs(ix)), first maximum */
                for (int l = 1; l < 3; ++l) if (abs(ix[l]) > abs(ix[dir])) dir = l;
                g->nb[c] = (int8_t)((dir + 1) * ix[dir]);
            } else if (sabs == 0) {
                igai = igai + 1;  /* 'souteigai' */
                g->nf[c] = 0;
            } else {
                double fx[3];
                int lg[3] = {i, j, k};
                for (int l = 0; l < 3; ++l) fx[l] = ix[l] == 0 ? 0.0 : fxx1(g, ix[l], l, lg);
                int dir = 0; /* maxloc(fxx) */
                for (int l = 1; l < 3; ++l) if (fx[l] > fx[dir]) dir = l;
                g->nb[c] = (int8_t)((dir + 1) * ix[dir]);
            }
        }
        if (igai) goto restart;
        break;
    }
}

/* a20: sp along the surface normal, set_initial_surface_position
 * (1sub-normal-ini.F90:23-40) with fb0 = fb. */
static double surface_pos(const G *g, long c, int i, int j, int k) {
    int nfbb = g->nb[c], anf = abs(nfbb), l = anf - 1;
    int idx[3] = {i, j, k};
    double f = g->F[c], fff;
    if (nfbb == -3 && g->fb[c] < 1.0) fff = 1.0 - g->fb[c] * (1.0 - f);
    else if (nfbb < 0) fff = f;
    else fff = 1.0 - f;
    double sp = g->xf[l][idx[l]] + g->dx[l][idx[l]] * fff;
    if (fabs(sp) < 1e-15) sp = 0.0;
    return sp;
}

/*
 * Builds the PCGPME inputs of one pass.  Returns N = mnoe (>= 0) or -1.
 *   out AL, LL (column stride lda, 6 slots [x-,x+,y-,y+,z-,z+]), D, B
 *   out row2cell[N] box index of each row, mno[ncell] row number (1-based) or 0
 *   out nf, nb per box cell
 * F is copied (the classification clips air cells' F).
 */
int oracle_build_system(int nx, int ny, int nz, const int8_t *solid, const double *F_in,
                        const double *fb, const double *a, const double *un,
                        const double *dx, const double *xf, int dxstride, double dt,
                        int8_t *nf, int8_t *nb, int *row2cell, int *mno, int lda,
                        double *AL, int *LL, double *D, double *B) {
    G g;
    memset(&g, 0, sizeof g);
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.X = nx + 2; g.Y = ny + 2; g.Z = nz + 2;
    g.ncell = (long)g.X * g.Y * g.Z;
    g.solid = solid;
    g.F = (double *)malloc(sizeof(double) * g.ncell);
    g.fln = (int *)malloc(sizeof(int) * g.ncell);
    g.sfn = (int *)malloc(sizeof(int) * g.ncell);
    if (!g.F || !g.fln || !g.sfn) { free(g.F); free(g.fln); free(g.sfn); return -1; }
    memcpy(g.F, F_in, sizeof(double) * g.ncell);
    g.fb = fb;
    for (int l = 0; l < 3; ++l) {
        g.a[l] = a + (size_t)l * g.ncell;
        g.un[l] = un + (size_t)l * g.ncell;
        g.dx[l] = dx + (size_t)l * dxstride;
        g.xf[l] = xf + (size_t)l * dxstride;
    }
    g.nf = nf; g.nb = nb;
    classify(&g);

    /* mp = (min dx>0)^2 (1sub-matrix.F90:17) over the cell sizes */
    double mp = 1e300;
    int nl3[3] = {nx, ny, nz};
    for (int l = 0; l < 3; ++l)
        for (int i = 1; i <= nl3[l]; ++i)
            if (g.dx[l][i] > 0.0 && g.dx[l][i] < mp) mp = g.dx[l][i];
    mp = mp * mp;

    /* row numbering (1sub-momentum.F90:278-296) */
    memset(mno, 0, sizeof(int) * g.ncell);
    int N = 0;
    for (int s = 0; s < g.ifln; ++s) {
        long c = g.fln[s];
        int i = (int)(c % g.X), j = (int)((c / g.X) % g.Y), k = (int)(c / ((long)g.X * g.Y));
        if (g.nf[c] == 0 || g.nf[c] == -1 || g.fb[c] == 0.0) continue;
        if (g.nf[c] == 2 && g.nb[c] != 0) {
            /* b = 0 (no orientation; fxx1 ties at zero) would index lf(-1,:)
             * in the reference; such a cell is numbered without the check */
            int l = abs(g.nb[c]) - 1, sg = g.nb[c] > 0 ? 1 : -1;
            if (fflag(&g, i + sg * E[l][0], j + sg * E[l][1], k + sg * E[l][2]) == -1) continue;
        }
        mno[c] = ++N;
        row2cell[N - 1] = (int)c;
    }
    if (N > lda) { free(g.F); free(g.fln); free(g.sfn); return -1; }

    for (int r = 0; r < N; ++r) {
        long c = row2cell[r];
        int i = (int)(c % g.X), j = (int)((c / g.X) % g.Y), k = (int)(c / ((long)g.X * g.Y));
        int lg[3] = {i, j, k};
        /* init_matrix (1sub-matrix.F90:32-64): ALU_C, LL_C (cell ids), DE_C */
        double alu_c[6], de_c = 0.0;
        long ll_c[6];
        for (int l = 0; l < 3; ++l) {
            for (int s2 = 0; s2 < 2; ++s2) {
                int ii = s2 == 0 ? -1 : 1, slot = 2 * l + s2;
                int ni = i + ii * E[l][0], nj = j + ii * E[l][1], nk = k + ii * E[l][2];
                int pi = i + (ii > 0) * E[l][0], pj = j + (ii > 0) * E[l][1], pk = k + (ii > 0) * E[l][2];
                int fnb = fflag(&g, ni, nj, nk);
                /* an object neighbour closes the slot unless it is a pressure-boundary cell */
                if ((fnb == -1 && bflag(&g, ni, nj, nk) != 4) || aval(&g, l, pi, pj, pk) == 0.0 ||
                    fbval(&g, ni, nj, nk) == 0.0) {
                    ll_c[slot] = -1;
                    alu_c[slot] = 0.0;
                } else {
                    ll_c[slot] = cid(&g, ni, nj, nk);
                    double h = g.dx[l][lg[l]], hm = g.dx[l][lg[l] + ii];
                    double cal_alu_f = 2.0 / h / (h + hm) * aval(&g, l, pi, pj, pk) / g.fb[c];
                    alu_c[slot] = -mp * cal_alu_f;
                    de_c = de_c - alu_c[slot];
                }
            }
        }
        /* cal_matrix_element (1sub-momentum.F90:481-548) */
        int fluid = g.nf[c] == 1;
        double de = fluid ? de_c : 1.0;
        for (int slot = 0; slot < 6; ++slot) {
            double alu = 0.0;
            int llu = 0;
            if (ll_c[slot] >= 0) {
                long nc = ll_c[slot];
                int np = mno[nc];
                if (g.nf[nc] == 0 || np == 0) {
                    if (fluid && g.nf[nc] != 2) de = de + alu_c[slot];
                } else {
                    llu = np;
                    if (fluid) {
                        alu = alu_c[slot];
                    } else {
                        int ll1 = slot + 1, l = (ll1 + ll1 % 2) / 2 - 1;
                        if (abs(g.nb[c]) - 1 == l) {
                            int ii = lg[l], m2 = ll1 % 2;
                            double sp = surface_pos(&g, c, i, j, k);
                            double fss = (double)(1 - 2 * m2) * (g.xf[l][ii - m2 + 1] - sp) / g.dx[l][ii];
                            double h = g.dx[l][ii], hm = g.dx[l][ii - 2 * m2 + 1];
                            alu = (fss - 0.5) * h / (fss * h + hm * 0.5) * (-1.0);
                        } else {
                            alu = -1.0e-20;
                        }
                    }
                }
            }
            AL[(size_t)slot * lda + r] = alu;
            LL[(size_t)slot * lda + r] = llu;
        }
        D[r] = de;
        /* RHS (1sub-momentum.F90:315-346), fluid rows only */
        double b = 0.0;
        if (fluid) {
            double rn;
            if (g.fb[c] != 0.0) {
                double axp = aval(&g, 0, i + 1, j, k), axm = g.a[0][c];
                double ayp = aval(&g, 1, i, j + 1, k), aym = g.a[1][c];
                double azp = aval(&g, 2, i, j, k + 1), azm = g.a[2][c];
                double up = unval(&g, 0, i + 1, j, k), um = g.un[0][c];
                double vp = unval(&g, 1, i, j + 1, k), vm = g.un[1][c];
                double wp = unval(&g, 2, i, j, k + 1), wm = g.un[2][c];
                rn = ((up * axp - um * axm) / g.dx[0][i] + (vp * ayp - vm * aym) / g.dx[1][j] +
                      (wp * azp - wm * azm) / g.dx[2][k]) / g.fb[c];
            } else {
                rn = -10000.0;
            }
            double ff = 1.0 - g.F[c];
            if (ff > 0.001) ff = 0.001;
            if (ff < -0.001) ff = -0.001;
            double dr1 = rn + ff / dt;
            double q = fabs(dr1 / dt);
            if (!(q <= 1e-15 && q != 0.0)) b = mp * dr1 / dt;
        }
        B[r] = b;
    }
    free(g.F); free(g.fln); free(g.sfn);
    return N;
}

/*
 * One SMAC pressure-correction pass after a solve (oracle):
 *   p(mno2mn(r)) -= xe(r)            src2LOWNS/1sub-momentum.F90:53-58
 *   estimate_velocity2               src2LOWNS/1sub-momentum.F90:590-651
 *     un(l,c) = momentum_d(un, -xe(nno(l)), -xe(nnn), dxc, dt)
 *             = un + dt*(pm - pc)/dxc, dxc = (dx(l,i) + dx(l,i-1))/2
 *     (lib/src/math_func.F90:251-255), over the cells of the case's main
 *     numbering that are neither air nor objects other than pressure-boundary
 *     cells:
 *       a pressure-boundary cell (nb = 4): on every axis whose low neighbour
 *         is flagged >= 1 (:611-623);
 *       any other cell: on every axis with a(l,c) != 0 whose low neighbour is
 *         flagged >= 1 (then |un| < 1e-15 -> 0) or is a pressure-boundary
 *         cell (:624-646).
 *   The reference reads xe(0) for an unnumbered cell; PCGPME never writes
 *   it and the caller allocates it without a value, so here it is 0.
 * and the RHS + DMAX of the next pass (1sub-momentum.F90:315-359) for the
 * given row numbering (rows numbered once per time step, ik = 0).
 * mno, row2cell, nf (flags) come from oracle_build_system; un is
 * 3*ncell (low-face velocities), p per cell.
 */
void oracle_velocity_correction(int nx, int ny, int nz, const int8_t *solid, const int8_t *nf,
                                const double *a, const double *dx, int dxstride, double dt,
                                const int *mno, const double *xe, double *un) {
    const int X = nx + 2, Y = ny + 2, Z = nz + 2;
    const long ncell = (long)X * Y * Z;
    const long step[3] = {1, X, (long)X * Y};
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < Y; ++j)
            for (int i = 0; i < X; ++i) {
                long c = i + (long)X * (j + (long)Y * k);
                if (solid[c] == 1 || (!solid[c] && nf[c] == 0)) continue;
                const int pbc = solid[c] == 4;  /* a pressure-boundary cell */
                if (solid[c] && !pbc) continue;
                int idx[3] = {i, j, k};
                double xec = -(mno[c] ? xe[mno[c] - 1] : 0.0);
                for (int l = 0; l < 3; ++l) {
                    if (!pbc && a[(size_t)l * ncell + c] == 0.0) continue;
                    /* the low neighbour; below the box it is the sentinel (f = -1, b = 0) */
                    int fm = -1, bm = 0;
                    long m = -1;
                    if (idx[l] >= 1) {
                        m = c - step[l];
                        fm = solid[m] ? -1 : nf[m];
                        bm = solid[m] == 1 ? 0 : solid[m];
                    }
                    const int corr = fm >= 1 || (!pbc && fm == -1 && bm == 4);
                    if (!corr) continue;
                    double dxc = (dx[(size_t)l * dxstride + idx[l]] + dx[(size_t)l * dxstride + idx[l] - 1]) * 0.5;
                    double pm = -(m >= 0 && mno[m] ? xe[mno[m] - 1] : 0.0);
                    double v = un[(size_t)l * ncell + c] + dt * (pm - xec) / dxc;
                    if (!pbc && fm >= 1 && fabs(v) < 1e-15) v = 0.0;
                    un[(size_t)l * ncell + c] = v;
                }
            }
}

/* RHS b (per row) and DMAX of a pass for a fixed numbering. */
double oracle_rhs(int nx, int ny, int nz, const int8_t *solid, const int8_t *nf, const double *F,
                  const double *fb, const double *a, const double *un, const double *dx, int dxstride,
                  double dt, double mp, int N, const int *row2cell, double *B) {
    const int X = nx + 2, Y = ny + 2;
    const long ncell = (long)X * Y * (nz + 2);
    double dmax = 0.0;
    for (int r = 0; r < N; ++r) {
        long c = row2cell[r];
        B[r] = 0.0;
        if (nf[c] != 1) continue;
        int i = (int)(c % X), j = (int)((c / X) % Y), k = (int)(c / ((long)X * Y));
        double rn;
        if (fb[c] != 0.0) {
            long cx = c + 1, cy = c + X, cz = c + (long)X * Y;
            /* a cell outside the case reads the sentinel's 0 */
            double axp = solid[cx] == 1 ? 0.0 : a[cx], ayp = solid[cy] == 1 ? 0.0 : a[ncell + cy];
            double azp = solid[cz] == 1 ? 0.0 : a[2 * ncell + cz];
            double up = solid[cx] == 1 ? 0.0 : un[cx], vp = solid[cy] == 1 ? 0.0 : un[ncell + cy];
            double wp = solid[cz] == 1 ? 0.0 : un[2 * ncell + cz];
            rn = ((up * axp - un[c] * a[c]) / dx[i] + (vp * ayp - un[ncell + c] * a[ncell + c]) / dx[dxstride + j] +
                  (wp * azp - un[2 * ncell + c] * a[2 * ncell + c]) / dx[2 * dxstride + k]) / fb[c];
        } else {
            rn = -10000.0;
        }
        double ff = 1.0 - F[c];
        if (ff > 0.001) ff = 0.001;
        if (ff < -0.001) ff = -0.001;
        double dr1 = rn + ff / dt;
        double q = fabs(dr1 / dt);
        if (!(q <= 1e-15 && q != 0.0)) B[r] = mp * dr1 / dt;
        if (fabs(dr1) > dmax) dmax = fabs(dr1);
    }
    return dmax;
}

/* round(num, dec), lib/src/xy_calc.f90:101-112: dfloat(nint(num*10**dec))
 * / 10**dec, nint rounding half away from zero. */
static double ref_round(double num, int dec) {
    double A = pow(10.0, (double)dec);
    double v = num * A;
    double t = trunc(v);
    double r = t;
    if (fabs(v - t) >= 0.5) r = t + (v > 0 ? 1.0 : -1.0);
    return r / A;
}

/* set_dx for one axis, src2LOWNS/1sub-main.F90:53-74.  x[0..m) holds the
 * face coordinates x(ls-2 .. le+1) (m = le-ls+4 >= 4), dx[0..m) receives
 * dx(ls-2 .. le+1).  x(le+1) = 0 is extrapolated first (:60-62); the number
 * of decimals is the first dec in 0..10 with (x(ls+1)-x(ls))/10**(4-dec) > 1
 * (:63-66; 11 if none); dx(i) = round(x(i+1)-x(i), dec) for i = ls-2..le and
 * dx(le+1) = dx(le) (:67-71).  Returns 0, or -1 if m < 4. */
int oracle_set_dx(double *x, int m, double *dx) {
    if (m < 4) return -1;
    const int ls = 2, le = m - 2; /* array positions of x(ls), x(le) */
    if (x[le + 1] == 0.0) x[le + 1] = x[le] + x[le] - x[le - 1];
    int dec;
    for (dec = 0; dec <= 10; ++dec)
        if ((x[ls + 1] - x[ls]) / pow(10.0, (double)(4 - dec)) > 1.0) break;
    for (int i = 0; i <= le; ++i) dx[i] = ref_round(x[i + 1] - x[i], dec);
    dx[le + 1] = dx[le];
    return 0;
}
