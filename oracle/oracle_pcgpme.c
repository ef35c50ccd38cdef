/*
 * oracle_pcgpme.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C CPU restatement of the 2CLOWNS-3D pressure solver PCGPME
 * (src2LOWNS/pcgpme_para.F90:1-354): row scaling with ghost indices, per-row
 * exchange sort, the five-stage product-form ("PM") preconditioner LDUSUB,
 * the 7-point product AXSUB, and the preconditioned Bi-CGSTAB iteration with
 * the reference's exits.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this file's library, and only as the checker.
 *
 * Conventions kept from the reference:
 *   AL(NA,NL), LL(NA,NL) column major, LL 1-based (0 = no neighbour);
 *   D, B length N; X length N+N2 (N2 >= 64, padding entries are zero);
 *   every floating-point expression is evaluated left to right as written in
 *   the Fortran, without fused multiply-add (the library is compiled with
 *   -ffp-contract=off, and so is the flang build of the reference it is
 *   compared against).
 * Reductions: each OpenMP thread owns one contiguous chunk
 * (schedule(static,(N-1)/NTHR+1), pcgpme_para.F90:60 and others) and the
 * per-chunk partial sums are combined in chunk order, so the result is
 * deterministic for a given thread count and equals the reference's serial
 * sum at one thread.
 *
 * Pinning (see tests/test_oracle_pcgpme.py): bit-exact against the flang
 * build of pcgpme_para.F90 at one thread (ITER = 1, 2, 5 and converged),
 * against dense numpy algebra for AXSUB and LDUSUB, against
 * numpy.linalg.solve for the converged X, and against the uniform-grid
 * known answer for the scaled/sorted coefficients.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static int g_nthr = 1;

void oracle_set_threads(int n) { g_nthr = n > 0 ? n : 1; }
int oracle_get_threads(void) { return g_nthr; }

#define A2(arr, i, k, NA) (arr)[(size_t)(k) * (size_t)(NA) + (size_t)(i)]

/* Chunk [lo,hi) of thread t for schedule(static,(n-1)/NTHR+1). */
static void chunk_of(int n, int t, int *lo, int *hi) {
    int c = (n - 1) / g_nthr + 1;
    long l = (long)t * c, h = l + c;
    if (l > n) l = n;
    if (h > n) h = n;
    *lo = (int)l;
    *hi = (int)h;
}

/* a3: row scaling with ghost-index substitution, pcgpme_para.F90:77-94.
 * a4: per-row exchange sort by |AL|, pcgpme_para.F90:96-112. */
void oracle_scale_sort(double *AL, int NL, int NA, int *LL, double *D, int N,
                       double *B, double PARM) {
    double P1 = PARM, P2;
    if (PARM < 1.0) P1 = 2.0 * 2.0;
    P2 = P1 + 1.0;
    int N1 = N + 1;
#pragma omp parallel num_threads(g_nthr)
    {
        int lo, hi;
        chunk_of(N, omp_get_thread_num(), &lo, &hi);
        for (int i = lo; i < hi; ++i) {
            int I = i + 1; /* Fortran row number */
            double w = 0.0;
            for (int k = 0; k < NL; ++k) {
                if (A2(LL, i, k, NA) == 0) A2(LL, i, k, NA) = N1 + I % 64;
                w = w + fabs(A2(AL, i, k, NA));
            }
            double temp = D[i];
            double m = (w > temp) ? w : temp; /* Fortran max(W,TEMP) */
            w = P2 / (m + temp * P1);
            for (int k = 0; k < NL; ++k) A2(AL, i, k, NA) = A2(AL, i, k, NA) * w;
            D[i] = D[i] * w;
            B[i] = B[i] * w;
        }
        for (int i = lo; i < hi; ++i) {
            for (int k = 0; k < NL - 1; ++k) {
                for (int kk = k + 1; kk < NL; ++kk) {
                    if (fabs(A2(AL, i, k, NA)) < fabs(A2(AL, i, kk, NA))) {
                        double t = A2(AL, i, k, NA);
                        int it = A2(LL, i, k, NA);
                        A2(AL, i, k, NA) = A2(AL, i, kk, NA);
                        A2(LL, i, k, NA) = A2(LL, i, kk, NA);
                        A2(AL, i, kk, NA) = t;
                        A2(LL, i, kk, NA) = it;
                    }
                }
            }
        }
    }
}

/* a6: AXSUB, Y = A*X in sorted slot order, pcgpme_para.F90:323-352.
 * LL must already be padded (no zeros).  X has N+64 entries at least.
 * Returns 0, or -1 for an unsupported NL (the reference STOPs there). */
int oracle_axsub(const double *AL, int NL, int NA, const int *LL,
                 const double *D, int N, const double *X, double *Y) {
    if (NL != 6 && NL != 4) return -1;
#pragma omp parallel num_threads(g_nthr)
    {
        int lo, hi;
        chunk_of(N, omp_get_thread_num(), &lo, &hi);
        if (NL == 6) {
            for (int i = lo; i < hi; ++i) {
                double y = X[i] * D[i];
                y = y + A2(AL, i, 0, NA) * X[A2(LL, i, 0, NA) - 1];
                y = y + A2(AL, i, 1, NA) * X[A2(LL, i, 1, NA) - 1];
                y = y + A2(AL, i, 2, NA) * X[A2(LL, i, 2, NA) - 1];
                y = y + A2(AL, i, 3, NA) * X[A2(LL, i, 3, NA) - 1];
                y = y + A2(AL, i, 4, NA) * X[A2(LL, i, 4, NA) - 1];
                y = y + A2(AL, i, 5, NA) * X[A2(LL, i, 5, NA) - 1];
                Y[i] = y;
            }
        } else {
            for (int i = lo; i < hi; ++i) {
                double y = X[i] * D[i];
                y = y + A2(AL, i, 0, NA) * X[A2(LL, i, 0, NA) - 1];
                y = y + A2(AL, i, 1, NA) * X[A2(LL, i, 1, NA) - 1];
                y = y + A2(AL, i, 2, NA) * X[A2(LL, i, 2, NA) - 1];
                y = y + A2(AL, i, 3, NA) * X[A2(LL, i, 3, NA) - 1];
                Y[i] = y;
            }
        }
    }
    return 0;
}

/* a7: LDUSUB, Q = PM*P, pcgpme_para.F90:263-320.  P and Q may alias (the
 * reference calls it in place); W is a scratch vector of N+64 entries whose
 * padding is zero on entry.  Stage j reads the previous stage's output at the
 * row's j-th largest off-diagonal. */
void oracle_ldusub(const double *AL, int NL, int NA, const int *LL, int N,
                   const double *P, double *Q, double *W) {
#pragma omp parallel num_threads(g_nthr)
    {
        int lo, hi;
        chunk_of(N, omp_get_thread_num(), &lo, &hi);
        for (int i = lo; i < hi; ++i)
            W[i] = P[i] - A2(AL, i, 0, NA) * P[A2(LL, i, 0, NA) - 1];
#pragma omp barrier
        for (int i = lo; i < hi; ++i)
            Q[i] = W[i] - A2(AL, i, 1, NA) * W[A2(LL, i, 1, NA) - 1];
#pragma omp barrier
        for (int i = lo; i < hi; ++i)
            W[i] = Q[i] - A2(AL, i, 2, NA) * Q[A2(LL, i, 2, NA) - 1];
#pragma omp barrier
        for (int i = lo; i < hi; ++i)
            Q[i] = W[i] - A2(AL, i, 3, NA) * W[A2(LL, i, 3, NA) - 1];
#pragma omp barrier
        if (NL > 4) {
            for (int i = lo; i < hi; ++i)
                W[i] = Q[i] - A2(AL, i, 4, NA) * Q[A2(LL, i, 4, NA) - 1] -
                       A2(AL, i, 5, NA) * Q[A2(LL, i, 5, NA) - 1];
#pragma omp barrier
            for (int i = lo; i < hi; ++i) Q[i] = W[i];
        }
    }
}

/* Deterministic chunk-ordered combination of per-thread partial sums. */
#define MAXT 1024

/* The preconditioned iteration shared by PCGPME and PCGPME_WILL on an
 * already scaled and sorted system: setup pcgpme_para.F90:115-145
 * (pcgpme_2DH.f90:121-150), iteration :148-247 (:153-247).  The two
 * differ only in BETA's rounding: beta_prod = 0 gives C1/AMU/RQ
 * (pcgpme_para.F90:241), 1 gives C1/(AMU*RQ) (pcgpme_2DH.f90:241).
 * itermax <= 0 runs no iteration.  Returns 0, or -2 if out of memory. */
static int pcgpme_iterate(const double *AL, int NL, int NA, const int *LL,
                          const double *D, int N, int N2, double *B, double *X,
                          double eps, int itermax, int beta_prod, int *K_out,
                          double *err_out, int *ier_out) {
    const double ZERO = 0.0;
    int IER = 0;
    double ERRP = ZERO, ERR = ZERO;
    size_t L = (size_t)N + (size_t)N2;
    double *WK = (double *)calloc(L * 5, sizeof(double));
    double *W = (double *)calloc(L, sizeof(double));
    if (!WK || !W) {
        free(WK);
        free(W);
        return -2;
    }
    double *WK1 = WK, *WK2 = WK + L, *WK3 = WK + 2 * L, *WK4 = WK + 3 * L,
           *WK5 = WK + 4 * L;

    for (int i = 0; i < N; ++i) WK1[i] = B[i];

    oracle_ldusub(AL, NL, NA, LL, N, WK1, WK1, W);   /* B = PM*B   :123 */
    oracle_axsub(AL, NL, NA, LL, D, N, X, WK3);      /* Q = A*X    :126 */
    oracle_ldusub(AL, NL, NA, LL, N, WK3, WK3, W);   /* Q = PM*Q   :127 */

    double part[MAXT][3];
    int nt = g_nthr;
    double BN = ZERO, C1 = ZERO;
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num(), lo, hi;
        chunk_of(N, t, &lo, &hi);
        double bn = ZERO, c1 = ZERO;
        for (int i = lo; i < hi; ++i) {                      /* :133-142 */
            double temp = WK1[i];
            bn = bn + temp * temp;
            temp = temp - WK3[i];
            B[i] = temp;
            c1 = c1 + temp * temp;
            WK1[i] = temp;
            WK2[i] = temp;
        }
        part[t][0] = bn;
        part[t][1] = c1;
    }
    for (int t = 0; t < nt; ++t) {
        BN = BN + part[t][0];
        C1 = C1 + part[t][1];
    }

    int K;
    for (K = 1; K <= itermax; ++K) {
        oracle_axsub(AL, NL, NA, LL, D, N, WK1, WK3);  /* Q = PM*A*P :150 */
        oracle_ldusub(AL, NL, NA, LL, N, WK3, WK3, W);
        double RQ = ZERO;
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num(), lo, hi;
            chunk_of(N, t, &lo, &hi);
            double rq = ZERO;
            for (int i = lo; i < hi; ++i) rq = rq + WK2[i] * WK3[i];
            part[t][0] = rq;
        }
        for (int t = 0; t < nt; ++t) RQ = RQ + part[t][0];
        double ALP = C1 / RQ;                                  /* :167 */
#pragma omp parallel num_threads(nt)
        {
            int lo, hi;
            chunk_of(N, omp_get_thread_num(), &lo, &hi);
            for (int i = lo; i < hi; ++i) WK4[i] = B[i] - ALP * WK3[i]; /* :173 */
        }
        oracle_axsub(AL, NL, NA, LL, D, N, WK4, WK5);  /* V = PM*A*E :179 */
        oracle_ldusub(AL, NL, NA, LL, N, WK5, WK5, W);
        double AMU = ZERO, SS = ZERO;
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num(), lo, hi;
            chunk_of(N, t, &lo, &hi);
            double amu = ZERO, ss = ZERO;
            for (int i = lo; i < hi; ++i) {                    /* :187-191 */
                double temp = WK5[i];
                amu = amu + WK4[i] * temp;
                ss = ss + temp * temp;
            }
            part[t][0] = amu;
            part[t][1] = ss;
        }
        for (int t = 0; t < nt; ++t) {
            AMU = AMU + part[t][0];
            SS = SS + part[t][1];
        }
        AMU = AMU / SS;                                         /* :197 */
        double RN = ZERO;
        C1 = ZERO;
#pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num(), lo, hi;
            chunk_of(N, t, &lo, &hi);
            double rn = ZERO, c1 = ZERO;
            for (int i = lo; i < hi; ++i) {                    /* :204-211 */
                double temp = WK4[i];
                X[i] = X[i] + ALP * WK1[i] + AMU * temp;
                temp = temp - AMU * WK5[i];
                B[i] = temp;
                rn = rn + temp * temp;
                c1 = c1 + temp * WK2[i];
            }
            part[t][0] = rn;
            part[t][1] = c1;
        }
        for (int t = 0; t < nt; ++t) {
            RN = RN + part[t][0];
            C1 = C1 + part[t][1];
        }
        ERR = sqrt(RN / BN);                                    /* :215 */
        if (K >= 10 && fabs(ERR - ERRP) <= ERR * 1.0e-6) goto done; /* :228 */
        if (ERR <= eps) goto done;                              /* :232 */
        ERRP = ERR;
        if (AMU == ZERO) goto done;                             /* :234 */
        if (C1 == ZERO) goto done;                              /* :240 */
        double BETA = beta_prod ? C1 / (AMU * RQ) : C1 / AMU / RQ; /* :241 */
#pragma omp parallel num_threads(nt)
        {
            int lo, hi;
            chunk_of(N, omp_get_thread_num(), &lo, &hi);
            for (int i = lo; i < hi; ++i)                       /* :244 */
                WK1[i] = B[i] + BETA * (WK1[i] - AMU * WK3[i]);
        }
    }
    IER = 3000;                                                 /* :251 */
done:
    *K_out = K;
    *err_out = ERR;
    *ier_out = IER;
    free(WK);
    free(W);
    return 0;
}

/* PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER), pcgpme_para.F90:1-257.
 * Same arguments (by value for scalars), same in-place mutation of AL, LL, D
 * and B, same exits.  Returns 0, or -1 if NL is not 4 or 6 (reference: STOP)
 * or N2 < 64 (the reference would index past X). */
int oracle_pcgpme(double *AL, int NL, int NA, int *LL, double *D, int N,
                  int N2, double *B, double *X, double *EPS, double PARM,
                  int *ITER, int *IER) {
    if (NL != 6 && NL != 4) return -1;
    if (N2 < 64 || g_nthr > MAXT) return -1;
    if (EPS[0] <= 1e-15) EPS[0] = 1e-9;                           /* :51 */
    if (*ITER <= 0) *ITER = (int)(10.0 * (sqrt((double)N) + 1.0)); /* :52 */
    oracle_scale_sort(AL, NL, NA, LL, D, N, B, PARM);
    int K = 0, ier = 0;
    double err = 0.0;
    int rc = pcgpme_iterate(AL, NL, NA, LL, D, N, N2, B, X, EPS[0], *ITER, 0,
                            &K, &err, &ier);
    if (rc) return rc;
    *IER = ier;
    *ITER = K;                                                  /* :254 */
    EPS[0] = err;
    return 0;
}

/* The output pass of PCGPME_WILL, pcgpme_2DH.f90:255-268, serial over i in
 * the caller's LL (not the sorted copy): zero |X(i)| <= 1e-14, then for each
 * neighbour in slot order -- skip itself, stop at the first 0, skip an equal
 * value, snap X(i) to a value within 1e-14 -- and store X(i) into B(i). */
void oracle_will_filter(const int *LL, int NL, int NA, int N, double *X,
                        double *B) {
    const double L_Limit = 1e-14;                               /* :47 */
    for (int i = 0; i < N; ++i) {
        if (fabs(X[i]) <= L_Limit) X[i] = 0.0;
        for (int k = 0; k < NL; ++k) {
            int j = A2(LL, i, k, NA);
            if (j == i + 1) continue;
            if (j == 0) break;
            if (X[i] == X[j - 1]) continue;
            if (fabs(X[i] - X[j - 1]) <= L_Limit) X[i] = X[j - 1];
            B[i] = X[i];
        }
    }
}

/* PCGPME_WILL(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER), pcgpme_2DH.f90:1-374:
 * AL, LL, D untouched (scaled and sorted copies, :56-76), X from zero, EPS
 * and ITER as given, BETA = C1/(AMU*RQ), then the output pass into B.
 * Returns 0 / -1 (bad NL or N2) / -2 (memory).  Not in the reference's
 * argument list, for the checks: *IER the local IER, *K_out the last K,
 * X_raw / X_out (N entries each, may be NULL) X before and after the
 * output pass. */
int oracle_pcgpme_will(const double *AL, int NL, int NA, const int *LL,
                       const double *D, int N, int N2, double *B, double EPS,
                       double PARM, int ITER, int *IER, int *K_out,
                       double *X_raw, double *X_out) {
    if (NL != 6 && NL != 4) return -1;
    if (N2 < 64 || g_nthr > MAXT || NA < N) return -1;
    size_t nm = (size_t)NA * (size_t)NL;
    double *ALT = (double *)malloc(nm * sizeof(double));
    int *LLT = (int *)malloc(nm * sizeof(int));
    double *DT = (double *)malloc((size_t)N * sizeof(double));
    double *X = (double *)calloc((size_t)N + (size_t)N2, sizeof(double));
    if (!ALT || !LLT || !DT || !X) {
        free(ALT); free(LLT); free(DT); free(X);
        return -2;
    }
    memcpy(ALT, AL, nm * sizeof(double));
    memcpy(LLT, LL, nm * sizeof(int));
    memcpy(DT, D, (size_t)N * sizeof(double));
    oracle_scale_sort(ALT, NL, NA, LLT, DT, N, B, PARM);
    int K = 0, ier = 0;
    double err = 0.0;
    int rc = pcgpme_iterate(ALT, NL, NA, LLT, DT, N, N2, B, X, EPS, ITER, 1,
                            &K, &err, &ier);
    if (!rc) {
        if (X_raw) memcpy(X_raw, X, (size_t)N * sizeof(double));
        oracle_will_filter(LL, NL, NA, N, X, B);
        if (IER) *IER = ier;
        if (K_out) *K_out = K;
        if (X_out) memcpy(X_out, X, (size_t)N * sizeof(double));
    }
    free(ALT); free(LLT); free(DT); free(X);
    return rc;
}

/* Fortran-mangled entry with the reference's by-reference argument list. */
int oracle_pcgpme_ref_args(double *AL, const int *NL, const int *NA, int *LL,
                           double *D, const int *N, const int *N2, double *B,
                           double *X, double *EPS, const double *PARM,
                           int *ITER, int *IER) {
    return oracle_pcgpme(AL, *NL, *NA, LL, D, *N, *N2, B, X, EPS, *PARM, ITER,
                         IER);
}
