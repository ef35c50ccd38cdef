/* oracle.h -- TEST INFRASTRUCTURE ONLY: declarations of the C oracle. */
#ifndef CLOWNS_ORACLE_H
#define CLOWNS_ORACLE_H
void oracle_set_threads(int n);
int oracle_get_threads(void);
void oracle_scale_sort(double *AL, int NL, int NA, int *LL, double *D, int N,
                       double *B, double PARM);
int oracle_axsub(const double *AL, int NL, int NA, const int *LL,
                 const double *D, int N, const double *X, double *Y);
void oracle_ldusub(const double *AL, int NL, int NA, const int *LL, int N,
                   const double *P, double *Q, double *W);
int oracle_pcgpme(double *AL, int NL, int NA, int *LL, double *D, int N,
                  int N2, double *B, double *X, double *EPS, double PARM,
                  int *ITER, int *IER);
#endif
