/* oracle.h -- TEST INFRASTRUCTURE ONLY: declarations of the C oracle. */
#ifndef CLOWNS_ORACLE_H
#define CLOWNS_ORACLE_H
#include <stdint.h>
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
void oracle_will_filter(const int *LL, int NL, int NA, int N, double *X,
                        double *B);
int oracle_pcgpme_will(const double *AL, int NL, int NA, const int *LL,
                       const double *D, int N, int N2, double *B, double EPS,
                       double PARM, int ITER, int *IER, int *K_out,
                       double *X_raw, double *X_out);
int oracle_set_dx(double *x, int m, double *dx);
int oracle_vof(int nx, int ny, int nz, int stride, const int8_t *solid, const int8_t *nf, const int8_t *nb,
               const double *F, const double *Fp, const double *fb, const double *a, const double *un,
               const double *dx, double *dt, int mode, const double *qs, const int8_t *qs_mask, int n_plic,
               const int32_t *plic_cells, const double *plic_vol, double *qf, double *Fn, int *retries,
               int *counts);
#endif
