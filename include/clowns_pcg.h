/*
 * clowns_pcg.h -- C-ABI of the MI355X (gfx950) pressure-Poisson solver that
 * replaces 2CLOWNS-3D's PCGPME.
 *
 * Two layers:
 *  1. clowns_pcgpme(): the drop-in.  Same argument list, layouts, in/out
 *     semantics and IER codes as the reference's external subroutine
 *     PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)
 *     (src2LOWNS/pcgpme_para.F90:1-58, 250-257), every argument passed by
 *     reference so that a Fortran ISO_C_BINDING interface binds it directly.
 *     Host arrays in, host arrays out; synchronous.
 *  2. clowns_solver_*(): the device-resident form of the same solve (matrix
 *     scaled and sorted once and kept on the GPU across pressure-correction
 *     passes and time steps, vectors already in HBM), the mixed-precision
 *     solve and the adjoint (transpose) solve.  Pointers named d_* are device
 *     pointers.
 *
 * Return values of int functions: 0 = success, otherwise a CLOWNS_E* code.
 * All functions are synchronous on return unless stated otherwise.
 */
#ifndef CLOWNS_PCG_H
#define CLOWNS_PCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- IER / return codes ------------------------------------------------
 * 0 and 3000 are the reference's (pcgpme_para.F90:53,251).  The others are
 * new: the reference STOPs on a bad NL (pcgpme_para.F90:347) and runs on
 * with NaN where these return early. */
#define CLOWNS_OK 0
#define CLOWNS_NOCONV 3000      /* ITER exhausted (reference IER=3000)      */
#define CLOWNS_ENONFINITE 4001  /* residual became NaN/Inf; early exit       */
#define CLOWNS_EBADNL 4002      /* NL not 4 or 6                             */
#define CLOWNS_EBADARG 4003     /* N<1, NA<N, N2<64, null pointer, bad flag  */
#define CLOWNS_EHIP 4004        /* HIP runtime error / no device             */
#define CLOWNS_ENOMEM 4005      /* device allocation failed                  */
#define CLOWNS_ESTRUCT 4006     /* transpose: pattern not structurally symmetric */
#define CLOWNS_EDECK 4007       /* unreadable deck file, or cells outside the box model */
#define CLOWNS_EVOF 4008        /* VOF update: the reference would STOP (F > F_lim on a
                                   surface cell, or dt below DT_lim) */

/* Replaces PCGPME (src2LOWNS/pcgpme_para.F90:1, argument declarations
 * :35-46).  AL(NA,NL), LL(NA,NL) column-major, LL 1-based (0 = none);
 * D, B length N; X length N+N2 (entries past N are treated as zero and never
 * read or written); EPS[0] tolerance in / achieved ERR out; ITER max in /
 * iterations out; IER out.  As in the reference, AL and D return scaled, LL
 * sorted with ghost indices N+1+mod(I,64), B the final scaled residual. */
void clowns_pcgpme(double *AL, const int32_t *NL, const int32_t *NA,
                   int32_t *LL, double *D, const int32_t *N, const int32_t *N2,
                   double *B, double *X, double *EPS, const double *PARM,
                   int32_t *ITER, int32_t *IER);

/* ---- device-resident solver ------------------------------------------- */
typedef struct clowns_solver clowns_solver;

typedef struct clowns_info {
    int32_t iters;        /* Bi-CGSTAB iterations (FP64), or inner total (mixed) */
    int32_t ier;          /* CLOWNS_OK / CLOWNS_NOCONV / CLOWNS_ENONFINITE       */
    int32_t refinements;  /* mixed: FP64 refinement steps                        */
    int32_t reserved;
    double err;           /* final sqrt(RN/BN), the reference's stopping value  */
    double ms;            /* device time of the solve (HIP events), ms          */
    double setup_ms;      /* part of ms spent before the first iteration        */
} clowns_info;

/* Solve flags */
#define CLOWNS_MIXED 1       /* FP32 inner PM-Bi-CGSTAB + FP64 refinement   */
#define CLOWNS_WARM_START 2  /* d_x holds the initial guess (else zero)      */

/* Create a solver bound to `device` (-1 = current).  `stream` is the
 * hipStream_t all its work goes to; NULL is the device's null stream (so the
 * solver orders with the caller's default-stream work, e.g. torch's). */
int clowns_solver_create(clowns_solver **out, int device, void *stream);
int clowns_solver_destroy(clowns_solver *h);
int clowns_solver_set_stream(clowns_solver *h, void *stream);

/* Load the operator from device arrays in the reference layout (AL/LL with
 * column stride lda >= N, LL 1-based, D length N).  Scales and sorts a copy
 * on the device (pcgpme_para.F90:77-112); the caller's arrays are not
 * modified.  The scaled/sorted matrix stays resident until the next call. */
int clowns_solver_set_matrix(clowns_solver *h, const double *d_AL, int32_t NL,
                             int32_t lda, const int32_t *d_LL,
                             const double *d_D, int32_t N, double parm);

/* Solve A x = b for the resident operator to the reference's stopping rule
 * sqrt(RN/BN) <= eps (pcgpme_para.F90:215-232).  d_b unscaled, length N;
 * d_x length N (in: initial guess when CLOWNS_WARM_START).  maxiter <= 0
 * gives the reference default 10*(sqrt(N)+1). */
int clowns_solver_solve(clowns_solver *h, const double *d_b, double *d_x,
                        double eps, int32_t maxiter, int32_t flags,
                        clowns_info *info);

/* clowns_solver_solve without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * time-stepping loops, harnesses that forbid it).  No counterpart in the
 * reference.  Same arguments and results as clowns_solver_solve (bit for
 * bit); the run's work vectors, scaled right-hand side, dot partials and
 * scalars live in a file-scope device array of the library, 58 MiB, which
 * every GPU the library runs on holds until the process ends.  Limits: at
 * most 524288 rows (CLOWNS_EBADARG beyond, with a message on stderr); the
 * operator comes from clowns_solver_set_matrix / clowns_solver_assemble as
 * usual (they allocate), and CLOWNS_MIXED needs the operator's FP32 copies,
 * which one clowns_solver_solve with CLOWNS_MIXED builds.  All calls on one
 * GPU share the array: calls that overlap on one GPU (from several threads
 * or streams) race. */
int clowns_solver_solve_noalloc(clowns_solver *h, const double *d_b, double *d_x,
                                double eps, int32_t maxiter, int32_t flags,
                                clowns_info *info);

/* Copy of the resident scaled/sorted operator (reference post-call layout,
 * column stride lda of set_matrix), the row scales W, and the final scaled
 * residual of the last solve.  Any pointer may be NULL. */
int clowns_solver_get_scaled(clowns_solver *h, double *d_AL, int32_t *d_LL,
                             double *d_D, double *d_W, double *d_B);

/* Adjoint: solve A^T lambda = g for the resident operator's transpose
 * (the operator is non-symmetric; its pattern must be structurally
 * symmetric), with the same PM-Bi-CGSTAB and stopping rule.  d_g, d_lambda
 * length N. */
int clowns_solver_solve_transpose(clowns_solver *h, const double *d_g,
                                  double *d_lambda, double eps,
                                  int32_t maxiter, int32_t flags,
                                  clowns_info *info);

/* Gradients of a loss through x = A^{-1} b, given lambda = A^{-T} dL/dx and
 * the solution x, in the caller's (unscaled, unsorted) layout of the operator
 * last passed to set_matrix:
 *   dL/db = lambda;  dL/dAL(i,k) = -lambda_i * x(LL(i,k));  dL/dD = -lambda_i x_i.
 * d_gAL uses column stride lda; entries with LL = 0 get 0. */
int clowns_grad_coefficients(clowns_solver *h, const double *d_lambda,
                             const double *d_x, double *d_gAL, double *d_gD);

/* Time `reps` back-to-back launches of one solver kernel over the resident
 * operator with HIP events on the solver's stream: kernel 0 = FP64 SpMV,
 * 1 = FP64 PM stage, 2 = FP32 SpMV, 3 = FP32 PM stage, 4 = plain FP64 copy
 * of N entries (reference for counter scaling), 5 = FP64 fused PM stage
 * pair, 6 = FP32 fused PM stage pair.  *avg_ms out. */
int clowns_solver_time_kernel(clowns_solver *h, int32_t kernel, int32_t reps,
                              double *avg_ms);

/* In-solve kernel durations.  enable = 1 arms the solver: every kernel
 * launch of the solves that follow is bracketed by HIP events on the
 * solver's stream.  enable = 0 disarms it and returns, per kernel class, the
 * summed device time ms[c] and the number count[c] of the launches that did
 * work (launches a run queued past its last iteration return at once and are
 * left out).  Classes c = 0..8 FP64, 9..17 FP32 (the mixed inner solve), in
 * this order: SpMV, single PM stage, fused PM stage pair, two-term or
 * dot-carrying PM stage, p update, e update, x/r update with its dots,
 * one-block finalize, true-residual SpMV.  ms and count hold 18 entries (or
 * are NULL).  No counterpart in the reference (pcgpme_para.F90 has no
 * timing): it measures the kernels behind the benchmark's roofline figure. */
#define CLOWNS_KPROF_CLASSES 18
int clowns_solver_kernel_profile(clowns_solver *h, int32_t enable, double *ms,
                                 int32_t *count);

/* Row patterns of the resident operator: rows whose six scaled, sorted
 * coefficients and diagonal repeat one of up to 16 table rows bit for bit,
 * for which the SpMV reads a one-byte pattern id instead of the
 * coefficients (results unchanged).  *npat table rows and
 * *rows rows on them; both 0 when the operator does not use them.  No
 * counterpart in the reference (a storage form of the operator). */
int clowns_solver_patterns(clowns_solver *h, int32_t *npat, int32_t *rows);

/* ---- grid geometry and assembly --------------------------------------
 * Box convention: cells i=0..nx+1, j=0..ny+1, k=0..nz+1 whose outer ring is
 * solid (closed walls); box index c = i + X*(j + Y*k), X = nx+2, Y = ny+2,
 * which is the reference's k-major row order (src2LOWNS/nfd2015.F90:69-78).
 * Per-cell fields: solid (int8: 0 a flow cell; otherwise a cell of flag -1:
 * 1 a cell outside the case, which reads as the reference's sentinel cell 0
 * with zero fields; any other value an object cell of the case, numbered by
 * the case reader (src2LOWNS/data_io.F90:171-199) and read with its own
 * fields, whose object code it is: for a boundary object with nfb = b
 * (1..199) b for 2..126 (4: a pressure-boundary cell), -2 for b = 1,
 * 124 - b for b >= 127; -1 / -128 the object cells appended to the numbering
 * with nfb 300 / 200; -127 a cell with nfb = -1), F (VOF fill), fb (cell
 * open ratio), a[3][ncell] (open ratio of the cell's LOW face per axis),
 * un[3][ncell] (velocity on that face); per-axis dx[3][stride] (cell sizes)
 * and xf[3][stride] (low-face coordinates), index = cell index along the
 * axis.  The classification gives a solid cell nf = -1 and nb = its object
 * code (0 outside the case). */
typedef struct clowns_grid_desc {
    int32_t nx, ny, nz;
    int32_t stride;     /* >= max(nx,ny,nz) + 2 */
    double dt;          /* time step of the RHS (1sub-momentum.F90:334-345) */
} clowns_grid_desc;

/* Replaces set_nflag (src2LOWNS/nfd2015.F90:7-306, serial semantics), the
 * ik==0 row numbering of cal_err_of_continuity (1sub-momentum.F90:273-296)
 * and the surface position of set_initial_surface_position
 * (1sub-normal-ini.F90:23-40).  Host arrays.  Outputs per cell nf (-1 solid,
 * 0 air, 1 fluid, 2 surface), nb (surface orientation +-1..3), mno (row,
 * 1-based, or 0) (nb of a solid cell: its object code, 0 outside the case);
 * per row row2cell (box index) and sp_row (surface position
 * along the normal for surface rows, else 0); nrows = mnoe; mp = (min dx)^2
 * (1sub-matrix.F90:17).  F is modified as the reference modifies it (air
 * cells clipped to >= 0).  cap = capacity of row2cell/sp_row. */
int clowns_classify(const clowns_grid_desc *g, const int8_t *solid, double *F,
                    const double *fb, const double *a, const double *dx,
                    const double *xf, int8_t *nf, int8_t *nb, int32_t *mno,
                    int32_t *row2cell, double *sp_row, int32_t cap,
                    int32_t *nrows, double *mp);

/* Device form of clowns_classify on the solver's stream (all arrays on the
 * GPU, same meaning; F modified in place as the reference does).  The
 * order-independent passes of set_nflag run one thread per cell and the row
 * numbering is a prefix sum in box order; if one of set_nflag's
 * order-dependent sweeps (nfd2015.F90:133-153, 167-243) would change a flag,
 * the exact serial pass runs on the host instead (*host_fallback = 1). */
int clowns_solver_classify(clowns_solver *h, const clowns_grid_desc *g,
                           const int8_t *d_solid, double *d_F,
                           const double *d_fb, const double *d_a,
                           const double *d_dx, const double *d_xf,
                           int8_t *d_nf, int8_t *d_nb, int32_t *d_mno,
                           int32_t *d_row2cell, double *d_sp_row, int32_t cap,
                           int32_t *nrows, double *mp, int32_t *host_fallback);

/* Device-resident grid for the assembly kernels (all pointers on the GPU). */
typedef struct clowns_grid_dev {
    int32_t nx, ny, nz, stride, nrows;
    int32_t reserved;
    double dt, mp;
    const int8_t *nf, *nb;
    const int32_t *mno, *row2cell;
    const double *sp_row, *F, *fb, *a, *un, *dx, *xf;
} clowns_grid_dev;

/* Replaces init_matrix + cal_matrix_element + the RHS of
 * cal_err_of_continuity (src2LOWNS/1sub-matrix.F90:3-87,
 * 1sub-momentum.F90:313-346, 465-551): writes the PCGPME inputs AL, LL
 * (column stride lda >= nrows, slots [x-,x+,y-,y+,z-,z+]), D and B (length
 * nrows) exactly as the reference's alu, llu, de and b.  `stream` is a
 * hipStream_t or NULL (default stream); synchronous on return. */
int clowns_assemble(const clowns_grid_dev *g, int32_t lda, double *d_AL,
                    int32_t *d_LL, double *d_D, double *d_B, void *stream);

/* The same assembly fused with PCGPME's row scaling and sort: the resident
 * operator of `h` is written in one pass (no unscaled copy in HBM unless
 * keep_raw, which the adjoint/gradient calls need); d_B gets the unscaled
 * right-hand side for clowns_solver_solve. */
int clowns_solver_assemble(clowns_solver *h, const clowns_grid_dev *g,
                           double parm, double *d_B, int32_t keep_raw);

/* Adjoint of the assembly: gradients of a loss L(x), x = A^{-1} b, with
 * respect to the face open ratios a and the cell open ratios fb from which
 * clowns_assemble / clowns_solver_assemble build A and b (cal_alu_f and the
 * diagonal of init_matrix, src2LOWNS/1sub-matrix.F90:50-61,82-87;
 * cal_matrix_element, 1sub-momentum.F90:481-548; the RHS,
 * 1sub-momentum.F90:315-346), given lambda = A^{-T} dL/dx (from
 * clowns_solver_solve_transpose) and x, both per row.  d_ga: 3*ncell entries
 * in the layout of g->a (face = the low face of a cell); d_gfb: ncell
 * entries; every entry is written, 0 where A and b do not depend on it
 * (solid cells, closed faces, surface rows).  The flags, numbering and
 * which faces are open stay as classified.  Either output may be NULL. */
int clowns_grad_open_ratios(clowns_solver *h, const clowns_grid_dev *g,
                            const double *d_lambda, const double *d_x,
                            double *d_ga, double *d_gfb);

/* The SMAC pressure-correction loop of one time step on the device
 * (src2LOWNS/1sub-momentum.F90:41-60 with check_dmax :419-463): the operator
 * is assembled once (fused with PCGPME's scaling and sort) and stays
 * resident; each pass computes the RHS and DMAX from d_un
 * (cal_err_of_continuity), stops when DMAX < d0, solves for xe (x0 = 0,
 * stopping rule eps; CLOWNS_MIXED selects mixed precision), updates
 * d_p(cell) -= xe and corrects d_un (estimate_velocity2, :590-651, with
 * pressure-boundary cells, object code 4).  d_un is 3*ncell low-face velocities
 * (used instead of g->un), d_p per cell.  More than max_passes solves ->
 * info->ier = 3000 (the reference stops at ik > 1000). */
typedef struct clowns_pc_info {
    int32_t passes;      /* solves done */
    int32_t iters;       /* total Bi-CGSTAB iterations */
    int32_t ier;
    int32_t reserved;
    double dmax_first, dmax_last;  /* DMAX before the first and after the last pass */
    double ms;           /* device time of the loop (HIP events) */
} clowns_pc_info;

int clowns_solver_pressure_correction(clowns_solver *h, const clowns_grid_dev *g,
                                      double *d_un, double *d_p, double d0,
                                      int32_t max_passes, double eps,
                                      int32_t flags, clowns_pc_info *info);

/* ---- 2DH dispersion solve: PCGPME_WILL ----------------------------------
 * Replaces PCGPME_WILL(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER)
 * (src2LOWNS/pcgpme_2DH.f90:1-41, interface 0mod-type_interface.F90:156-164),
 * called by disp_adjust (1sub-2DH-momentum.f90:325-326) with NL = 4,
 * NA = N, N2 = 64, PARM = 5, ITER = 1000.  The same PM-Bi-CGSTAB as PCGPME
 * with the reference's differences: AL, LL, D are inputs only; X starts at
 * zero; EPS and ITER are used as given (no defaults; ITER <= 0 runs no
 * iteration); BETA = C1/(AMU*RQ); and the output pass
 * (pcgpme_2DH.f90:255-268): |X(i)| <= 1e-14 -> 0, X(i) snapped to a
 * neighbour within 1e-14 in caller order, serial over i, and B(i) = X(i) for
 * every row with a differing neighbour, else B(i) keeps the final scaled
 * residual.  B: right-hand side in, that output out.  IER (not in the
 * reference's argument list, which has no status) gets 0 / 3000 / CLOWNS_E*. */
void clowns_pcgpme_will(const double *AL, const int32_t *NL, const int32_t *NA,
                        const int32_t *LL, const double *D, const int32_t *N,
                        const int32_t *N2, double *B, const double *EPS,
                        const double *PARM, const int32_t *ITER, int32_t *IER);

/* PCGPME_WILL on the operator last passed to clowns_solver_set_matrix (or
 * clowns_solver_assemble with keep_raw): d_b unscaled right-hand side, d_out
 * the reference's B on return (may be d_b).  *filter_rounds (may be NULL):
 * device rounds of the output pass, -1 if it ran on the host (a chain of
 * more than 4096 successive snaps). */
int clowns_solver_solve_will(clowns_solver *h, const double *d_b, double *d_out,
                             double eps, int32_t iter, clowns_info *info,
                             int32_t *filter_rounds);

/* The output pass of PCGPME_WILL alone, on the device: from the solution
 * d_X and final residual d_R (length N) and the caller-order LL (column
 * stride lda, entries in [0, N], NL <= 8), writes the filtered X to d_Xout
 * (may be NULL; must not be d_X) and B to d_B (may be d_R).  Exact: the
 * result equals the serial pass bit for bit. */
int clowns_will_filter(clowns_solver *h, const int32_t *d_LL, int32_t NL,
                       int32_t lda, int32_t N, const double *d_X,
                       const double *d_R, double *d_Xout, double *d_B,
                       int32_t *rounds);

/* Replaces set_dx (src2LOWNS/1sub-main.F90:53-74) for one axis, with the
 * reference's round (lib/src/xy_calc.f90:101-112): x[0..m) are the face
 * coordinates x(ls-2 .. le+1) (m = le-ls+4 >= 4), dx[0..m) receives
 * dx(ls-2 .. le+1).  A zero x(le+1) is extrapolated in place first; dx is
 * rounded to the number of decimals the first cell size suggests.  Host
 * function, start-up geometry. */
int clowns_set_dx(double *x, int32_t m, double *dx);

/* ---- the reference's case files -------------------------------------
 * Fortran unformatted sequential files (4-byte record lengths around each
 * record; big-endian in the reference build, either order read).  Grid
 * file: read_data (src2LOWNS/data_io.F90:98-128), ".xyz" (name ends in 'z':
 * nf/nfb dense in the data file) or ".xyzn" (list of nf = -1 cells,
 * output_xyzn :1531-1560).  Data file: read_data :130-156, :219-300 /
 * output_data :708-783 (T, g, nu; flags when dense; U, F, P of the numbered
 * cells; dt).  Open-ratio file: read_fdata :304-330 (a0, fb0).  Cells with
 * nf = -1 become solid cells (data_io.F90:168-199): nfb = 0 ones outside the
 * case (solid value 1); boundary objects (1 <= nfb < 200, numbered with the
 * flow cells 1..inne) and appended objects (nfb >= 200, numbered
 * inne+1..inne2: U and a0 only, fb 0) with their object code (the grid
 * descriptor's convention); nfb = -1 ones, which the reference maps to its
 * second sentinel cell (flag -1, b = -1), code -127 with zero fields. */
typedef struct clowns_deck_desc {
    int32_t is, js, ks, ie, je, ke;  /* the grid file's extents */
    int32_t nx, ny, nz;              /* box interior: ie-is, je-js, ke-ks */
    int32_t stride;                  /* max(nx,ny,nz)+2: axis array stride */
    int32_t sparse;                  /* 1: .xyzn grid file, 0: .xyz */
    int32_t big_endian;              /* byte order of the grid file */
    int32_t inne;                    /* flow cells and boundary objects (1..inne) */
    int32_t inne2;                   /* with the appended object cells */
    int32_t has_dt0;                 /* t0 > 0 and the data file ends with dt0 */
    double t0, g, nu, dt0;           /* the data file's header (with t0 = 0 the
                                        reference takes g, nu from its own
                                        parameter file instead) */
} clowns_deck_desc;

/* Read the headers and the cell flags (box dims, numbering count). */
int clowns_deck_open(const char *grid_path, const char *data_path,
                     clowns_deck_desc *d);
/* Fill the box arrays (ncell = (nx+2)(ny+2)(nz+2), axis arrays
 * [3][stride]): solid, F, un[3][ncell] (the data file's U), p, a[3][ncell]
 * and fb (from fdata_path, or 1 for numbered cells and 0 for solid ones when
 * it is NULL or ""), x = low-face coordinates X(l, ls-1+c) and dx from
 * set_dx (clowns_set_dx).  d from clowns_deck_open; dt0 is filled in. */
int clowns_deck_read(const char *grid_path, const char *data_path,
                     const char *fdata_path, clowns_deck_desc *d,
                     int8_t *solid, double *F, double *un, double *p,
                     double *a, double *fb, double *x, double *dx);
/* The grid file's faces as stored: faces[l*(stride+2) + q] = X(l, ls-2+q),
 * q = 0 .. n_l+3 (read_data, src2LOWNS/data_io.F90:113-116); the rest 0. */
int clowns_deck_faces(const char *grid_path, const clowns_deck_desc *d,
                      double *faces);
/* Write a sparse (.xyzn) grid file in output_xyzn's format
 * (src2LOWNS/data_io.F90:1505-1560, iflag = 0): the extents of d, faces as
 * clowns_deck_faces returns them, and every solid box cell as nf = -1,
 * nfb = 0. */
int clowns_deck_write_grid(const char *path, const clowns_deck_desc *d,
                           const int8_t *solid, const double *faces,
                           int32_t big_endian);
/* Write a data file in output_data's format from box arrays: nf/nb the cell
 * flags (clowns_classify), un, F, p; t and dt its first and last records;
 * d->sparse selects the .xyzn form. */
int clowns_deck_write_data(const char *path, const clowns_deck_desc *d,
                           const int8_t *nf, const int8_t *nb,
                           const double *un, const double *F,
                           const double *p, double t, double dt,
                           int32_t big_endian);

/* The VOF update cal_vof (src2LOWNS/1sub-vof.F90:10-38) on the device, up to
 * its closing set_nflag (run clowns_solver_classify for that):
 * cal_quvwf/QFF face fluxes, the surface stage, quvwf_modify (in the
 * reference's serial order), cal_new_f with its dt retry and
 * simple_slide(0) (SLIDE0 builds).
 *   g        the classified grid: nf, nb, F (current), fb, a, un (the new
 *            face velocities), dx, dt (nrows, mno, row2cell, sp_row, xf unused)
 *   d_Fp     F of the previous step (fp)
 *   mode 1   quvwf_kaidan (builds without NORMAL);
 *   mode 0   NORMAL builds: qf = d_qs where d_qs_mask != 0 ([3][ncell]), the
 *            PLIC face fluxes cal_idou2/a of quvwf_normal (:119-154); both
 *            NULL for none
 *   d_qf     out [3][ncell]: the face fluxes qf after the update
 *   d_Fnew   out [ncell]: the new F (may be g->F)
 * info->dt is dt after cal_new_f's reductions (dt/1.1 per retry; the
 * reference also moves t by the same amount).  CLOWNS_EVOF where the
 * reference STOPs.  Grids with object cells of the case (a flag -1 cell
 * with a nonzero object code) give CLOWNS_EBADARG: QFF's branches for
 * object donors (1sub-vof.F90:77-97) are not restated.  Synchronous on
 * return. */
typedef struct clowns_vof_info {
    int32_t retries;   /* cal_new_f dt reductions */
    int32_t modified;  /* surface cells corrected by quvwf_modify */
    int32_t slid;      /* cells that moved volume down in simple_slide */
    int32_t rounds;    /* device rounds of the serial quvwf_modify walk */
    double dt, ms;
} clowns_vof_info;
int clowns_solver_vof(clowns_solver *h, const clowns_grid_dev *g,
                      const double *d_Fp, int32_t mode, const double *d_qs,
                      const int8_t *d_qs_mask, double *d_qf, double *d_Fnew,
                      clowns_vof_info *info);

/* clowns_solver_vof for NORMAL builds with quvwf_normal itself on the device
 * (src2LOWNS/1sub-vof.F90:119-154): d_plic_cells lists the n_plic surface
 * cells of the PLIC reconstruction (the reference's sfnoi), d_plic_vol
 * holds per listed cell the six segment volumes cal_idou2 takes from its
 * polyhedron clipping (src2LOWNS/1sub-normal.F90:29-99), face l, low
 * (ll = 0) or high (ll = 1) at [6 m + 2 l + ll].  On each face the cell
 * drains through, qf = sign(u) V / area / dt / a (0 where |u| dt < 1e-9 or
 * V = 0); the rest as clowns_solver_vof.  The listed cells must be flow
 * cells (CLOWNS_EBADARG otherwise).
 * PARITY UNPINNED: this surface stage is restated from quvwf_normal's
 * source and checked only against the same formula written out by hand
 * (and, downstream of the fluxes, against the pinned given-flux mode); the
 * reference's PLIC state (the segment volumes it clips from its polyhedra,
 * src2LOWNS/1sub-normal*.F90) is not built here, so no reference output
 * exists to compare it with.  clowns_solver_vof (kaidan surface stage, or
 * caller-given PLIC fluxes) is the pinned path. */
int clowns_solver_vof_plic(clowns_solver *h, const clowns_grid_dev *g,
                           const double *d_Fp, int32_t n_plic,
                           const int32_t *d_plic_cells,
                           const double *d_plic_vol, double *d_qf,
                           double *d_Fnew, clowns_vof_info *info);

/* Average duration of one launch of a VOF-update kernel over `reps`
 * launches (HIP events on the solver's stream), on the inputs of
 * clowns_solver_vof: kernel 0 the face fluxes (kaidan surface stage),
 * 1 cal_new_f's pass, 2 simple_slide, 3 a plain copy of F into d_Fnew
 * (the copy rate of the same arrays).  Writes d_qf / d_Fnew. */
int clowns_solver_vof_time_kernel(clowns_solver *h, const clowns_grid_dev *g,
                                  const double *d_Fp, int32_t kernel,
                                  int32_t reps, double *d_qf, double *d_Fnew,
                                  float *ms_per_launch);

/* The per-operator kernel forms the solver chose for the resident operator
 * when it first solved at this size: out[0], out[1] pairwise-fused PM
 * stages (FP64, FP32: 1 fused, 0 single); out[2], out[3] streamed
 * (non-temporal) arrays (FP64, FP32: bit 0 the PM stages, bit 1 the SpMV);
 * out[4], out[5] the width of the neighbour-index codes (FP64, FP32: 8, 16
 * or 32); out[6] 1 if the FP64 PM stages read pattern rows' coefficients
 * from the row-pattern table (two 128-row chunks a wavefront), else 0;
 * out[7] the FP32 SpMV's form: 0 column-major, 1 row-major copies, 2
 * row-major in two 128-row chunks a wavefront, plus 4 when it
 * reads pattern rows' coefficients from the row-pattern table.  -1 where
 * not yet decided.  No counterpart in the reference (reporting). */
int clowns_solver_tuning(clowns_solver *h, int32_t out[8]);

/* Run the resident operator's gather kernels (SpMV, PM stages) over their
 * 512-row blocks in the given order: d_order[G] in device memory, a
 * permutation of 0..G-1 with G = ceil(N/512) rounded up to a multiple of 8
 * (the kernels' block count; checked).  Results are unchanged bit for bit;
 * only which rows run close together in time changes, and with it how often
 * a row's +-plane neighbours are still in L2.  d_order = NULL or G = 0
 * restores the natural order; an order whose G no longer matches the
 * resident operator is ignored.  No counterpart in the reference
 * (scheduling). */
int clowns_solver_set_block_order(clowns_solver *h, const int32_t *d_order, int32_t G);

/* Width of the neighbour indices the resident operator's FP64 solve kernels
 * read: 8 when they use the 8-bit codes (per 512-row block, four offset
 * classes with 6-bit steps; the few slots that do not fit read the 32-bit
 * LL; chosen per operator when it is first solved, or forced), 16 for the
 * 16-bit codes (every offset fits; always for box grids of up to 16383
 * cells a line), else 32; negative CLOWNS_E* on error.  No counterpart in
 * the reference (diagnostics). */
int clowns_solver_index_bits(clowns_solver *h);

/* Library build string, e.g. "clowns_pcg gfx950 ...". */
const char *clowns_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CLOWNS_PCG_H */
