// vof_kernels.hpp -- gfx950 kernels of the VOF update cal_vof
// (src2LOWNS/1sub-vof.F90:10-38) on the box layout, without its closing
// set_nflag (that is the classification in classify_kernels.hpp):
//
//   k_vof_flux       cal_quvwf + QFF (1sub-vof.F90:40-116) with the surface
//                    stage fused: quvwf_kaidan (:157-183) or the caller's
//                    PLIC face fluxes in place of quvwf_normal (:119-154)
//   k_vof_mod_eval   quvwf_modify (:186-213): one round of the serial walk
//   k_vof_mod_apply  quvwf_modify's flux corrections, in the serial order
//   k_vof_newf       cal_new_f (:215-269), one pass of its dt retry loop
//   k_vof_slide      simple_slide(0) + cal_fslide_m0 (:302-414), one thread
//                    per (i, j) column walking k downwards
//
// Layout: box index c = i + X*(j + Y*k); un, a, qf are [3][ncell] and hold
// the low face of c in each direction; dx is [3][stride].  A solid cell
// (nf = -1) is the reference's numbering sentinel mn = 0: read as a
// neighbour it gives a = 0 and a face flux of 0.
//
// quvwf_modify walks the surface list in order (ascending c, the k, j, i
// scan of set_nflag) and a correction changes the flux that later cells
// read.  A correction by cell r lands on a face of r; the only later cell
// that reads that face is t = r + e_l, when r points along -e_l (nfb =
// -(l+1)).  So a surface cell's evaluation depends only on its at most
// three "predecessors" r = c - e_l with nfb(r) = -(l+1) that are surface
// cells: rounds evaluate every cell whose predecessors are decided (most
// cells in the first round), and the corrections are applied afterwards per
// face, predecessor first -- the serial result, bit for bit.
//
// No fused multiply-add (-ffp-contract=off): every expression rounds as the
// reference's does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clowns {

constexpr double VOF_TINY = 1e-15;          // TINY, 0mod-variables.F90:39
constexpr double VOF_FLIM = 1.0 + 1e-5;     // F_lim, 0mod-variables.F90:40
constexpr double VOF_DTLIM = 1e-6;          // DT_lim, 0mod-variables.F90:41

struct VofDev {
    int nx, ny, nz, stride;
    long ncell;
    const int8_t *nf, *nb;
    const double *F, *Fp, *fb, *a, *un, *dx;
    double *qf, *Fn;
    double *D;  // df_vof of every open cell after quvwf_modify (k_vof_newf -> k_vof_slide)
};

// Launches of k_vof_flux / k_vof_newf / k_vof_slide have a multiple of 8
// blocks; the hardware deals blocks to the 8 XCDs round robin, so block b
// takes the logical index (b mod 8) * G/8 + b/8: each XCD streams one
// contiguous range, and the +-y neighbours (one block away) sit in its own L2.
__device__ __forceinline__ int vof_block() {
    const int b = blockIdx.x;
    return (b & 7) * (gridDim.x >> 3) + (b >> 3);
}

// counters: [0] retry (cal_new_f's icon = 30), [1] stop (F > F_lim on a
// surface cell), [2] correction aimed at a solid cell, [3] pending cells of
// a modify round, [4] corrected cells, [5] slid cells, [6] invalid flags
enum { VC_RETRY = 0, VC_STOP = 1, VC_SOLID = 2, VC_PENDING = 3, VC_MODIFIED = 4, VC_SLID = 5, VC_BAD = 6,
       VC_N = 8 };

__device__ inline void vof_pos(const VofDev &v, long c, int p[3]) {
    const int X = v.nx + 2, Y = v.ny + 2;
    p[0] = (int)(c % X);
    p[1] = (int)((c / X) % Y);
    p[2] = (int)(c / ((long)X * Y));
}

// df_vof(nn) = fp - dt*renzoku_fv(...) (1sub-vof.F90:271-300,
// lib/src/math_func.F90:51-54); the low faces of c may carry a pending
// correction of a predecessor (sub[l], applied as qf - sub).
__device__ inline double vof_df(const VofDev &v, long c, const int p[3], double dt, const double sub[3],
                                const bool has_sub[3]) {
    const long st[3] = {1, (long)(v.nx + 2), (long)(v.nx + 2) * (v.ny + 2)};
    double qp[3], qm[3], ap[3], am[3];
    for (int l = 0; l < 3; ++l) {
        const long h = c + st[l];
        const size_t oc = (size_t)l * v.ncell + c, oh = (size_t)l * v.ncell + h;
        const bool hs = v.nf[h] == -1;
        qp[l] = hs ? 0.0 : v.qf[oh];
        ap[l] = hs ? 0.0 : v.a[oh];
        const double q = v.qf[oc];
        qm[l] = has_sub[l] ? q - sub[l] : q;
        am[l] = v.a[oc];
    }
    const double r = ((qp[0] * ap[0] - qm[0] * am[0]) / v.dx[p[0]] +
                      (qp[1] * ap[1] - qm[1] * am[1]) / v.dx[(size_t)v.stride + p[1]] +
                      (qp[2] * ap[2] - qm[2] * am[2]) / v.dx[2 * (size_t)v.stride + p[2]]) / v.fb[c];
    return v.Fp[c] - dt * r;
}

__device__ inline double vof_df(const VofDev &v, long c, const int p[3], double dt) {
    const double z[3] = {0.0, 0.0, 0.0};
    const bool n[3] = {false, false, false};
    return vof_df(v, c, p, dt, z, n);
}

__device__ inline double vof_sign(double a, double b) { return signbit(b) ? -fabs(a) : fabs(a); }

// MODE 0: caller's PLIC face fluxes where qmask != 0 (qs/qmask may be null);
// MODE 1: quvwf_kaidan.
template <int MODE>
__global__ void __launch_bounds__(256)
k_vof_flux(VofDev v, double dt, const double *__restrict__ qs, const int8_t *__restrict__ qmask) {
    const long c = (long)vof_block() * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    const long st[3] = {1, (long)(v.nx + 2), (long)(v.nx + 2) * (v.ny + 2)};
    const size_t n = (size_t)v.ncell;
    if (v.nf[c] == -1) {
        for (int l = 0; l < 3; ++l) v.qf[l * n + c] = 0.0;
        return;
    }
    int p[3];
    vof_pos(v, c, p);
    for (int l = 0; l < 3; ++l) {
        const size_t o = l * n + c;
        const double u = v.un[o];
        const long m = c - st[l];
        // QFF: the donor is the upwind cell (c - e_l for u > 0, c for u < 0)
        const bool neg = signbit(u);
        const long nnd = neg ? c : m;
        double q = 0.0;
        if (u != 0.0 && v.nf[m] != -1) {
            const int fd = v.nf[nnd], bd = v.nb[nnd];
            const double Fd = v.F[nnd];
            if (fd == 2 && (bd < 0 ? -bd : bd) == l + 1) {
                const double dxx = v.dx[(size_t)l * v.stride + (neg ? p[l] : p[l] - 1)];
                if ((double)bd * u < 0.0) {
                    const double t = fabs(u) - (1.0 - Fd) * dxx / dt;
                    q = vof_sign(1.0, u) * (t > 0.0 ? t : 0.0);
                } else {
                    const double t = Fd * dxx / dt, au = fabs(u);
                    q = vof_sign(1.0, u) * (au < t ? au : t);
                }
            } else {
                q = u * Fd;
            }
        }
        if (MODE == 1) {
            if (l < 2) {
                const double av = v.a[o];
                if (av > 0.0 && av < 1.0) {
                    if (u < 0.0 && v.nf[c] == 2 && av != v.fb[c]) {
                        const double t = av - v.fb[c] + v.F[c] * v.fb[c];
                        q = u * (t > 0.0 ? t : 0.0) / av;
                    } else if (u > 0.0 && v.nf[m] == 2 && av != v.fb[m]) {
                        const double t = av - v.fb[m] + v.F[m] * v.fb[m];
                        q = u * (t > 0.0 ? t : 0.0) / av;
                    }
                }
            }
        } else {
            if (qmask && qmask[o]) q = qs[o];
        }
        v.qf[o] = q;
    }
}

// quvwf_normal (1sub-vof.F90:119-154) on the listed surface cells, with
// cal_idou2's flux (1sub-normal.F90:29-99) from the caller's PLIC segment
// volumes vol[6 m + 2 l + ll] (face l, ll = 0 low / 1 high of cells[m]):
// on each face the cell drains through (un (1 - 2 ll) < 0, a != 0),
// qf = sign(u) V / area / dt / a, 0 for |u| dt < SMALL or V = 0.  Runs after
// k_vof_flux<0>.  Two listed cells never write one face: the shared face's
// velocity drains only one of them.
__global__ void __launch_bounds__(256)
k_vof_plic(VofDev v, double dt, int n, const int *__restrict__ cells, const double *__restrict__ vol) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const long c = cells[m];
    const long st[3] = {1, (long)(v.nx + 2), (long)(v.nx + 2) * (v.ny + 2)};
    int p[3];
    vof_pos(v, c, p);
    const size_t S = (size_t)v.stride, nc = (size_t)v.ncell;
    const double dxx = v.dx[p[0]], dyy = v.dx[S + p[1]], dzz = v.dx[2 * S + p[2]];
    for (int l = 0; l < 3; ++l)
        for (int ll = 0; ll < 2; ++ll) {
            const long nn = c + ll * st[l];
            const bool solid = v.nf[nn] == -1;
            const double u = solid ? 0.0 : v.un[l * nc + nn];
            const double a = solid ? 0.0 : v.a[l * nc + nn];
            const double lsn = 1.0 - 2.0 * ll;
            if (u * lsn >= 0.0 || a == 0.0) continue;
            double idou = 0.0;
            if (!(fabs(u) * dt < 1e-9)) {  // SMALL, 0mod-variables.F90:39
                const double area = l == 2 ? dxx * dyy : (l == 1 ? dxx * dzz : dzz * dyy);
                const double vs = vol[6 * (size_t)m + 2 * l + ll];
                idou = vs == 0.0 ? 0.0 : vof_sign(1.0, u) * vs / area / dt;
            }
            v.qf[l * nc + nn] = idou / a;
        }
}

// Each listed PLIC cell must be a flow cell (interior: its neighbour reads
// stay in the box, the ring being solid).
__global__ void __launch_bounds__(256)
k_vof_check_plic(VofDev v, int n, const int *__restrict__ cells, int *__restrict__ cnt) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const long c = cells[m];
    if (c < 0 || c >= v.ncell || v.nf[c] == -1) atomicOr(&cnt[VC_BAD], 1);
}

// The flags every other kernel relies on for its neighbour reads: the box
// ring is solid (nf = -1), nf is -1..2, and a surface cell has 1 <= |nb| <= 3.
__global__ void __launch_bounds__(256) k_vof_check(VofDev v, int *__restrict__ cnt) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    int p[3];
    vof_pos(v, c, p);
    const int f = v.nf[c], b = v.nb[c];
    const bool ring = p[0] == 0 || p[0] == v.nx + 1 || p[1] == 0 || p[1] == v.ny + 1 || p[2] == 0 || p[2] == v.nz + 1;
    // object cells of the case (flag -1 with a nonzero object code) are not
    // covered: QFF's object branches (1sub-vof.F90:77-97) are not restated
    const bool bad = (ring && f != -1) || f < -1 || f > 2 || (f == 2 && (b == 0 || b < -3 || b > 3)) ||
                     (f == -1 && b != 0);
    if (bad) atomicOr(&cnt[VC_BAD], 1);
}

__global__ void __launch_bounds__(256) k_vof_state_init(long n, int *__restrict__ state) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) state[c] = 0;
}

// state: 0 pending, else 2*round + (1 if the cell's correction term is set).
// A cell counts a predecessor as decided only if it was decided in an
// earlier round (earlier launch), so the term it reads is complete.
__device__ inline bool vof_decided(int s, int round) { return s != 0 && (s >> 1) < round; }

__global__ void __launch_bounds__(256)
k_vof_mod_eval(VofDev v, double dt, int round, int *__restrict__ state, double *__restrict__ term,
               int *__restrict__ cnt) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell || v.nf[c] != 2 || state[c] != 0) return;
    const long st[3] = {1, (long)(v.nx + 2), (long)(v.nx + 2) * (v.ny + 2)};
    double sub[3] = {0.0, 0.0, 0.0};
    bool has[3] = {false, false, false};
    for (int l = 0; l < 3; ++l) {
        const long r = c - st[l];
        if (v.nf[r] == 2 && v.nb[r] == -(l + 1)) {
            const int s = state[r];
            if (!vof_decided(s, round)) {
                atomicAdd(&cnt[VC_PENDING], 1);
                return;
            }
            if (s & 1) {
                sub[l] = term[r];
                has[l] = true;
            }
        }
    }
    int p[3];
    vof_pos(v, c, p);
    const double d = vof_df(v, c, p, dt, sub, has);
    if (d < VOF_FLIM) {
        state[c] = 2 * round;
        return;
    }
    const int b = v.nb[c], l = (b < 0 ? -b : b) - 1;
    const long n1 = b > 0 ? c : c + st[l];
    if (v.nf[n1] == -1) atomicOr(&cnt[VC_SOLID], 1);
    const double s = b > 0 ? 1.0 : -1.0;
    term[c] = v.dx[(size_t)l * v.stride + p[l]] * (d - 1.0) / dt * v.fb[c] / v.a[(size_t)l * v.ncell + n1] * s;
    state[c] = 2 * round + 1;
    atomicAdd(&cnt[VC_MODIFIED], 1);
}

__global__ void __launch_bounds__(256)
k_vof_mod_apply(VofDev v, const int *__restrict__ state, const double *__restrict__ term) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= v.ncell || v.nf[c] != 2 || !(state[c] & 1)) return;
    const long st[3] = {1, (long)(v.nx + 2), (long)(v.nx + 2) * (v.ny + 2)};
    const int b = v.nb[c], l = (b < 0 ? -b : b) - 1;
    const size_t n = (size_t)v.ncell;
    if (b > 0) {
        // own low face: a predecessor's correction first, then this one
        double q = v.qf[l * n + c];
        const long r = c - st[l];
        if (v.nf[r] == 2 && v.nb[r] == -(l + 1) && (state[r] & 1)) q = q - term[r];
        v.qf[l * n + c] = q - term[c];
    } else {
        const long t = c + st[l];
        if (v.nf[t] == -1) return;  // reported through VC_SOLID
        if (v.nf[t] == 2 && v.nb[t] == l + 1 && (state[t] & 1)) return;  // t applies both
        v.qf[l * n + t] = v.qf[l * n + t] - term[c];
    }
}

__global__ void __launch_bounds__(256) k_vof_newf(VofDev v, double dt, int *__restrict__ cnt) {
    const long c = (long)vof_block() * blockDim.x + threadIdx.x;
    if (c >= v.ncell) return;
    const int f = v.nf[c];
    if (f == -1) {
        v.Fn[c] = v.F[c];
        return;
    }
    int p[3];
    vof_pos(v, c, p);
    double d = vof_df(v, c, p, dt);
    v.D[c] = d;
    if (-d * dt * v.fb[c] > 0.4) {
        atomicOr(&cnt[VC_RETRY], 1);
    } else if (fabs(d * dt) < VOF_TINY && d != 0.0) {
        d = 0.0;
    }
    if (d > VOF_FLIM && f == 2) atomicOr(&cnt[VC_STOP], 1);
    v.Fn[c] = d;
}

// simple_slide(0): one thread per column, k downwards.  cal_fslide_m0
// recomputes df_vof of the cell and of the one below before using them;
// those equal D (k_vof_newf's values on the same fluxes) unless a move in
// this column changed a face since: the cell's high face is changed by a
// move one level up (then its df_vof is the value that move left for the
// cell below), the lower cell's faces only by moves not yet made.
__global__ void __launch_bounds__(256) k_vof_slide(VofDev v, double dt, int *__restrict__ cnt) {
    const int col = vof_block() * blockDim.x + threadIdx.x;
    if (col >= v.nx * v.ny) return;
    const int i = 1 + col % v.nx, j = 1 + col / v.nx;
    const long X = v.nx + 2, Y = v.ny + 2, sz = X * Y;
    const size_t S = (size_t)v.stride;
    const double dx0 = v.dx[i], dx1 = v.dx[S + j];
    int moved = 0;
    bool changed_above = false;  // qf(2, c + e_z) moved by the previous level
    double carried = 0.0;        // df_vof(c) after that move
    for (int k = v.nz; k >= 1; --k) {
        const long c = i + X * (j + Y * k), m = c - sz;
        const int f = v.nf[c];
        if (f == -1) break;
        if (f == 1 || v.nf[m] == -1) {
            changed_above = false;
            continue;
        }
        const double vmesh = dx0 * dx1 * v.dx[2 * S + k];
        const double dvc = vmesh * v.fb[c];
        const double dvm = dx0 * dx1 * v.dx[2 * S + k - 1] * v.fb[m];
        if (dvc * dvm == 0.0) {
            changed_above = false;
            continue;
        }
        const double fc = changed_above ? carried : v.D[c];
        const double fm = v.D[m];
        v.Fn[c] = fc;
        v.Fn[m] = fm;
        const double t1 = dvm * (1.0 - fm), t2 = dvc * fc;
        const double f2 = t1 < t2 ? t1 : t2;
        if (fabs(f2) == 0.0) {
            changed_above = false;
            continue;
        }
        const size_t o = 2 * (size_t)v.ncell + c;
        v.qf[o] = v.qf[o] + -1.0 * f2 / vmesh / dt * v.dx[2 * S + k] / v.a[o];
        const int pc[3] = {i, j, k}, pm[3] = {i, j, k - 1};
        const double ft = vof_df(v, c, pc, dt);
        v.Fn[c] = fabs(ft) < VOF_TINY ? 0.0 : ft;
        carried = vof_df(v, m, pm, dt);
        v.Fn[m] = carried;
        changed_above = true;
        ++moved;
    }
    if (moved) atomicAdd(&cnt[VC_SLID], moved);
}

} // namespace clowns
