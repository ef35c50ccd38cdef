// classify_kernels.hpp -- device form of the cell classification and row
// numbering (set_nflag, src2LOWNS/nfd2015.F90:7-306; numbering,
// src2LOWNS/1sub-momentum.F90:278-296; surface position,
// src2LOWNS/1sub-normal-ini.F90:23-40), one thread per box cell.
//
// set_nflag is defined by its serial order (its OpenMP loops race).  The
// passes map to the GPU as follows:
//   initial flags (:36-56)        per cell; F read unclipped (in the serial
//                                 order the cell above is visited later), the
//                                 air clip F = max(0, F) applied after
//   surface detection (:69-102)   per cell; tests neighbours for == 0 only,
//                                 which the pass never changes
//   unsupported cells (:110-126)  per surface cell; tests == 1 and == -1 only,
//                                 which a 2 -> 0 drop never changes
//   one-sided cells (:135-152)    order dependent.  A 2 -> 0 drop can only make
//                                 a later (lower-index) cell's keep test true,
//                                 never false, so every cell the serial sweep
//                                 drops is also dropped by an evaluation on the
//                                 pre-sweep flags.  If that evaluation drops
//                                 nothing, the serial sweep drops nothing.
//   orientation (:167-243)        order dependent only through 'souteigai'
//                                 cells becoming air; if none occurs on the
//                                 pre-sweep flags, none occurs serially.
// When either order-dependent pass would change a flag, the host runs the
// exact serial classification instead (the caller is told).
// Numbering: rows are the flow cells of the row list in box order (the
// reference's fln order), so mno is an exclusive prefix sum over the box.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clowns {

struct ClsGrid {
    int nx, ny, nz, X, Y, stride;
    long ncell;
    const int8_t *solid;
    const double *F, *fb, *a, *dx, *xf;
    int8_t *nf, *nb;
};

__device__ __forceinline__ int cls_flag(const ClsGrid &g, long c) {
    return g.solid[c] ? -1 : g.nf[c];
}
// solid[c] = 1: outside the case (reads as the sentinel: 0); other nonzero
// values: an object cell of the case, read with its own values.
__device__ __forceinline__ double cls_face(const ClsGrid &g, int l, long c) {
    return g.solid[c] == 1 ? 0.0 : g.a[(size_t)l * g.ncell + c];
}
__device__ __forceinline__ void cls_coords(const ClsGrid &g, long c, int p[3]) {
    p[0] = (int)(c % g.X);
    p[1] = (int)((c / g.X) % g.Y);
    p[2] = (int)(c / ((long)g.X * g.Y));
}
__device__ __forceinline__ bool cls_interior(const ClsGrid &g, const int p[3]) {
    return p[0] >= 1 && p[0] <= g.nx && p[1] >= 1 && p[1] <= g.ny && p[2] >= 1 && p[2] <= g.nz;
}

// ext[0] = max i, ext[1] = max k of fluid cells (block maximum, then one
// atomic per block); ext[4] set if any F < 0.
__global__ void __launch_bounds__(256) k_cls_init(ClsGrid g, int *__restrict__ ext) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    int mi = 0, mk = 0;
    if (c < g.ncell) {
        if (g.solid[c]) {
            g.nf[c] = -1;
            g.nb[c] = g.solid[c] == 1 ? 0 : g.solid[c];  // object code
        } else {
            int p[3];
            cls_coords(g, c, p);
            const long up = c + (long)g.X * g.Y, dn = c - (long)g.X * g.Y;
            const double f = g.F[c];
            const bool air =
                f <= 1e-12 || (f * g.fb[c] < 1.0e-3 && g.solid[dn] && (g.solid[up] == 1 ? 0.0 : g.F[up]) == 0.0);
            if (f < 0.0) atomicOr(ext + 4, 1);  // clipping order matters: exact host path
            g.nb[c] = 0;
            if (air) {
                g.nf[c] = 0;
            } else {
                g.nf[c] = 1;
                mi = p[0];
                mk = p[2];
            }
        }
    }
    __shared__ int li[256], lk[256];
    li[threadIdx.x] = mi;
    lk[threadIdx.x] = mk;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            li[threadIdx.x] = max(li[threadIdx.x], li[threadIdx.x + s]);
            lk[threadIdx.x] = max(lk[threadIdx.x], lk[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && (li[0] | lk[0])) {
        atomicMax(ext, li[0]);
        atomicMax(ext + 1, lk[0]);
    }
}

__global__ void __launch_bounds__(256) k_cls_clip(long n, const int8_t *__restrict__ nf, double *__restrict__ F) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    if (nf[c] == 0 && F[c] < 0.0) F[c] = 0.0;
}

// Surface detection over the row-list range (k <= kmax, i <= imax).
__global__ void __launch_bounds__(256) k_cls_surface(ClsGrid g, int imax, int kmax) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= g.ncell) return;
    int p[3];
    cls_coords(g, c, p);
    if (!cls_interior(g, p) || p[0] > imax || p[2] > kmax) return;
    if (cls_flag(g, c) < 1) return;
    const long step[3] = {1, g.X, (long)g.X * g.Y};
    bool surface = false;
    for (int l = 0; l < 3 && !surface; ++l) {
        const long up = c + step[l], dn = c - step[l];
        surface = (cls_flag(g, up) == 0 && cls_face(g, l, up) > 0.0) ||
                  (cls_flag(g, dn) == 0 && g.a[(size_t)l * g.ncell + c] > 0.0);
    }
    g.nf[c] = surface ? 2 : 1;
    g.nb[c] = surface ? -3 : 0;
}

// Surface cells with no fluid neighbour (and no object below) become air.
__global__ void __launch_bounds__(256) k_cls_unsupported(ClsGrid g, int *__restrict__ changed) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= g.ncell || g.solid[c] || g.nf[c] != 2) return;
    const long step[3] = {1, g.X, (long)g.X * g.Y};
    if (cls_flag(g, c - step[2]) == -1) return;
    for (int l = 0; l < 3; ++l)
        if (cls_flag(g, c + step[l]) == 1 || cls_flag(g, c - step[l]) == 1) return;
    g.nf[c] = 0;
    g.nb[c] = 0;
    atomicOr(changed, 1);
}

// One-sided test on the pre-sweep flags; any drop -> *changed (host redoes).
__global__ void __launch_bounds__(256) k_cls_onesided(ClsGrid g, int *__restrict__ changed) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= g.ncell || g.solid[c] || g.nf[c] != 2) return;
    const long step[3] = {1, g.X, (long)g.X * g.Y};
    if (cls_flag(g, c - step[2]) == -1 && cls_flag(g, c + step[2]) == 0) return;
    for (int l = 0; l < 3; ++l) {
        const int fp = cls_flag(g, c + step[l]), fm = cls_flag(g, c - step[l]);
        if ((fp == 1 && fm == 0) || (fm == 1 && fp == 0)) return;
    }
    atomicOr(changed, 1);
}

__device__ double cls_fxx1(const ClsGrid &g, int ix, int ll, const int lg[3]) {
    const long step[3] = {1, g.X, (long)g.X * g.Y};
    auto h = [&](int l, int i) { return g.dx[(size_t)l * g.stride + i]; };
    double s = 0.0;
    for (int l = 0; l < 3; ++l) {
        if (l == ll) continue;
        const long base = (lg[0] + (long)g.X * (lg[1] + (long)g.Y * lg[2])) + ix * step[ll];
        for (int ii = -1; ii <= 1; ++ii) {
            const long n = base + ii * step[l];
            const double f = g.solid[n] == 1 ? 0.0 : g.F[n];
            const double fb = g.solid[n] == 1 ? 0.0 : g.fb[n];
            s = s + h(l, lg[l] + ii) * f * fb;
        }
        for (int m = 0; m < 3; ++m)
            if (m != l && m != ll) s = s * h(m, lg[m]);
    }
    return s * h(ll, lg[ll] + ix);
}

// Orientation of every surface cell; a cell without one ('souteigai') sets
// *changed and is left for the host.
__global__ void __launch_bounds__(256) k_cls_orient(ClsGrid g, int *__restrict__ changed) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= g.ncell || g.solid[c] || g.nf[c] != 2) return;
    const long step[3] = {1, g.X, (long)g.X * g.Y};
    int ix[3] = {0, 0, 0};
    for (int l = 0; l < 3; ++l) {
        const long up = c + step[l], dn = c - step[l];
        if (cls_flag(g, dn) == 1 && cls_flag(g, up) == 0 && cls_face(g, l, up) > 0.0)
            ix[l] = -1;
        else if (cls_flag(g, dn) == 0 && cls_flag(g, up) == 1 && g.a[(size_t)l * g.ncell + c] > 0.0)
            ix[l] = 1;
    }
    if (ix[0] == 0 && ix[1] == 0 && ix[2] == 0 && cls_flag(g, c - step[2]) == -1 &&
        cls_flag(g, c + step[2]) == 0) {
        if (g.fb[c] != 0.0) {
            ix[2] = -1;
        } else {
            for (int l = 0; l < 2; ++l) {
                const long up = c + step[l], dn = c - step[l];
                if (cls_flag(g, dn) > 0 && cls_flag(g, up) == 0 && cls_face(g, l, up) > 0.0)
                    ix[l] = -1;
                else if (cls_flag(g, dn) == 0 && cls_flag(g, up) > 0 && g.a[(size_t)l * g.ncell + c] > 0.0)
                    ix[l] = 1;
            }
            if (ix[0] == 0 && ix[1] == 0) ix[2] = -1;
        }
    }
    const int n = abs(ix[0]) + abs(ix[1]) + abs(ix[2]);
    int dir = 0;
    if (n == 1) {
        for (int l = 1; l < 3; ++l)
            if (abs(ix[l]) > abs(ix[dir])) dir = l;
    } else if (n == 0) {
        atomicOr(changed, 1);
        return;
    } else {
        int p[3];
        cls_coords(g, c, p);
        double w[3];
        for (int l = 0; l < 3; ++l) w[l] = ix[l] == 0 ? 0.0 : cls_fxx1(g, ix[l], l, p);
        for (int l = 1; l < 3; ++l)
            if (w[l] > w[dir]) dir = l;
    }
    g.nb[c] = (int8_t)((dir + 1) * ix[dir]);
}

// Row test (1sub-momentum.F90:286-290) for cells of the row-list range.
__device__ __forceinline__ int cls_is_row(const ClsGrid &g, long c, int imax, int kmax) {
    if (g.solid[c]) return 0;
    const int f = g.nf[c];
    if (f < 1 || g.fb[c] == 0.0) return 0;
    int p[3];
    cls_coords(g, c, p);
    if (!cls_interior(g, p) || p[0] > imax || p[2] > kmax) return 0;
    if (f == 2 && g.nb[c] != 0) {
        const long step[3] = {1, g.X, (long)g.X * g.Y};
        const int b = g.nb[c], l = abs(b) - 1, sg = b > 0 ? 1 : -1;
        if (cls_flag(g, c + sg * step[l]) == -1) return 0;
    }
    return 1;
}

constexpr int SCAN_BS = 256, SCAN_PER = 8;  // cells per block: 2048

__global__ void __launch_bounds__(SCAN_BS)
k_cls_count(ClsGrid g, int imax, int kmax, int *__restrict__ bcount) {
    const long base = (long)blockIdx.x * SCAN_BS * SCAN_PER;
    int n = 0;
    for (int u = 0; u < SCAN_PER; ++u) {
        const long c = base + (long)u * SCAN_BS + threadIdx.x;
        if (c < g.ncell) n += cls_is_row(g, c, imax, kmax);
    }
    __shared__ int lds[SCAN_BS];
    lds[threadIdx.x] = n;
    __syncthreads();
    for (int s = SCAN_BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bcount[blockIdx.x] = lds[0];
}

// Exclusive scan of the block counts in one block; total in bcount[nb].
__global__ void __launch_bounds__(1024) k_cls_scan_blocks(int nb, int *__restrict__ bcount) {
    __shared__ int lds[1024];
    const int per = (nb + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = min(nb, lo + per);
    int s = 0;
    for (int k = lo; k < hi; ++k) s += bcount[k];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += v;
        __syncthreads();
    }
    int run = lds[threadIdx.x] - s;  // exclusive prefix of this thread's range
    for (int k = lo; k < hi; ++k) {
        const int v = bcount[k];
        bcount[k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) bcount[nb] = lds[1023];
}

// Row numbers in box order; sp_row for surface rows (1sub-normal-ini.F90:27-38).
__global__ void __launch_bounds__(SCAN_BS)
k_cls_number(ClsGrid g, int imax, int kmax, const int *__restrict__ bstart, int cap,
             int32_t *__restrict__ mno, int32_t *__restrict__ row2cell, double *__restrict__ sp_row) {
    const long base = (long)blockIdx.x * SCAN_BS * SCAN_PER;
    __shared__ int lds[SCAN_BS];
    int run = bstart[blockIdx.x];
    for (int u = 0; u < SCAN_PER; ++u) {
        const long c = base + (long)u * SCAN_BS + threadIdx.x;
        const int r = c < g.ncell ? cls_is_row(g, c, imax, kmax) : 0;
        lds[threadIdx.x] = r;
        __syncthreads();
        for (int off = 1; off < SCAN_BS; off <<= 1) {
            const int v = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
            __syncthreads();
            lds[threadIdx.x] += v;
            __syncthreads();
        }
        const int incl = lds[threadIdx.x];
        const int total = lds[SCAN_BS - 1];
        if (c < g.ncell) {
            if (r) {
                const int row = run + incl - 1;  // 0-based
                mno[c] = row + 1;
                if (row < cap) {
                    row2cell[row] = (int32_t)c;
                    double s = 0.0;
                    if (g.nf[c] == 2 && g.nb[c] != 0) {
                        int p[3];
                        cls_coords(g, c, p);
                        const int b = g.nb[c], l = abs(b) - 1;
                        const double f = g.F[c], fbc = g.fb[c];
                        double fff;
                        if (b == -3 && fbc < 1.0)
                            fff = 1.0 - fbc * (1.0 - f);
                        else if (b < 0)
                            fff = f;
                        else
                            fff = 1.0 - f;
                        s = g.xf[(size_t)l * g.stride + p[l]] + g.dx[(size_t)l * g.stride + p[l]] * fff;
                        if (fabs(s) < 1e-15) s = 0.0;
                    }
                    sp_row[row] = s;
                }
            } else {
                mno[c] = 0;
            }
        }
        run += total;
        __syncthreads();
    }
}

} // namespace clowns
