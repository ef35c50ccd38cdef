// assembly_kernels.hpp -- per-row assembly of the pressure system on the GPU.
//
// One thread per matrix row builds what the reference builds in three
// places: the static coefficients of init_matrix (src2LOWNS/1sub-matrix.F90:
// 32-64, cal_alu_f :82-87), the per-pass row of cal_matrix_element
// (src2LOWNS/1sub-momentum.F90:481-548, cal_alu 1sub-matrix.F90:75-80) and the
// right-hand side of cal_err_of_continuity (1sub-momentum.F90:315-346,
// renzoku_fv lib/src/math_func.F90:51-54).  Neighbours are fixed offsets on
// the box (+-1, +-X, +-X*Y), so the gathers of flags, row numbers, open ratios
// and velocities are coalesced across consecutive rows.  Expressions keep the
// reference's order; with -ffp-contract=off the output equals the reference
// bit for bit.
//
// The FUSED form continues in registers with PCGPME's row scaling, ghost
// indices and exchange sort (pcgpme_para.F90:77-112) and writes the resident
// operator directly: the unscaled system never round-trips through HBM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clowns {

struct GridDev {
    int nx, ny, nz, stride, N;
    long ncell;
    double dt, mp;
    const int8_t *nf, *nb;
    const int32_t *mno, *row2cell;
    const double *sp_row, *F, *fb, *a, *un, *dx, *xf;
};

// A cell of the case (the reference's mn != 0): a flow cell, or an object
// cell, whose nb is its nonzero object code (4: a pressure-boundary cell).
// A cell outside the case reads as the reference's sentinel cell 0: its
// fb, a and un are 0.
__device__ __forceinline__ bool live(const GridDev &g, long c) { return g.nf[c] != -1 || g.nb[c] != 0; }

template <bool FUSED>
__global__ void __launch_bounds__(256)
k_assemble(GridDev g, int lda, double *__restrict__ AL, int *__restrict__ LL,
           double *__restrict__ D, double *__restrict__ B, double *__restrict__ W,
           double *__restrict__ AL0, int *__restrict__ LL0, double *__restrict__ D0,
           double P1, double P2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.N) return;
    const int X = g.nx + 2, Y = g.ny + 2;
    const long c = g.row2cell[r];
    const int p[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
    const long step[3] = {1, X, (long)X * Y};
    const int fc = g.nf[c];
    const bool fluid = fc == 1;
    const double fbc = g.fb[c];
    const double mp = g.mp;

    double alu[6];
    int llu[6];
    double de_c = 0.0;
    // init_matrix: ALU_C, then cal_matrix_element per slot [x-,x+,y-,y+,z-,z+]
    double de_adj[6];  // ALU_C added back for air neighbours (Neumann)
    int adj = 0;
#pragma unroll
    for (int slot = 0; slot < 6; ++slot) {
        const int l = slot >> 1, side = (slot & 1) ? 1 : -1;
        const long nc = c + side * step[l];
        const long pc = side > 0 ? nc : c;  // cell whose low face is shared
        const int fn = g.nf[nc];
        const double af = live(g, pc) ? g.a[(size_t)l * g.ncell + pc] : 0.0;
        const double fbn = live(g, nc) ? g.fb[nc] : 0.0;
        alu[slot] = 0.0;
        llu[slot] = 0;
        de_adj[slot] = 0.0;
        // LL_C = 0: an object neighbour other than a pressure-boundary cell,
        // a closed face or an empty neighbour (1sub-matrix.F90:41-46)
        if ((fn == -1 && g.nb[nc] != 4) || af == 0.0 || fbn == 0.0) continue;
        const double h = g.dx[(size_t)l * g.stride + p[l]];
        const double hm = g.dx[(size_t)l * g.stride + p[l] + side];
        const double alu_c = -mp * (2.0 / h / (h + hm) * af / fbc);
        de_c = de_c - alu_c;
        const int np = g.mno[nc];
        if (fn == 0 || np == 0) {
            if (fluid && fn != 2) {
                de_adj[slot] = alu_c;
                adj |= 1 << slot;
            }
            continue;
        }
        llu[slot] = np;
        if (fluid) {
            alu[slot] = alu_c;
        } else if (abs((int)g.nb[c]) - 1 == l) {
            // surface row, normal axis: linear extrapolation through sp
            const int m2 = (slot + 1) & 1;  // mod(ll,2), ll = slot+1
            const int ii = p[l];
            const double sp = g.sp_row[r];
            const double fss = (double)(1 - 2 * m2) * (g.xf[(size_t)l * g.stride + ii - m2 + 1] - sp) /
                               g.dx[(size_t)l * g.stride + ii];
            const double hh = g.dx[(size_t)l * g.stride + ii];
            const double hm2 = g.dx[(size_t)l * g.stride + ii - 2 * m2 + 1];
            alu[slot] = (fss - 0.5) * hh / (fss * hh + hm2 * 0.5) * (-1.0);
        } else {
            alu[slot] = -1.0e-20;
        }
    }
    double de = fluid ? de_c : 1.0;
#pragma unroll
    for (int slot = 0; slot < 6; ++slot)
        if (adj & (1 << slot)) de = de + de_adj[slot];
    // right-hand side, fluid rows only
    double b = 0.0;
    if (fluid) {
        double rn;
        if (fbc != 0.0) {
            const size_t n = (size_t)g.ncell;
            auto A = [&](int l, long cc) { return live(g, cc) ? g.a[l * n + cc] : 0.0; };
            auto U = [&](int l, long cc) { return live(g, cc) ? g.un[l * n + cc] : 0.0; };
            const long cx = c + 1, cy = c + X, cz = c + (long)X * Y;
            rn = ((U(0, cx) * A(0, cx) - g.un[c] * g.a[c]) / g.dx[p[0]] +
                  (U(1, cy) * A(1, cy) - g.un[n + c] * g.a[n + c]) / g.dx[(size_t)g.stride + p[1]] +
                  (U(2, cz) * A(2, cz) - g.un[2 * n + c] * g.a[2 * n + c]) /
                      g.dx[2 * (size_t)g.stride + p[2]]) /
                 fbc;
        } else {
            rn = -10000.0;
        }
        double ff = 1.0 - g.F[c];
        ff = ff > 0.001 ? 0.001 : (ff < -0.001 ? -0.001 : ff);
        const double dr1 = rn + ff / g.dt;
        const double q = fabs(dr1 / g.dt);
        if (!(q <= 1e-15 && q != 0.0)) b = g.mp * dr1 / g.dt;
    }
    B[r] = b;
    if (!FUSED || AL0) {
        double *oA = FUSED ? AL0 : AL;
        int *oL = FUSED ? LL0 : LL;
        double *oD = FUSED ? D0 : D;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            oA[(size_t)s * lda + r] = alu[s];
            oL[(size_t)s * lda + r] = llu[s];
        }
        oD[r] = de;
    }
    if (!FUSED) return;
    // PCGPME row scaling + ghost indices + exchange sort (pcgpme_para.F90:77-112)
    const int I = r + 1;
    double w = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        if (llu[s] == 0) llu[s] = g.N + 1 + I % 64;
        w = w + fabs(alu[s]);
    }
    const double m = (w > de) ? w : de;
    w = P2 / (m + de * P1);
#pragma unroll
    for (int s = 0; s < 6; ++s) alu[s] = alu[s] * w;
    D[r] = de * w;
    W[r] = w;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
#pragma unroll
        for (int t = s + 1; t < 6; ++t) {
            if (fabs(alu[s]) < fabs(alu[t])) {
                const double tv = alu[s];
                const int ti = llu[s];
                alu[s] = alu[t];
                llu[s] = llu[t];
                alu[t] = tv;
                llu[t] = ti;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        AL[(size_t)s * lda + r] = alu[s];
        LL[(size_t)s * lda + r] = llu[s];
    }
}

// ---------------------------------------------- pressure-correction pass --
// p(mno2mn(r)) -= xe(r)  (src2LOWNS/1sub-momentum.F90:53-58)
__global__ void __launch_bounds__(256)
k_pressure_update(int N, const int32_t *__restrict__ row2cell, const double *__restrict__ xe,
                  double *__restrict__ p) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) p[row2cell[r]] -= xe[r];
}

// estimate_velocity2 (src2LOWNS/1sub-momentum.F90:590-651):
// un(l,c) = un + dt*(pm - pc)/dxc with pc = -xe(row of c), pm = -xe(row of
// the low neighbour) (0 for an unnumbered cell: the reference's
// never-written xe(0)), dxc = (dx_i + dx_{i-1})/2 (momentum_d,
// lib/src/math_func.F90:251-255), over the cells that are neither air nor
// objects other than pressure-boundary cells (:600-602):
//   a pressure-boundary cell (object code 4): every axis whose low neighbour
//     is fluid or surface (:611-623);
//   a fluid or surface cell: every open low face (a != 0) whose neighbour is
//     fluid or surface (then |un| < 1e-15 -> 0) or a pressure-boundary cell
//     (:624-646).
// A low neighbour below the box is the sentinel (flag -1, code 0).
__global__ void __launch_bounds__(256)
k_velocity_correction(GridDev g, const double *__restrict__ xe, double *__restrict__ un) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ncell) return;
    const int fc = g.nf[c];
    if (fc == 0) return;
    const bool pbc = fc == -1 && g.nb[c] == 4;
    if (fc == -1 && !pbc) return;
    const int X = g.nx + 2, Y = g.ny + 2;
    const int p[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
    const long step[3] = {1, X, (long)X * Y};
    const int nc = g.mno[c];
    const double xec = -(nc ? xe[nc - 1] : 0.0);
    for (int l = 0; l < 3; ++l) {
        const size_t o = (size_t)l * g.ncell + c;
        if (!pbc && g.a[o] == 0.0) continue;
        if (p[l] < 1) continue;
        const long m = c - step[l];
        const int fm = g.nf[m];
        const bool flow = fm >= 1;
        if (!flow && (pbc || !(fm == -1 && g.nb[m] == 4))) continue;
        const double dxc = (g.dx[(size_t)l * g.stride + p[l]] + g.dx[(size_t)l * g.stride + p[l] - 1]) * 0.5;
        const int nm = g.mno[m];
        const double pm = -(nm ? xe[nm - 1] : 0.0);
        double v = un[o] + g.dt * (pm - xec) / dxc;
        if (!pbc && flow && fabs(v) < 1e-15) v = 0.0;
        un[o] = v;
    }
}

// Right-hand side and DMAX of a pressure-correction pass for the fixed row
// numbering (cal_err_of_continuity, 1sub-momentum.F90:315-359); per-block
// maxima of |DR1| go to part.
__global__ void __launch_bounds__(256)
k_rhs(GridDev g, double *__restrict__ B, double *__restrict__ part) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double dm = 0.0;
    if (r < g.N) {
        const long c = g.row2cell[r];
        double b = 0.0;
        if (g.nf[c] == 1) {
            const int X = g.nx + 2, Y = g.ny + 2;
            const int p[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
            const double fbc = g.fb[c];
            double rn;
            if (fbc != 0.0) {
                const size_t n = (size_t)g.ncell;
                auto A = [&](int l, long cc) { return live(g, cc) ? g.a[l * n + cc] : 0.0; };
                auto U = [&](int l, long cc) { return live(g, cc) ? g.un[l * n + cc] : 0.0; };
                const long cx = c + 1, cy = c + X, cz = c + (long)X * Y;
                rn = ((U(0, cx) * A(0, cx) - g.un[c] * g.a[c]) / g.dx[p[0]] +
                      (U(1, cy) * A(1, cy) - g.un[n + c] * g.a[n + c]) / g.dx[(size_t)g.stride + p[1]] +
                      (U(2, cz) * A(2, cz) - g.un[2 * n + c] * g.a[2 * n + c]) / g.dx[2 * (size_t)g.stride + p[2]]) /
                     fbc;
            } else {
                rn = -10000.0;
            }
            double ff = 1.0 - g.F[c];
            ff = ff > 0.001 ? 0.001 : (ff < -0.001 ? -0.001 : ff);
            const double dr1 = rn + ff / g.dt;
            const double q = fabs(dr1 / g.dt);
            if (!(q <= 1e-15 && q != 0.0)) b = g.mp * dr1 / g.dt;
            dm = fabs(dr1);
        }
        B[r] = b;
    }
    __shared__ double lds[256];
    lds[threadIdx.x] = dm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// ------------------------------------------- gradients of the open ratios --
// dL/da and dL/dfb through the assembly above, given lambda = A^{-T} dL/dx
// and x for the operator and right-hand side assembled from g (the adjoint
// of init_matrix's cal_alu_f, 1sub-matrix.F90:50-61,82-87, of
// cal_matrix_element's diagonal, 1sub-momentum.F90:486-509, and of the RHS,
// 1sub-momentum.F90:315-346).  Only fluid rows depend on them: a surface
// row's D is 1 and its coefficients come from sp or are the constant rate.
// With gAL(r,s) = -lambda_r x(LL(r,s)), gD(r) = -lambda_r x_r and, per open
// slot, ALU_C = -mp 2/h/(h+hm) a/fb_c:
//   dL/dALU_C = gAL (numbered neighbour) + gD dD/dALU_C, where
//   dD/dALU_C = -1, or 0 for the Neumann slots (air or unnumbered
//   non-surface neighbour: D -= ALU_C, then D += ALU_C);
//   dALU_C/da = -mp 2/h/(h+hm)/fb_c,  dALU_C/dfb_c = -ALU_C/fb_c;
//   db/da = +-(mp/dt) u/dx/fb_c (high/low face), db/dfb_c = -(mp/dt) rn/fb_c
//   (zero where the reference zeroes b, |DR1/dt| <= 1e-15).
// The flags, the row numbering and which faces are open are fixed by the
// classification: their dependence on a and fb is discrete.
struct RhsRow {
    double rn;   // renzoku_n (divergence / fb)
    bool live;   // b = mp DR1/dt (not zeroed)
};

__device__ __forceinline__ RhsRow rhs_row(const GridDev &g, long c, const int (&p)[3]) {
    const int X = g.nx + 2, Y = g.ny + 2;
    const size_t n = (size_t)g.ncell;
    const double fbc = g.fb[c];
    double rn = -10000.0;
    if (fbc != 0.0) {
        auto A = [&](int l, long cc) { return live(g, cc) ? g.a[l * n + cc] : 0.0; };
        auto U = [&](int l, long cc) { return live(g, cc) ? g.un[l * n + cc] : 0.0; };
        const long cx = c + 1, cy = c + X, cz = c + (long)X * Y;
        rn = ((U(0, cx) * A(0, cx) - g.un[c] * g.a[c]) / g.dx[p[0]] +
              (U(1, cy) * A(1, cy) - g.un[n + c] * g.a[n + c]) / g.dx[(size_t)g.stride + p[1]] +
              (U(2, cz) * A(2, cz) - g.un[2 * n + c] * g.a[2 * n + c]) / g.dx[2 * (size_t)g.stride + p[2]]) /
             fbc;
    }
    double ff = 1.0 - g.F[c];
    ff = ff > 0.001 ? 0.001 : (ff < -0.001 ? -0.001 : ff);
    const double dr1 = rn + ff / g.dt;
    const double q = fabs(dr1 / g.dt);
    return {rn, fbc != 0.0 && !(q <= 1e-15 && q != 0.0)};
}

// dL/dALU_C of cell q's slot toward neighbour nb (0 for a closed slot) and
// dALU_C/da; q a fluid row.
__device__ __forceinline__ void slot_grad(const GridDev &g, long q, long nb, long pc, int l, int pq, int pn,
                                          double lam, double xq, const double *__restrict__ x, double &galu,
                                          double &dalu_da, double &alu) {
    galu = dalu_da = alu = 0.0;
    const int fn = g.nf[nb];
    const double af = live(g, pc) ? g.a[(size_t)l * g.ncell + pc] : 0.0;
    const double fbn = live(g, nb) ? g.fb[nb] : 0.0;
    if ((fn == -1 && g.nb[nb] != 4) || af == 0.0 || fbn == 0.0) return;  // LL_C = 0
    const double h = g.dx[(size_t)l * g.stride + pq];
    const double hm = g.dx[(size_t)l * g.stride + pn];
    const double fbq = g.fb[q];
    dalu_da = -g.mp * (2.0 / h / (h + hm) / fbq);
    alu = -g.mp * (2.0 / h / (h + hm) * af / fbq);
    const int np = g.mno[nb];
    const bool numbered = !(fn == 0 || np == 0);
    const double gD = -lam * xq;
    if (numbered)
        galu = -lam * x[np - 1] + gD * -1.0;
    else
        galu = gD * (fn != 2 ? 0.0 : -1.0);
}

// dL/da per face: face (l, pc) is the low face of cell pc; the cells on its
// two sides are cm = pc - step[l] (its high face) and pc (its low face).
__global__ void __launch_bounds__(256)
k_grad_open_faces(GridDev g, const double *__restrict__ lam, const double *__restrict__ x, double *__restrict__ ga) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * g.ncell) return;
    const int l = (int)(t / g.ncell);
    const long pc = t - (long)l * g.ncell;
    const int X = g.nx + 2, Y = g.ny + 2;
    const int p[3] = {(int)(pc % X), (int)((pc / X) % Y), (int)(pc / ((long)X * Y))};
    const long step[3] = {1, X, (long)X * Y};
    double d = 0.0;
    if (p[l] >= 1 && live(g, pc)) {
        const long cm = pc - step[l];
        for (int side = 0; side < 2; ++side) {
            const long q = side == 0 ? cm : pc;      // the row's cell
            const long nb = side == 0 ? pc : cm;     // its neighbour across the face
            if (g.nf[q] != 1) continue;
            const int r = g.mno[q];
            if (r == 0) continue;
            const double lr = lam[r - 1], xr = x[r - 1];
            const int pq = side == 0 ? p[l] - 1 : p[l];
            const int pn = side == 0 ? p[l] : p[l] - 1;
            double galu, dalu_da, alu;
            slot_grad(g, q, nb, pc, l, pq, pn, lr, xr, x, galu, dalu_da, alu);
            d += galu * dalu_da;
            // the right-hand side of row q
            int pqv[3] = {p[0], p[1], p[2]};
            pqv[l] = pq;
            const RhsRow rr = rhs_row(g, q, pqv);
            if (rr.live) {
                const double u = g.un[(size_t)l * g.ncell + pc];
                const double dxq = g.dx[(size_t)l * g.stride + pq];
                d += lr * (g.mp / g.dt) * ((side == 0 ? u : -u) / dxq / g.fb[q]);
            }
        }
    }
    ga[t] = d;
}

// dL/dfb per cell: the cell's own row only (a neighbour's fb decides only
// whether a slot is open).
__global__ void __launch_bounds__(256)
k_grad_open_cells(GridDev g, const double *__restrict__ lam, const double *__restrict__ x,
                  double *__restrict__ gfb) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.ncell) return;
    double d = 0.0;
    const int r = g.mno[c];
    if (g.nf[c] == 1 && r != 0) {
        const int X = g.nx + 2, Y = g.ny + 2;
        const int p[3] = {(int)(c % X), (int)((c / X) % Y), (int)(c / ((long)X * Y))};
        const long step[3] = {1, X, (long)X * Y};
        const double lr = lam[r - 1], xr = x[r - 1], fbc = g.fb[c];
        for (int slot = 0; slot < 6; ++slot) {
            const int l = slot >> 1, side = (slot & 1) ? 1 : -1;
            const long nb = c + side * step[l];
            const long pc = side > 0 ? nb : c;
            double galu, dalu_da, alu;
            slot_grad(g, c, nb, pc, l, p[l], p[l] + side, lr, xr, x, galu, dalu_da, alu);
            d += galu * (-alu / fbc);
        }
        const RhsRow rr = rhs_row(g, c, p);
        if (rr.live) d += lr * (g.mp / g.dt) * (-rr.rn / fbc);
    }
    gfb[c] = d;
}

__global__ void __launch_bounds__(1024) k_max_fin(const double *__restrict__ part, int G, double *__restrict__ out) {
    __shared__ double lds[1024];
    double m = 0.0;
    for (int k = threadIdx.x; k < G; k += 1024) m = fmax(m, part[k]);
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = lds[0];
}

} // namespace clowns
