// pcg_kernels.hpp -- gfx950 device kernels of the PCGPME solve.
//
// Reference: src2LOWNS/pcgpme_para.F90 (PCGPME, LDUSUB, AXSUB).  Every pass
// of the reference's OpenMP loops maps to one streaming kernel here:
//   k_scale_sort   row scaling + ghost indices + per-row exchange sort (:77-112)
//   k_spmv_v2      AXSUB, y = A x in sorted slot order (:323-352), two rows
//                  a thread; k_spmv_rows32 its FP32 form from row-major
//                  copies (mixed inner solve), k_spmv_rows32_p2 that form
//                  in two chunks a wavefront
//   k_pm_v2        one stage of LDUSUB (:279-318); the last stage carries the
//                  dot products that follow it ((r^,q) :157-160,
//                  (e,v),(v,v) :186-191); k_pm_v4, k_pm2_v2 its FP32
//                  four-row and fused-pair forms, k_pm_p2 its FP64 form
//                  with the row-pattern table
//   k_pupd_vec     p = r + beta (p - omega q) (:244)
//   k_eupd_vec     e = r - alpha q (:173)
//   k_update_vec   x += alpha p + omega e; r = e - omega v; (r,r), (r^,r)
//                  (:203-211)
//   k_fin_*        the scalar steps (alpha, omega, ERR, exits, beta), one
//                  block, deterministic order
// Layout: matrix columns are structure-of-arrays with column stride lda
// (identical to the reference's AL(NA,NL)/LL(NA,NL)); LL stays 1-based, so a
// gather reads x[LL-1]; every vector has N+64 entries whose last 64 (the
// ghost rows N+1+mod(I,64), :80) are zero.
// Arithmetic: expressions keep the Fortran's left-to-right order; the
// library is compiled with -ffp-contract=off so single FP64 kernels agree
// bit for bit with the reference.  Reductions: per-thread sums in row order,
// a fixed wavefront/block tree, per-block partials combined by one block in
// fixed order, so a solve is deterministic run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace clowns {

constexpr int BS = 256;        // threads per block (4 wavefronts of 64)
constexpr int FIN_BS = 1024;   // threads of the one-block finalize kernels
constexpr int PAD = 64;        // ghost rows after N (pcgpme_para.F90:80)

// Scalars of one Bi-CGSTAB run, resident in device memory.  Names follow
// pcgpme_para.F90:47.
struct Scal {
    double BN, C1, RQ, ALP, AMU, SS, RN, ERR, ERRP, BETA, EPS;
    double RN0;        // refinement: ||r|| of the outer residual (scale)
    int K;             // last completed iteration
    int itermax;
    int done;          // 1 = converged / exited: every later kernel returns
    int ier;
    int kout;          // ITER returned (the reference returns ITER+1 on 3000)
    int beta_prod;     // 1: BETA = C1/(AMU*RQ) (pcgpme_2DH.f90:241), else C1/AMU/RQ
    int pad_[2];
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Giving XCD x the contiguous logical blocks [x*G/8, (x+1)*G/8) keeps the
// +-1, +-row and +-plane neighbours of a row in the L2 that streamed them.
// G is a multiple of 8.
__device__ __forceinline__ int logical_block() {
    const int b = blockIdx.x;
    return (b & 7) * (gridDim.x >> 3) + (b >> 3);
}

// Row block of a gather kernel's workgroup: the XCD-contiguous logical block,
// mapped through `order` (a permutation of the row blocks) when the solver
// runs the operator in a block order of its own.  Results do not depend on
// it: every row's arithmetic, and the dot partials' index, follow the row
// block.
__device__ __forceinline__ int row_block(const int *order) {
    const int lb = logical_block();
    return order ? order[lb] : lb;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the block, fixed order; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NT / 64; ++k) s += lds[k];
    }
    return s;
}

// ---------------------------------------------------------------- setup --
// a3 + a4: scale row i by W = P2/(max(sum|AL|, D) + P1*D), substitute ghost
// indices for LL = 0, then the reference's (unstable) exchange sort by |AL|.
template <int NL>
__global__ void __launch_bounds__(BS)
k_scale_sort(int N, int lda, double *__restrict__ AL, int *__restrict__ LL,
             double *__restrict__ D, double *__restrict__ Wsc, double P1,
             double P2) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    const int I = i + 1;
    double a[NL];
    int l[NL];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        a[k] = AL[(size_t)k * lda + i];
        l[k] = LL[(size_t)k * lda + i];
        if (l[k] == 0) l[k] = N + 1 + I % 64;
        w = w + fabs(a[k]);
    }
    const double temp = D[i];
    const double m = (w > temp) ? w : temp;
    w = P2 / (m + temp * P1);
#pragma unroll
    for (int k = 0; k < NL; ++k) a[k] = a[k] * w;
    D[i] = temp * w;
    Wsc[i] = w;
#pragma unroll
    for (int k = 0; k < NL - 1; ++k) {
#pragma unroll
        for (int kk = k + 1; kk < NL; ++kk) {
            if (fabs(a[k]) < fabs(a[kk])) {
                const double t = a[k];
                const int it = l[k];
                a[k] = a[kk];
                l[k] = l[kk];
                a[kk] = t;
                l[kk] = it;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        AL[(size_t)k * lda + i] = a[k];
        LL[(size_t)k * lda + i] = l[k];
    }
}

// Caller LL must name rows 1..N or 0 (no neighbour); anything else would be
// gathered out of bounds.  Sets *bad.
__global__ void __launch_bounds__(BS)
k_check_ll(int N, int lda, int NL, const int *__restrict__ LL, int *__restrict__ bad) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    int ok = 1;
    for (int k = 0; k < NL; ++k) {
        const int v = LL[(size_t)k * lda + i];
        ok &= (v >= 0 && v <= N) ? 1 : 0;
    }
    if (!ok) atomicOr(bad, 1);
}

// FP32 copy of the scaled, sorted FP64 operator (the mixed solve's inner
// matrix); LL is shared.
__global__ void __launch_bounds__(BS)
k_to_f32(int N, int lda, int NL, const double *__restrict__ AL,
         const double *__restrict__ D, float *__restrict__ AL32,
         float *__restrict__ D32) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    for (int k = 0; k < NL; ++k)
        AL32[(size_t)k * lda + i] = (float)AL[(size_t)k * lda + i];
    D32[i] = (float)D[i];
}

// Half-precision copy of the scaled, sorted PM coefficients (the mixed
// solve's inner preconditioner, ldc2).
__global__ void __launch_bounds__(BS)
k_to_f16(int N, int lda, int NL, const double *__restrict__ AL, _Float16 *__restrict__ ALh) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    for (int k = 0; k < NL; ++k) ALh[(size_t)k * lda + i] = (_Float16)(float)AL[(size_t)k * lda + i];
}

// b_s = W o b (pcgpme_para.F90:92); x0 into the padded solve vector.
__global__ void __launch_bounds__(BS)
k_scale_b(int N, const double *__restrict__ b, const double *__restrict__ Wsc,
          double *__restrict__ bs, const double *__restrict__ x_in,
          double *__restrict__ x) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    bs[i] = b[i] * Wsc[i];
    if (x) x[i] = x_in ? x_in[i] : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(BS) k_fill(size_t n, T *__restrict__ v, T val) {
    const size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    if (i < n) v[i] = val;
}

template <typename T>
__global__ void __launch_bounds__(BS)
k_copy(int N, const T *__restrict__ a, T *__restrict__ b) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < N) b[i] = a[i];
}

// ------------------------------------------------ two rows per thread --
// The same PM stages with each thread owning two consecutive rows and
// loading both with one 8-byte (FP32) or 16-byte (FP64) access: twice the
// bytes in flight per memory instruction, which the FP32 stages need to get
// near the copy rate.  Per-row arithmetic is unchanged (bit-identical rows);
// only the dot products' summation order differs from the one-row kernels.
// Column starts and vectors must be 2-element aligned (lda even).
template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
template <> struct V2<int> { using t = int2; };

template <typename T>
__device__ __forceinline__ typename V2<T>::t ld2(const T *p) {
    return *reinterpret_cast<const typename V2<T>::t *>(p);
}

template <typename T> struct NV2;
template <> struct NV2<float> { typedef float t __attribute__((ext_vector_type(2))); };
template <> struct NV2<double> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct NV2<int> { typedef int t __attribute__((ext_vector_type(2))); };
template <> struct NV2<short> { typedef short t __attribute__((ext_vector_type(2))); };
// Two consecutive elements of an array the kernel streams once (operator
// values, index codes, the dot vector, outputs).  ST: non-temporal loads and
// stores, which keep those streams from displacing the gathered vector in
// L2 -- faster once the operator no longer fits the 256 MB last-level cache,
// slower while it does (the non-temporal accesses are not kept there
// either); the solver times both per operator (tune_stream).
template <bool ST, typename T>
__device__ __forceinline__ typename NV2<T>::t ldn2(const T *p) {
    const typename NV2<T>::t *q = reinterpret_cast<const typename NV2<T>::t *>(p);
    if (ST) return __builtin_nontemporal_load(q);
    return *q;
}
// Any vector (or scalar) the kernel streams once, with the same choice.
template <bool ST, typename V>
__device__ __forceinline__ V ldv(const V *p) {
    if (ST) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool ST, typename V>
__device__ __forceinline__ void stv(V *p, V v) {
    if (ST)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
typedef float f4v __attribute__((ext_vector_type(4)));
// two floats at any 4-byte boundary (a dword-aligned 8-byte load)
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));

template <bool ST, typename T>
__device__ __forceinline__ typename V2<T>::t lds2(const T *p) {
    const typename NV2<T>::t v = ldn2<ST>(p);
    typename V2<T>::t o;
    o.x = v.x;
    o.y = v.y;
    return o;
}
template <bool ST, typename T>
__device__ __forceinline__ void sts2(T *p, typename V2<T>::t o) {
    typename NV2<T>::t v;
    v.x = o.x;
    v.y = o.y;
    if (ST)
        __builtin_nontemporal_store(v, reinterpret_cast<typename NV2<T>::t *>(p));
    else
        *reinterpret_cast<typename NV2<T>::t *>(p) = v;
}

// Two consecutive PM coefficients as T.  C = T, or C = _Float16: the
// half-precision copy of the preconditioner's coefficients that the FP32
// inner solve of the mixed form may use (the preconditioner there need only
// approximate the FP64 one: the outer FP64 loop measures the true residual
// with the exact one).
template <typename T, bool ST, typename C>
__device__ __forceinline__ typename V2<T>::t ldc2(const C *p) {
    if constexpr (std::is_same<C, T>::value) {
        return lds2<ST>(p);
    } else {
        typedef C c2 __attribute__((ext_vector_type(2)));
        const c2 *q = reinterpret_cast<const c2 *>(p);
        c2 v;
        if (ST)
            v = __builtin_nontemporal_load(q);
        else
            v = *q;
        typename V2<T>::t o;
        o.x = (T)v.x;
        o.y = (T)v.y;
        return o;
    }
}

// Neighbour index sources for the two-row kernels, 0-based, ghosts (the
// reference's N+1+mod(I,64), pcgpme_para.F90:80) at the zero rows >= N.
// Ix32: the sorted LL itself.  Ix16: the same indices coded in 16 bits per
// slot (k_ll16_build): an offset from the row when |offset| < 16384, else an
// offset from the row block's nearest far offset of that sign (z
// neighbours, up to a plane away), or the ghost code, or -- for the rare
// slot whose far offset spreads too far from its block's (under a ragged
// surface) -- the wide code, which reads that slot from the 32-bit LL.  Half
// the index bytes of every SpMV and PM stage; the gathered element is the
// same.
struct Ix32 {
    const int *L;
    __device__ __forceinline__ Ix32 col(size_t off) const { return {L + off}; }
    template <bool ST = false> __device__ __forceinline__ int2 pair(int i) const {
        const NV2<int>::t j = ldn2<ST>(L + i);
        return make_int2(j.x - 1, j.y - 1);
    }
    // four consecutive rows (i a multiple of 4): one 16-byte load
    template <bool ST = false> __device__ __forceinline__ int4 quad(int i) const {
        typedef int i4 __attribute__((ext_vector_type(4)));
        const i4 j = ldv<ST>(reinterpret_cast<const i4 *>(L + i));
        return make_int4(j.x - 1, j.y - 1, j.z - 1, j.w - 1);
    }
    __device__ __forceinline__ int one(int i) const { return L[i] - 1; }
    // Phased form (k_spmv_v2): the pair's raw codes, then 1-based indices.
    typedef NV2<int>::t Raw;
    typedef int Base;
    __device__ __forceinline__ Base blk(int) const { return 0; }
    // With the code block's base passed in (see Ix16::pair_b).
    __device__ __forceinline__ Base blk_at(int, int) const { return 0; }
    template <bool ST = false> __device__ __forceinline__ int2 pair_b(int i, Base) const { return pair<ST>(i); }
    template <bool ST = false> __device__ __forceinline__ int4 quad_b(int i, Base) const { return quad<ST>(i); }
    template <bool ST = false> __device__ __forceinline__ Raw raw(int i) const { return ldn2<ST>(L + i); }
    __device__ __forceinline__ int2 dec1(Raw v, int, Base) const { return make_int2(v.x, v.y); }
    static constexpr bool TABLE = false;
};

constexpr int LL16_ROWS = 512;  // rows per base block (two-row kernels: one block)
constexpr int LL16_GHOST = -32768;
constexpr int LL16_WIDE = -32767;

struct Ix16 {
    static constexpr int ROWS = LL16_ROWS;
    static constexpr int WIDE = LL16_WIDE;
    static constexpr bool TABLE = false;
    const short *L;
    const int *L32;  // the same column, 32-bit (slots coded LL16_WIDE)
    const int2 *base;
    int N;
    __device__ __forceinline__ Ix16 col(size_t off) const { return {L + off, L32 + off, base, N}; }
    // Decoded indices are clamped to [0, N] (N: a zero ghost row), so a
    // coding error can only give a wrong value, never an access outside the
    // vectors.
    // Selects rather than branches for the common codes.
    __device__ __forceinline__ int dec(int v, int i, int2 b) const {
        int j = i + v;
        j = v >= 16384 ? i + b.x + (v - 16384) : j;
        j = v < -16384 ? i + b.y + (v + 16385) : j;
        j = v == LL16_GHOST ? N : j;
        if (v == LL16_WIDE) j = L32[i] - 1;
        return min(max(j, 0), N);
    }
    template <bool ST = false> __device__ __forceinline__ int2 pair(int i) const {
        const int2 b = base[i / LL16_ROWS];
        const NV2<short>::t v = ldn2<ST>(L + i);
        return make_int2(dec(v.x, i, b), dec(v.y, i + 1, b));
    }
    // four consecutive rows (i a multiple of 4): one 8-byte load
    template <bool ST = false> __device__ __forceinline__ int4 quad(int i) const {
        const int2 b = base[i / LL16_ROWS];
        typedef short s4 __attribute__((ext_vector_type(4)));
        const s4 v = ldv<ST>(reinterpret_cast<const s4 *>(L + i));
        return make_int4(dec(v.x, i, b), dec(v.y, i + 1, b), dec(v.z, i + 2, b), dec(v.w, i + 3, b));
    }
    __device__ __forceinline__ int one(int i) const { return dec(L[i], i, base[i / LL16_ROWS]); }
    // pair() and quad() with the code block's base passed in: a kernel whose
    // workgroup (or wavefront) rows lie in one code block reads it once, as
    // a scalar load at a uniform block index, instead of a per-lane load of
    // the same element that the stream loads must wait behind.  blk_at
    // reads block b's base when the block holds rows (b ROWS < N).
    __device__ __forceinline__ int2 blk_at(int b, int n) const {
        return (long)b * LL16_ROWS < n ? base[b] : make_int2(0, 0);
    }
    template <bool ST = false> __device__ __forceinline__ int2 pair_b(int i, int2 b) const {
        const NV2<short>::t v = ldn2<ST>(L + i);
        return make_int2(dec(v.x, i, b), dec(v.y, i + 1, b));
    }
    template <bool ST = false> __device__ __forceinline__ int4 quad_b(int i, int2 b) const {
        typedef short s4 __attribute__((ext_vector_type(4)));
        const s4 v = ldv<ST>(reinterpret_cast<const s4 *>(L + i));
        return make_int4(dec(v.x, i, b), dec(v.y, i + 1, b), dec(v.z, i + 2, b), dec(v.w, i + 3, b));
    }
    // Phased form for kernels that load several slots' codes before using
    // any (k_spmv_v2): raw() the pair's codes, dec1() their indices 1-based
    // (clamped to [1, N + 1]).  A wide slot's 32-bit index is loaded into
    // the result without being touched, so the wait for it falls at the
    // gather, after every slot's wide loads are in flight, not one slot at
    // a time.
    typedef NV2<short>::t Raw;
    typedef int2 Base;
    __device__ __forceinline__ Base blk(int i) const { return base[i / LL16_ROWS]; }
    template <bool ST = false> __device__ __forceinline__ Raw raw(int i) const { return ldn2<ST>(L + i); }
    __device__ __forceinline__ int near1(int v, int i, int2 b) const {
        int j = i + v;
        j = v >= 16384 ? i + b.x + (v - 16384) : j;
        j = v < -16384 ? i + b.y + (v + 16385) : j;
        j = v == LL16_GHOST ? N : j;
        return min(max(j, 0), N) + 1;
    }
    __device__ __forceinline__ int2 dec1(Raw v, int i, Base b) const {
        int2 j = make_int2(near1(v.x, i, b), near1(v.y, i + 1, b));
        if (v.x == LL16_WIDE) j.x = L32[i];
        if (v.y == LL16_WIDE) j.y = L32[i + 1];
        return j;
    }
};

// Build the 16-bit form of the sorted LL: one block per LL16_ROWS rows finds
// its nearest far offsets of each sign, then codes every slot; *nwide counts
// the slots left on the 32-bit LL.  *bad = 1 for an index outside [1, N]
// that is not a ghost (cannot happen after the checks and the sort; the
// operator then keeps the 32-bit path).
__global__ void __launch_bounds__(BS)
k_ll16_build(int N, int lda, int NL, const int *__restrict__ LL, short *__restrict__ L16,
             int2 *__restrict__ base, int *__restrict__ bad, int *__restrict__ nwide) {
    const int r0 = blockIdx.x * LL16_ROWS;
    int bp = 0x7fffffff, bn = -0x7fffffff - 1;
    for (int u = threadIdx.x; u < LL16_ROWS; u += BS) {
        const int i = r0 + u;
        if (i >= N) break;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            if (j < 1 || j > N) continue;
            const int off = j - 1 - i;
            if (off >= 16384) bp = min(bp, off);
            if (off < -16384) bn = max(bn, off);
        }
    }
    __shared__ int lp[BS], ln[BS];
    lp[threadIdx.x] = bp;
    ln[threadIdx.x] = bn;
    __syncthreads();
    for (int st = BS / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            lp[threadIdx.x] = min(lp[threadIdx.x], lp[threadIdx.x + st]);
            ln[threadIdx.x] = max(ln[threadIdx.x], ln[threadIdx.x + st]);
        }
        __syncthreads();
    }
    bp = lp[0];
    bn = ln[0];
    int wide = 0, invalid = 0;
    for (int u = threadIdx.x; u < LL16_ROWS; u += BS) {
        const int i = r0 + u;
        if (i >= N) break;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            int v;
            if (j > N) {
                v = LL16_GHOST;
            } else if (j < 1) {
                v = 0;
                invalid = 1;
            } else {
                const int off = j - 1 - i;
                if (off >= 16384) {
                    const int d = off - bp;
                    v = d <= 16383 ? 16384 + d : LL16_WIDE;
                } else if (off >= -16384) {
                    v = off;
                } else {
                    const int d = bn - off;
                    v = d <= 16381 ? -16385 - d : LL16_WIDE;
                }
                wide += v == LL16_WIDE;
            }
            L16[(size_t)k * lda + i] = (short)v;
        }
    }
    if (invalid) atomicOr(bad, 1);
    if (wide) atomicAdd(nwide, wide);
    if (threadIdx.x == 0) base[blockIdx.x] = make_int2(bp, bn);
}

// Ix8: the sorted LL in 8 bits per slot.  Each 512-row block sorts its
// neighbour offsets into four classes -- near and far, each sign -- with the
// split between near and far at the geometric mean of the block's smallest
// and largest non-unit offset magnitudes (the grid-line and plane offsets of
// a box grid: the classes stand in for y-, y+, z-, z+, which the sorted slots
// no longer name).  A class codes its offsets as 6-bit steps from its
// smallest member.  Codes: class*64 + step (far+: steps 0..59), 252/253 the
// offsets -1/+1, 254 a ghost, 255 wide (the slot is read from the 32-bit
// LL).  On the box configurations 0.2-0.5% of slots are wide.  Same gathered
// element as Ix32 and Ix16; half the index bytes of Ix16.
constexpr int LL8_ROWS = 512;
constexpr int LL8_XM = 252, LL8_XP = 253, LL8_GHOST = 254, LL8_WIDE = 255;

struct Ix8 {
    static constexpr int ROWS = LL8_ROWS;
    static constexpr int WIDE = LL8_WIDE;
    const uint8_t *L;
    const int *L32;   // the same column, 32-bit (slots coded LL8_WIDE)
    const int4 *base; // per block: window starts of near-, near+, far-, far+
    int N;
    __device__ __forceinline__ Ix8 col(size_t off) const { return {L + off, L32 + off, base, N}; }
    // Decoded indices are clamped to [0, N] (N: a zero ghost row).
    __device__ __forceinline__ int dec(int c, int i, int4 b) const {
        const int cls = c >> 6;
        const int bs = cls == 0 ? b.x : (cls == 1 ? b.y : (cls == 2 ? b.z : b.w));
        int j = i + bs + (c & 63);
        j = c == LL8_XM ? i - 1 : j;
        j = c == LL8_XP ? i + 1 : j;
        j = c == LL8_GHOST ? N : j;
        if (c == LL8_WIDE) j = L32[i] - 1;
        return min(max(j, 0), N);
    }
    template <bool ST = false> __device__ __forceinline__ int2 pair(int i) const {
        const int4 b = base[i / LL8_ROWS];
        const uint16_t v = ST ? __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(L + i))
                              : *reinterpret_cast<const uint16_t *>(L + i);
        return make_int2(dec(v & 255, i, b), dec(v >> 8, i + 1, b));
    }
    // four consecutive rows (i a multiple of 4): one 4-byte load
    template <bool ST = false> __device__ __forceinline__ int4 quad(int i) const {
        const int4 b = base[i / LL8_ROWS];
        const unsigned v = ldv<ST>(reinterpret_cast<const unsigned *>(L + i));
        return make_int4(dec(v & 255, i, b), dec((v >> 8) & 255, i + 1, b), dec((v >> 16) & 255, i + 2, b),
                         dec(v >> 24, i + 3, b));
    }
    __device__ __forceinline__ int one(int i) const { return dec(L[i], i, base[i / LL8_ROWS]); }
    // With the code block's base passed in, as Ix16's.
    __device__ __forceinline__ int4 blk_at(int b, int n) const {
        return (long)b * LL8_ROWS < n ? base[b] : make_int4(0, 0, 0, 0);
    }
    template <bool ST = false> __device__ __forceinline__ int2 pair_b(int i, int4 b) const {
        const uint16_t v = ldv<ST>(reinterpret_cast<const uint16_t *>(L + i));
        return make_int2(dec(v & 255, i, b), dec(v >> 8, i + 1, b));
    }
    template <bool ST = false> __device__ __forceinline__ int4 quad_b(int i, int4 b) const {
        const unsigned v = ldv<ST>(reinterpret_cast<const unsigned *>(L + i));
        return make_int4(dec(v & 255, i, b), dec((v >> 8) & 255, i + 1, b), dec((v >> 16) & 255, i + 2, b),
                         dec(v >> 24, i + 3, b));
    }
    // Phased form, as Ix16's.
    typedef uint16_t Raw;
    typedef int4 Base;
    __device__ __forceinline__ Base blk(int i) const { return base[i / LL8_ROWS]; }
    template <bool ST = false> __device__ __forceinline__ Raw raw(int i) const {
        return ldv<ST>(reinterpret_cast<const uint16_t *>(L + i));
    }
    __device__ __forceinline__ int near1(int c, int i, int4 b) const {
        const int cls = c >> 6;
        const int bs = cls == 0 ? b.x : (cls == 1 ? b.y : (cls == 2 ? b.z : b.w));
        int j = i + bs + (c & 63);
        j = c == LL8_XM ? i - 1 : j;
        j = c == LL8_XP ? i + 1 : j;
        j = c == LL8_GHOST ? N : j;
        return min(max(j, 0), N) + 1;
    }
    __device__ __forceinline__ int2 dec1(Raw v, int i, Base b) const {
        const int c0 = v & 255, c1 = v >> 8;
        int2 j = make_int2(near1(c0, i, b), near1(c1, i + 1, b));
        if (c0 == LL8_WIDE) j.x = L32[i];
        if (c1 == LL8_WIDE) j.y = L32[i + 1];
        return j;
    }
    // Decoding through a table in LDS, for kernels that stage one per code
    // block: T[c] is code c's offset from its row (the ghost code's is large
    // enough for the clamp to give N; the wide code's is not used).  One
    // LDS read, an add and a clamp per slot, in place of the selects
    // above; the kernels' pace is set by instruction issue about as much
    // as by bytes.
    static constexpr bool TABLE = true;
    static __device__ __forceinline__ int tab_entry(int c, int4 b) {
        const int cls = c >> 6;
        const int bs = cls == 0 ? b.x : (cls == 1 ? b.y : (cls == 2 ? b.z : b.w));
        int off = bs + (c & 63);
        off = c == LL8_XM ? -1 : off;
        off = c == LL8_XP ? 1 : off;
        off = c == LL8_GHOST ? (1 << 30) : off;
        return off;
    }
    // 0-based index of code c (not the wide code) at row i
    __device__ __forceinline__ int near_t(int c, int i, const int *T) const { return min(max(i + T[c], 0), N); }
};

// Build the 8-bit form of the sorted LL: one block per LL8_ROWS rows finds
// its near/far split and class window starts, then codes every slot; *nwide
// counts wide slots, *bad = 1 for an index outside [1, N] that is not a
// ghost (the operator then keeps the wider codes).
__global__ void __launch_bounds__(BS)
k_ll8_build(int N, int lda, int NL, const int *__restrict__ LL, uint8_t *__restrict__ L8,
            int4 *__restrict__ base, int *__restrict__ bad, int *__restrict__ nwide) {
    __shared__ long long red[BS][4];
    const int r0 = blockIdx.x * LL8_ROWS;
    const long long BIG = 0x7fffffffffffLL;
    auto reduce_min4 = [&](long long (&v)[4]) {
        for (int a = 0; a < 4; ++a) red[threadIdx.x][a] = v[a];
        __syncthreads();
        for (int st = BS / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st)
                for (int a = 0; a < 4; ++a) red[threadIdx.x][a] = min(red[threadIdx.x][a], red[threadIdx.x + st][a]);
            __syncthreads();
        }
        for (int a = 0; a < 4; ++a) v[a] = red[0][a];
        __syncthreads();
    };
    // pass 1: smallest and largest non-unit magnitude (max as min of -m)
    long long m[4] = {BIG, BIG, BIG, BIG};
    for (int u = threadIdx.x; u < LL8_ROWS; u += BS) {
        const int i = r0 + u;
        if (i >= N) break;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            if (j < 1 || j > N) continue;
            const long long mag = llabs((long long)(j - 1 - i));
            if (mag <= 1) continue;
            m[0] = min(m[0], mag);
            m[1] = min(m[1], -mag);
        }
    }
    reduce_min4(m);
    const long long mmin = m[0], mmax = -m[1];
    const double split = (mmin < BIG && mmax > 4 * mmin) ? sqrt((double)mmin * (double)mmax) : (double)mmax + 1.0;
    auto cls_of = [&](int off) -> int {
        const bool far = (double)llabs((long long)off) > split;
        return (far ? 2 : 0) + (off > 0 ? 1 : 0);
    };
    // pass 2: each class's smallest member
    long long cb[4] = {BIG, BIG, BIG, BIG};
    for (int u = threadIdx.x; u < LL8_ROWS; u += BS) {
        const int i = r0 + u;
        if (i >= N) break;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            if (j < 1 || j > N) continue;
            const int off = j - 1 - i;
            if (off == 1 || off == -1) continue;
            const int c = cls_of(off);
            cb[c] = min(cb[c], (long long)off);
        }
    }
    reduce_min4(cb);
    int4 b;
    b.x = cb[0] < BIG ? (int)cb[0] : 0;
    b.y = cb[1] < BIG ? (int)cb[1] : 0;
    b.z = cb[2] < BIG ? (int)cb[2] : 0;
    b.w = cb[3] < BIG ? (int)cb[3] : 0;
    const int bs[4] = {b.x, b.y, b.z, b.w};
    int wide = 0, invalid = 0;
    for (int u = threadIdx.x; u < LL8_ROWS; u += BS) {
        const int i = r0 + u;
        if (i >= N) break;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            int v;
            if (j > N) {
                v = LL8_GHOST;
            } else if (j < 1) {
                v = LL8_GHOST;
                invalid = 1;
            } else {
                const int off = j - 1 - i;
                if (off == -1) {
                    v = LL8_XM;
                } else if (off == 1) {
                    v = LL8_XP;
                } else {
                    const int c = cls_of(off);
                    const long long d = (long long)off - bs[c];
                    v = (d >= 0 && d < (c == 3 ? 60 : 64)) ? c * 64 + (int)d : LL8_WIDE;
                }
                wide += v == LL8_WIDE;
            }
            L8[(size_t)k * lda + i] = (uint8_t)v;
        }
    }
    if (invalid) atomicOr(bad, 1);
    if (wide) atomicAdd(nwide, wide);
    if (threadIdx.x == 0) base[blockIdx.x] = b;
}

template <typename T, int NT, int DOT, typename IX, bool ST, typename C = T>
__global__ void __launch_bounds__(BS)
k_pm_v2(const Scal *__restrict__ sc, int N, const C *__restrict__ a0, IX l0,
        const C *__restrict__ a1, IX l1, const T *__restrict__ in,
        T *__restrict__ out, const T *__restrict__ w, double *__restrict__ part,
        const int *__restrict__ order) {
    const int done = sc->done;  // read together with the block order
    const int lb = row_block(order);
    if (done) return;
    const int i = (lb * BS + threadIdx.x) * 2;
    const auto bb = l0.blk_at(lb, N);  // the row block is the code block; one base array for all columns
    double s0 = 0.0, s1 = 0.0;
    if (i + 1 < N) {
        const auto c0 = ldc2<T, ST>(a0 + i);
        const int2 j0 = l0.template pair_b<ST>(i, bb);
        const auto xi = ld2(in + i);
        T o0 = xi.x - c0.x * in[j0.x];
        T o1 = xi.y - c0.y * in[j0.y];
        if (NT == 2) {
            const auto c1 = ldc2<T, ST>(a1 + i);
            const int2 j1 = l1.template pair_b<ST>(i, bb);
            o0 = o0 - c1.x * in[j1.x];
            o1 = o1 - c1.y * in[j1.y];
        }
        typename V2<T>::t o;
        o.x = o0;
        o.y = o1;
        sts2<ST>(out + i, o);
        if (DOT == 1 || DOT == 2) {
            const auto wi = lds2<ST>(w + i);
            s0 += (double)wi.x * (double)o0;
            s0 += (double)wi.y * (double)o1;
        }
        if (DOT == 2) {
            s1 += (double)o0 * (double)o0;
            s1 += (double)o1 * (double)o1;
        }
        if (DOT == 3) {
            s0 += (double)o0 * (double)o0;
            s0 += (double)o1 * (double)o1;
        }
    } else if (i < N) {
        T o = in[i] - (T)a0[i] * in[l0.one(i)];
        if (NT == 2) o = o - (T)a1[i] * in[l1.one(i)];
        out[i] = o;
        if (DOT == 1 || DOT == 2) s0 += (double)w[i] * (double)o;
        if (DOT == 2) s1 += (double)o * (double)o;
        if (DOT == 3) s0 += (double)o * (double)o;
    }
    if (DOT != 0) {
        __shared__ double lds[BS / 64];
        const double t0 = block_sum<BS>(s0, lds);
        if (DOT == 2) {
            const double t1 = block_sum<BS>(s1, lds);
            if (threadIdx.x == 0) part[gridDim.x + lb] = t1;
        }
        if (threadIdx.x == 0) part[lb] = t0;
    }
}

// ------------------------------------------------------- row patterns --
// After PCGPME's scaling and sort most rows of an assembled pressure
// operator repeat one of a few rows exactly: every interior cell of a
// uniform region gives the same six coefficients and diagonal (at
// 512x512x96 with bathymetry 95.4% of the rows fall on four such rows, at
// 256x256x64 with an obstacle 97.3%).  A per-operator table of up to PAT_MAX
// rows (six coefficients and D, FP64) and a one-byte pattern id per row
// (PAT_NONE: the row's own values) let the SpMV read one byte for those rows
// instead of 56 (FP64) or 24 (FP32) bytes of coefficients.  (The PM stages
// read one coefficient a row: measured on MI355X, an id in its place gains
// nothing in FP64 and 6% in FP32 with 8-bit index codes, lost 7% with
// 16-bit ones; they keep the coefficient columns.)  The values it takes from the table are
// the same bits (or the same FP32 / half roundings of them) the arrays
// hold, so results do not change.
constexpr int PAT_MAX = 16;
constexpr int PAT_NONE = 255;
constexpr int PAT_W = 7;  // values per table row: AL(1..6), D

// Sampled rows of the scaled, sorted operator for the table: row s * stride
// for s < S, its six coefficients and D.
__global__ void __launch_bounds__(BS)
k_pat_sample(int N, int lda, const double *__restrict__ AL, const double *__restrict__ D, int S, int stride,
             double *__restrict__ keys) {
    const int s = blockIdx.x * BS + threadIdx.x;
    if (s >= S) return;
    const long row = (long)s * stride;
    if (row >= N) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) keys[(size_t)s * PAT_W + k] = AL[(size_t)k * lda + row];
    keys[(size_t)s * PAT_W + 6] = D[row];
}

// Pattern id of every row: the table row whose seven values equal the
// row's bit for bit, else PAT_NONE; *nnone counts the latter.  The ids past
// N (up to a multiple of 4 plus 64) are PAT_NONE.
__global__ void __launch_bounds__(BS)
k_pat_assign(int N, int npad, int lda, const double *__restrict__ AL, const double *__restrict__ D,
             const double *__restrict__ tab, int npat, uint8_t *__restrict__ pid, int *__restrict__ nnone) {
    __shared__ unsigned long long t[PAT_MAX * PAT_W];
    if (threadIdx.x < PAT_MAX * PAT_W)
        t[threadIdx.x] = threadIdx.x < npat * PAT_W ? (unsigned long long)__double_as_longlong(tab[threadIdx.x]) : 0ull;
    __syncthreads();
    const int i = blockIdx.x * BS + threadIdx.x;
    int none = 0;
    if (i < N) {
        unsigned long long v[PAT_W];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = (unsigned long long)__double_as_longlong(AL[(size_t)k * lda + i]);
        v[6] = (unsigned long long)__double_as_longlong(D[i]);
        int id = PAT_NONE;
        for (int p = 0; p < npat && id == PAT_NONE; ++p) {
            bool eq = true;
#pragma unroll
            for (int k = 0; k < PAT_W; ++k) eq = eq && v[k] == t[p * PAT_W + k];
            if (eq) id = p;
        }
        pid[i] = (uint8_t)id;
        none = id == PAT_NONE;
    } else if (i < npad) {
        pid[i] = (uint8_t)PAT_NONE;
    }
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    if (none) atomicAdd(&cnt, 1);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(nnone, cnt);
}

// FP64 PM stage for an operator with row patterns: k_pm_v2's stage (NT, DOT
// as there) with two changes.  A wavefront takes two chunks of 128
// consecutive rows, two rows a lane in each, and issues both chunks' loads
// before either chunk's gathers -- twice the bytes in flight per wavefront;
// and a pattern row takes its coefficient from the table (column `col` of
// ptab, and col + 1 for NT = 2, staged in LDS), so the coefficient column
// is read only for pairs holding a row of their own (at 512x512x96 with
// bathymetry 95% of the rows are pattern rows and 9% of the column's cache
// lines hold another).  Measured on MI355X at 13.1M rows, natural row order,
// streamed, single stage: 50.7 us against 57.6 us for k_pm_v2; the two
// chunks without the table 60.2 us, the table with one chunk 53.8 us.
// A workgroup does the work of two k_pm_v2 workgroups -- row blocks 2g and
// 2g+1 of the block order, wavefront pair h = 0, 1 taking block 2g + h --
// and every row's arithmetic is k_pm_v2's.  Dot partials: chunk c of
// wavefront q holds exactly the rows of k_pm_v2's wavefront 2(q&1) + c of
// the same 512-row block, so each chunk's wave_sum equals that wavefront's
// and the four are added in k_pm_v2's order: the partials, and the solve,
// are bit-identical.  G: the row blocks (k_pm_v2's grid); the grid here is
// geo_half(G) workgroups, those past G/2 doing nothing.
template <int NT, int DOT, typename IX, bool ST>
__global__ void __launch_bounds__(BS)
k_pm_p2(const Scal *__restrict__ sc, int N, int G, const double *__restrict__ a0, IX l0,
        const double *__restrict__ a1, IX l1, int col, const double *__restrict__ in,
        double *__restrict__ out, const double *__restrict__ w, double *__restrict__ part,
        const int *__restrict__ order, const uint8_t *__restrict__ pid, const double *__restrict__ ptab) {
    __shared__ double t0[PAT_MAX], t1[PAT_MAX];
    const int done = sc->done;
    if (threadIdx.x < PAT_MAX) {
        t0[threadIdx.x] = ptab[threadIdx.x * PAT_W + col];
        if (NT == 2) t1[threadIdx.x] = ptab[threadIdx.x * PAT_W + col + 1];
    }
    __syncthreads();
    if (done) return;
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // wavefront-uniform: the code block's base is one scalar load
    const int blk = __builtin_amdgcn_readfirstlane(2 * logical_block() + (q >> 1));
    const bool live = blk < G;
    const int rb = live ? (order ? order[blk] : blk) : 0;
    const int i0 = rb * (2 * BS) + (q & 1) * (4 * 64) + 2 * lane;
    const auto bb = l0.blk_at(rb, N);  // one base array for all columns
    bool full[2];
    uint16_t pp[2];
    typename IX::Raw r0[2], r1[2];
    typename IX::Base b0[2];
    double2 xv[2], wv[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int i = i0 + 128 * c;
        full[c] = live && i + 1 < N;
        if (full[c]) {
            pp[c] = ldv<ST>(reinterpret_cast<const uint16_t *>(pid + i));
            b0[c] = bb;
            r0[c] = l0.template raw<ST>(i);
            if (NT == 2) r1[c] = l1.template raw<ST>(i);
            xv[c] = ld2(in + i);
            if (DOT == 1 || DOT == 2) wv[c] = lds2<ST>(w + i);
        }
    }
    // coefficients of the pairs with a row of their own
    double2 c0[2], c1[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        c0[c] = make_double2(0.0, 0.0);
        c1[c] = make_double2(0.0, 0.0);
        if (full[c] && ((pp[c] & 255) >= PAT_MAX || (pp[c] >> 8) >= PAT_MAX)) {
            c0[c] = lds2<ST>(a0 + i0 + 128 * c);
            if (NT == 2) c1[c] = lds2<ST>(a1 + i0 + 128 * c);
        }
    }
    double g0[2][2], g1[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (full[c]) {
            const int i = i0 + 128 * c;
            const int2 j0 = l0.dec1(r0[c], i, b0[c]);
            g0[c][0] = in[j0.x - 1];
            g0[c][1] = in[j0.y - 1];
            if (NT == 2) {
                const int2 j1 = l1.dec1(r1[c], i, b0[c]);  // one base array for all columns
                g1[c][0] = in[j1.x - 1];
                g1[c][1] = in[j1.y - 1];
            }
        }
    }
    double s0[2] = {0.0, 0.0}, s1[2] = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int i = i0 + 128 * c;
        if (full[c]) {
            const int p0 = pp[c] & 255, p1 = pp[c] >> 8;
            const double a00 = p0 < PAT_MAX ? t0[p0] : c0[c].x;
            const double a01 = p1 < PAT_MAX ? t0[p1] : c0[c].y;
            double o0 = xv[c].x - a00 * g0[c][0];
            double o1 = xv[c].y - a01 * g0[c][1];
            if (NT == 2) {
                const double a10 = p0 < PAT_MAX ? t1[p0] : c1[c].x;
                const double a11 = p1 < PAT_MAX ? t1[p1] : c1[c].y;
                o0 = o0 - a10 * g1[c][0];
                o1 = o1 - a11 * g1[c][1];
            }
            double2 o;
            o.x = o0;
            o.y = o1;
            sts2<ST>(out + i, o);
            if (DOT == 1 || DOT == 2) {
                s0[c] += wv[c].x * o0;
                s0[c] += wv[c].y * o1;
            }
            if (DOT == 2) {
                s1[c] += o0 * o0;
                s1[c] += o1 * o1;
            }
            if (DOT == 3) {
                s0[c] += o0 * o0;
                s0[c] += o1 * o1;
            }
        } else if (live && i < N) {
            double o = in[i] - a0[i] * in[l0.one(i)];
            if (NT == 2) o = o - a1[i] * in[l1.one(i)];
            out[i] = o;
            if (DOT == 1 || DOT == 2) s0[c] += w[i] * o;
            if (DOT == 2) s1[c] += o * o;
            if (DOT == 3) s0[c] += o * o;
        }
    }
    if (DOT != 0) {
        // [k_pm_v2 wavefront of the row block][dot], per half of the workgroup
        __shared__ double ws[2][4][2];
        const int h = q >> 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double v0 = wave_sum(s0[c]);
            const double v1 = DOT == 2 ? wave_sum(s1[c]) : 0.0;
            if (lane == 0) {
                ws[h][2 * (q & 1) + c][0] = v0;
                ws[h][2 * (q & 1) + c][1] = v1;
            }
        }
        __syncthreads();
        if (live && (threadIdx.x & 127) == 0) {
            double u0 = 0.0, u1 = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                u0 += ws[h][k][0];
                u1 += ws[h][k][1];
            }
            part[rb] = u0;
            if (DOT == 2) part[G + rb] = u1;
        }
    }
}

// Four consecutive PM coefficients as FP32: one 8-byte (half precision) or
// 16-byte load.
template <bool ST, typename C>
__device__ __forceinline__ f4v ldc4(const C *p) {
    typedef C c4 __attribute__((ext_vector_type(4)));
    const c4 v = ldv<ST>(reinterpret_cast<const c4 *>(p));
    f4v o;
    o.x = (float)v.x;
    o.y = (float)v.y;
    o.z = (float)v.z;
    o.w = (float)v.w;
    return o;
}

// FP32 PM stage with four consecutive rows per thread, the form of the mixed
// inner solve: the coefficients, the index codes and the input of the four
// rows are one load each (the FP32 stages' pace is set by the number of
// memory instructions more than by bytes: measured on MI355X at 13.1M rows,
// 33.6 us with two loads of each against 28.3 us with one).  NT, DOT as in
// k_pm_v2; the dot partials of block b go to part[b] (and part[G + b]).
template <int NT, int DOT, typename IX, bool ST, typename C>
__global__ void __launch_bounds__(BS)
k_pm_v4(const Scal *__restrict__ sc, int N, const C *__restrict__ a0, IX l0, const C *__restrict__ a1, IX l1,
        const float *__restrict__ in, float *__restrict__ out, const float *__restrict__ w,
        double *__restrict__ part) {
    if (sc->done) return;
    const int lb = logical_block();
    const int i = (lb * BS + threadIdx.x) * 4;
    // a wavefront's 256 rows lie in one code block: its base is one scalar load
    const auto bb = l0.blk_at(__builtin_amdgcn_readfirstlane(i / 512), N);
    double s0 = 0.0, s1 = 0.0;
    auto dots = [&](float o, float wi) {
        if (DOT == 1 || DOT == 2) s0 += (double)wi * (double)o;
        if (DOT == 2) s1 += (double)o * (double)o;
        if (DOT == 3) s0 += (double)o * (double)o;
    };
    const bool full = i + 3 < N;
    f4v c0, c1, x;
    int4 j0, j1;
    if (full) {
        c0 = ldc4<ST>(a0 + i);
        j0 = l0.template quad_b<ST>(i, bb);
        x = *reinterpret_cast<const f4v *>(in + i);
        if (NT == 2) {
            c1 = ldc4<ST>(a1 + i);
            j1 = l1.template quad_b<ST>(i, bb);
        }
    }
    if (full) {
        f4v o;
        o.x = x.x - c0.x * in[j0.x];
        o.y = x.y - c0.y * in[j0.y];
        o.z = x.z - c0.z * in[j0.z];
        o.w = x.w - c0.w * in[j0.w];
        if (NT == 2) {
            o.x = o.x - c1.x * in[j1.x];
            o.y = o.y - c1.y * in[j1.y];
            o.z = o.z - c1.z * in[j1.z];
            o.w = o.w - c1.w * in[j1.w];
        }
        stv<ST>(reinterpret_cast<f4v *>(out + i), o);
        if (DOT != 0) {
            f4v wv = {0.f, 0.f, 0.f, 0.f};
            if (DOT == 1 || DOT == 2) wv = ldv<ST>(reinterpret_cast<const f4v *>(w + i));
            dots(o.x, wv.x);
            dots(o.y, wv.y);
            dots(o.z, wv.z);
            dots(o.w, wv.w);
        }
    } else {
        for (int r = i; r < N && r < i + 4; ++r) {
            float o = in[r] - (float)a0[r] * in[l0.one(r)];
            if (NT == 2) o = o - (float)a1[r] * in[l1.one(r)];
            out[r] = o;
            dots(o, (DOT == 1 || DOT == 2) ? w[r] : 0.f);
        }
    }
    if (DOT != 0) {
        __shared__ double lds[BS / 64];
        const double t0 = block_sum<BS>(s0, lds);
        if (DOT == 2) {
            const double t1 = block_sum<BS>(s1, lds);
            if (threadIdx.x == 0) part[gridDim.x + lb] = t1;
        }
        if (threadIdx.x == 0) part[lb] = t0;
    }
}

template <typename T, typename IX, bool ST, typename C = T>
__global__ void __launch_bounds__(BS)
k_pm2_v2(const Scal *__restrict__ sc, int N, const C *__restrict__ a0, IX l0,
         const C *__restrict__ a1, IX l1, const T *__restrict__ in, T *__restrict__ out,
         const int *__restrict__ order) {
    const int done = sc->done;  // read together with the block order
    const int rb = row_block(order);
    const int i = (rb * BS + threadIdx.x) * 2;
    if (done) return;
    const auto bb = l0.blk_at(rb, N);  // the row block is the code block
    // stage 1 at row r, and at its stage-2 neighbour m (a ghost m reads 0)
    auto one = [&](T c0, int jj0, T c1, int m, T xr) -> T {
        const T o1i = xr - c0 * in[jj0];
        T o1m = T(0);
        if (m < N) o1m = in[m] - (T)a0[m] * in[l0.one(m)];
        return o1i - c1 * o1m;
    };
    if (i + 1 < N) {
        const auto c0 = ldc2<T, ST>(a0 + i);
        const int2 j0 = l0.template pair_b<ST>(i, bb);
        const auto c1 = ldc2<T, ST>(a1 + i);
        const int2 j1 = l1.template pair_b<ST>(i, bb);
        const auto xi = ld2(in + i);
        typename V2<T>::t o;
        o.x = one(c0.x, j0.x, c1.x, j1.x, xi.x);
        o.y = one(c0.y, j0.y, c1.y, j1.y, xi.y);
        sts2<ST>(out + i, o);
    } else if (i < N) {
        out[i] = one((T)a0[i], l0.one(i), (T)a1[i], l1.one(i), in[i]);
    }
}

// AXSUB (pcgpme_para.F90:323-352) with two consecutive rows per thread:
// y = A x in sorted slot order (MODE 0), or the true residual y = v1 - A x
// (MODE 3, the mixed solve's FP64 refinement).
// PAT (NL = 6): the coefficients and D of pattern rows come from the table
// ptab (pid the rows' pattern ids).
// A wave's dependent memory round trips set this kernel's pace more than its
// bytes (its row streams are a third of a PM stage's): the done flag, the
// block order and the table are read together, the row streams (codes,
// pattern ids, x) are in flight before the table is staged in LDS, and only
// the gathers wait for the codes.
template <typename T, int NL, int MODE, typename IX, bool ST, bool PAT = false>
__global__ void __launch_bounds__(BS)
k_spmv_v2(const Scal *__restrict__ sc, int N, int lda, const T *__restrict__ AL, IX LL,
          const T *__restrict__ D, const T *__restrict__ x, const T *__restrict__ v1, T *__restrict__ y,
          const int *__restrict__ order, const uint8_t *__restrict__ pid = nullptr,
          const double *__restrict__ ptab = nullptr) {
    static_assert(MODE == 0 || MODE == 3, "y = A x or y = v1 - A x");
    static_assert(!PAT || NL == 6, "patterns of six-slot rows");
    const int done = sc->done;
    const int rb = row_block(order);
    const int i = (rb * BS + threadIdx.x) * 2;
    __shared__ T tab[PAT ? PAT_MAX * PAT_W : 1];
    double tv = 0.0;  // converted when staged, so nothing waits for it here
    if (PAT && threadIdx.x < PAT_MAX * PAT_W) tv = ptab[threadIdx.x];
    if (done) return;
    const bool full = i + 1 < N;
    typename IX::Raw raw[NL];
    typename IX::Base bb;
    typename V2<T>::t xv;
    int p0 = PAT_NONE, p1 = PAT_NONE;
    if (full) {
        if constexpr (!IX::TABLE) bb = LL.blk_at(rb, N);  // the row block is the code block
#pragma unroll
        for (int k = 0; k < NL; ++k) raw[k] = LL.col((size_t)k * lda).template raw<ST>(i);
        if (PAT) {
            const uint16_t pp = ldv<ST>(reinterpret_cast<const uint16_t *>(pid + i));
            p0 = pp & 255;
            p1 = pp >> 8;
        }
        xv = ld2(x + i);
    }
    // Ix8: the block's decode table (its 2 * BS rows are one code block)
    __shared__ int DT[IX::TABLE ? 256 : 1];
    if constexpr (IX::TABLE) {
        // the row block is the code block (uniform: one scalar load of its base)
        static_assert(2 * BS == IX::ROWS, "a block is one code block");
        DT[threadIdx.x] = (long)rb * IX::ROWS < N ? IX::tab_entry(threadIdx.x, LL.base[rb]) : 0;
    }
    if (PAT || IX::TABLE) {
        if (PAT && threadIdx.x < PAT_MAX * PAT_W) tab[threadIdx.x] = (T)tv;
        __syncthreads();
    }
    if (full) {
        int2 l[NL];  // 1-based
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            if constexpr (IX::TABLE) {
                const IX c = LL.col((size_t)k * lda);
                const int c0 = raw[k] & 255, c1 = raw[k] >> 8;
                l[k] = make_int2(c.near_t(c0, i, DT) + 1, c.near_t(c1, i + 1, DT) + 1);
                if (c0 == IX::WIDE) l[k].x = c.L32[i];
                if (c1 == IX::WIDE) l[k].y = c.L32[i + 1];
            } else {
                l[k] = LL.col((size_t)k * lda).dec1(raw[k], i, bb);
            }
        }
        typename V2<T>::t a[NL];
        typename V2<T>::t d;
        if (PAT && p0 < PAT_MAX && p1 < PAT_MAX) {
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                a[k].x = tab[p0 * PAT_W + k];
                a[k].y = tab[p1 * PAT_W + k];
            }
            d.x = tab[p0 * PAT_W + 6];
            d.y = tab[p1 * PAT_W + 6];
        } else {
#pragma unroll
            for (int k = 0; k < NL; ++k) a[k] = lds2<ST>(AL + (size_t)k * lda + i);
            d = lds2<ST>(D + i);
            if (PAT) {
                if (p0 < PAT_MAX) {
#pragma unroll
                    for (int k = 0; k < NL; ++k) a[k].x = tab[p0 * PAT_W + k];
                    d.x = tab[p0 * PAT_W + 6];
                }
                if (p1 < PAT_MAX) {
#pragma unroll
                    for (int k = 0; k < NL; ++k) a[k].y = tab[p1 * PAT_W + k];
                    d.y = tab[p1 * PAT_W + 6];
                }
            }
        }
        T acc0 = xv.x * d.x, acc1 = xv.y * d.y;
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            acc0 = acc0 + a[k].x * x[l[k].x - 1];
            acc1 = acc1 + a[k].y * x[l[k].y - 1];
        }
        typename V2<T>::t o;
        if (MODE == 3) {
            const auto bv = ld2(v1 + i);
            o.x = bv.x - acc0;
            o.y = bv.y - acc1;
        } else {
            o.x = acc0;
            o.y = acc1;
        }
        sts2<ST>(y + i, o);
    } else if (i < N) {
        T acc = x[i] * D[i];
#pragma unroll
        for (int k = 0; k < NL; ++k) acc = acc + AL[(size_t)k * lda + i] * x[LL.col((size_t)k * lda).one(i)];
        y[i] = MODE == 3 ? v1[i] - acc : acc;
    }
}

// Row-major copies for the FP32 SpMV of the mixed inner solve: the six
// coefficients of a row together (ALr[6 i + k]) and its six 16-bit codes
// (Lr16[6 i + k]) and, when the operator has them, 8-bit codes (Lr8).  Two
// rows are then three 16-byte loads of coefficients and three 8- (or 4-)
// byte loads of codes instead of twelve, the FP32 kernels' pace being set
// more by the number of memory instructions than by bytes.  *nonunit counts
// rows whose FP32 diagonal is not 1: after PCGPME's scaling every row with
// max(sum|AL|, D) = D has D W = P2 D / ((1 + P1) D) = 1 up to one rounding,
// which FP32 drops, so the SpMV can skip the diagonal stream when it is 0.
__global__ void __launch_bounds__(BS)
k_to_rows32(int N, int lda, const double *__restrict__ AL, const short *__restrict__ L16,
            const uint8_t *__restrict__ L8, const float *__restrict__ D32, float *__restrict__ ALr,
            short *__restrict__ Lr16, uint8_t *__restrict__ Lr8, int *__restrict__ nonunit) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        ALr[(size_t)6 * i + k] = (float)AL[(size_t)k * lda + i];
        Lr16[(size_t)6 * i + k] = L16[(size_t)k * lda + i];
        if (L8) Lr8[(size_t)6 * i + k] = L8[(size_t)k * lda + i];
    }
    if (D32[i] != 1.0f) atomicAdd(nonunit, 1);
}

// The codes of two consecutive rows (i even) from the row-major copies.
template <bool ST>
__device__ __forceinline__ void row_codes(const short *__restrict__ Lr, int i, int (&cr)[2][6]) {
    typedef short s4 __attribute__((ext_vector_type(4)));
    const s4 *cp = reinterpret_cast<const s4 *>(Lr + (size_t)6 * i);
    const s4 c0 = ldv<ST>(cp), c1 = ldv<ST>(cp + 1), c2 = ldv<ST>(cp + 2);
    const int v[12] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y, c2.z, c2.w};
#pragma unroll
    for (int k = 0; k < 12; ++k) cr[k / 6][k % 6] = v[k];
}
template <bool ST>
__device__ __forceinline__ void row_codes(const uint8_t *__restrict__ Lr, int i, int (&cr)[2][6]) {
    const unsigned *cp = reinterpret_cast<const unsigned *>(Lr + (size_t)6 * i);
    const unsigned w[3] = {ldv<ST>(cp), ldv<ST>(cp + 1), ldv<ST>(cp + 2)};
#pragma unroll
    for (int k = 0; k < 12; ++k) cr[k / 6][k % 6] = (int)((w[k / 4] >> (8 * (k % 4))) & 255u);
}

// FP32 y = A x (NL = 6) from the row-major copies, two rows per thread; IX
// (Ix16 with Lr16, or Ix8 with Lr8) decodes the codes, wide slots reading
// the column-major LL.  With *nonunit == 0 every diagonal is 1 and is not
// read (x * 1 = x exactly).  Same per-row arithmetic as k_spmv_v2.
// PAT: the coefficients of pattern rows come from the table ptab (FP32
// roundings of its values, as ALr holds them).
template <typename IX, typename CT, bool ST, bool PAT = false>
__global__ void __launch_bounds__(BS)
k_spmv_rows32(const Scal *__restrict__ sc, int N, int lda, const float *__restrict__ ALr,
              const CT *__restrict__ Lr, IX LL, const float *__restrict__ D, const int *__restrict__ nonunit,
              const float *__restrict__ x, float *__restrict__ y, const uint8_t *__restrict__ pid = nullptr,
              const double *__restrict__ ptab = nullptr) {
    const int done = sc->done;
    const bool unit = *nonunit == 0;
    __shared__ float tab[PAT ? PAT_MAX * PAT_W : 1];
    double tv = 0.0;  // converted when staged, so nothing waits for it here
    if (PAT && threadIdx.x < PAT_MAX * PAT_W) tv = ptab[threadIdx.x];
    if (done) return;
    const int lb = logical_block();  // the row block is the code block (2 BS rows)
    const int i = (lb * BS + threadIdx.x) * 2;
    const bool full = i + 1 < N;
    int cr[2][6];
    int p[2] = {PAT_NONE, PAT_NONE};
    float2 xv;
    std::remove_cv_t<std::remove_reference_t<decltype(*LL.base)>> b;
    if (full) {
        if (PAT) {
            const uint16_t pp = ldv<ST>(reinterpret_cast<const uint16_t *>(pid + i));
            p[0] = pp & 255;
            p[1] = pp >> 8;
        }
        row_codes<ST>(Lr, i, cr);
        if constexpr (!IX::TABLE) b = LL.blk_at(lb, N);
        xv = ld2(x + i);
    }
    // Ix8: the block's decode table (its 2 * BS rows are one code block)
    __shared__ int T[IX::TABLE ? 256 : 1];
    if constexpr (IX::TABLE) {
        static_assert(2 * BS == IX::ROWS, "a block is one code block");
        T[threadIdx.x] = (long)lb * IX::ROWS < N ? IX::tab_entry(threadIdx.x, LL.base[lb]) : 0;
    }
    if (PAT || IX::TABLE) {
        if (PAT && threadIdx.x < PAT_MAX * PAT_W) tab[threadIdx.x] = (float)tv;
        __syncthreads();
    }
    if (full) {
        float ar[2][6];
        if (PAT && p[0] < PAT_MAX && p[1] < PAT_MAX) {
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int k = 0; k < 6; ++k) ar[r][k] = tab[p[r] * PAT_W + k];
        } else {
            const f4v *ap = reinterpret_cast<const f4v *>(ALr + (size_t)6 * i);
            const f4v a0 = ldv<ST>(ap), a1 = ldv<ST>(ap + 1), a2 = ldv<ST>(ap + 2);
            const float av[12] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int k = 0; k < 6; ++k) ar[r][k] = (PAT && p[r] < PAT_MAX) ? tab[p[r] * PAT_W + k] : av[6 * r + k];
        }
        float acc0 = xv.x, acc1 = xv.y;
        if (!unit) {
            const float2 d = lds2<ST>(D + i);
            acc0 = xv.x * d.x;
            acc1 = xv.y * d.y;
        }
        int j[2][6];  // 1-based, as k_spmv_v2's
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const IX c = LL.col((size_t)k * lda);
            if constexpr (IX::TABLE) {
                j[0][k] = c.near_t(cr[0][k], i, T) + 1;
                j[1][k] = c.near_t(cr[1][k], i + 1, T) + 1;
            } else {
                j[0][k] = c.near1(cr[0][k], i, b);
                j[1][k] = c.near1(cr[1][k], i + 1, b);
            }
            if (cr[0][k] == IX::WIDE) j[0][k] = c.L32[i];
            if (cr[1][k] == IX::WIDE) j[1][k] = c.L32[i + 1];
        }
        // the two rows' neighbours in a slot are usually adjacent (same
        // offset): one 8-byte load takes both, a second load only where not
        float g[2][6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const f2u v = *reinterpret_cast<const f2u *>(x + (j[0][k] - 1));
            g[0][k] = v.x;
            g[1][k] = v.y;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (j[1][k] != j[0][k] + 1) g[1][k] = x[j[1][k] - 1];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            acc0 = acc0 + ar[0][k] * g[0][k];
            acc1 = acc1 + ar[1][k] * g[1][k];
        }
        float2 o;
        o.x = acc0;
        o.y = acc1;
        sts2<ST>(y + i, o);
    } else if (i < N) {
        float acc = unit ? x[i] : x[i] * D[i];
#pragma unroll
        for (int k = 0; k < 6; ++k) acc = acc + ALr[(size_t)6 * i + k] * x[LL.col((size_t)k * lda).one(i)];
        y[i] = acc;
    }
}

// k_spmv_rows32 in the two-chunk form: a wavefront takes 256 rows of one
// 512-row code block as two chunks of 128 (two rows a lane), and issues both
// chunks' pattern ids, codes and x before the first chunk's gathers; a
// workgroup takes two code blocks (grid: geo_half of k_spmv_rows32's).  IX:
// Ix8 with Lr8 (the blocks' decode tables in LDS) or Ix16 with Lr16 (the
// blocks' bases).  Same per-row arithmetic, so y is bit-identical.
template <typename IX, typename CT, bool ST, bool PAT>
__global__ void __launch_bounds__(BS)
k_spmv_rows32_p2(const Scal *__restrict__ sc, int N, int lda, const float *__restrict__ ALr,
                 const CT *__restrict__ Lr, IX LL, const float *__restrict__ D, const int *__restrict__ nonunit,
                 const float *__restrict__ x, float *__restrict__ y, const uint8_t *__restrict__ pid,
                 const double *__restrict__ ptab) {
    static_assert(2 * BS == IX::ROWS, "half a workgroup is one code block");
    const int done = sc->done;
    const bool unit = *nonunit == 0;
    __shared__ float tab[PAT ? PAT_MAX * PAT_W : 1];
    __shared__ int T[IX::TABLE ? 2 : 1][IX::TABLE ? 256 : 1];
    double tv = 0.0;
    if (PAT && threadIdx.x < PAT_MAX * PAT_W) tv = ptab[threadIdx.x];
    if (done) return;
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63, h = q >> 1;
    const int g = logical_block();
    const int i0 = (2 * g + h) * IX::ROWS + (q & 1) * 256 + 2 * lane;
    // Ix16: the wavefront's code block base (one of the workgroup's two)
    std::remove_cv_t<std::remove_reference_t<decltype(*LL.base)>> bw;
    if constexpr (!IX::TABLE) {
        const auto b0 = LL.blk_at(2 * g, N), b1 = LL.blk_at(2 * g + 1, N);
        bw = h ? b1 : b0;
    }
    bool full[2];
    int cr[2][2][6];
    int pp[2] = {0, 0};
    float2 xv[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int i = i0 + 128 * c;
        full[c] = i + 1 < N;
        if (full[c]) {
            if (PAT) pp[c] = ldv<ST>(reinterpret_cast<const uint16_t *>(pid + i));
            row_codes<ST>(Lr, i, cr[c]);
            xv[c] = ld2(x + i);
        }
    }
    if constexpr (IX::TABLE) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const long r0 = (long)(2 * g + b) * IX::ROWS;
            T[b][threadIdx.x] = r0 < N ? IX::tab_entry(threadIdx.x, LL.base[2 * g + b]) : 0;
        }
    }
    if (PAT && threadIdx.x < PAT_MAX * PAT_W) tab[threadIdx.x] = (float)tv;
    if (PAT || IX::TABLE) __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int i = i0 + 128 * c;
        if (full[c]) {
            const int p[2] = {PAT ? (pp[c] & 255) : PAT_NONE, PAT ? (pp[c] >> 8) : PAT_NONE};
            float ar[2][6];
            if (PAT && p[0] < PAT_MAX && p[1] < PAT_MAX) {
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int k = 0; k < 6; ++k) ar[r][k] = tab[p[r] * PAT_W + k];
            } else {
                const f4v *ap = reinterpret_cast<const f4v *>(ALr + (size_t)6 * i);
                const f4v a0 = ldv<ST>(ap), a1 = ldv<ST>(ap + 1), a2 = ldv<ST>(ap + 2);
                const float av[12] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        ar[r][k] = (PAT && p[r] < PAT_MAX) ? tab[p[r] * PAT_W + k] : av[6 * r + k];
            }
            float acc0 = xv[c].x, acc1 = xv[c].y;
            if (!unit) {
                const float2 d = lds2<ST>(D + i);
                acc0 = xv[c].x * d.x;
                acc1 = xv[c].y * d.y;
            }
            int j[2][6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const IX cc = LL.col((size_t)k * lda);
                if constexpr (IX::TABLE) {
                    j[0][k] = cc.near_t(cr[c][0][k], i, T[h]) + 1;
                    j[1][k] = cc.near_t(cr[c][1][k], i + 1, T[h]) + 1;
                } else {
                    j[0][k] = cc.near1(cr[c][0][k], i, bw);
                    j[1][k] = cc.near1(cr[c][1][k], i + 1, bw);
                }
                if (cr[c][0][k] == IX::WIDE) j[0][k] = cc.L32[i];
                if (cr[c][1][k] == IX::WIDE) j[1][k] = cc.L32[i + 1];
            }
            float gv[2][6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const f2u v = *reinterpret_cast<const f2u *>(x + (j[0][k] - 1));
                gv[0][k] = v.x;
                gv[1][k] = v.y;
            }
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (j[1][k] != j[0][k] + 1) gv[1][k] = x[j[1][k] - 1];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                acc0 = acc0 + ar[0][k] * gv[0][k];
                acc1 = acc1 + ar[1][k] * gv[1][k];
            }
            float2 o;
            o.x = acc0;
            o.y = acc1;
            sts2<ST>(y + i, o);
        } else if (i < N) {
            float acc = unit ? x[i] : x[i] * D[i];
#pragma unroll
            for (int k = 0; k < 6; ++k) acc = acc + ALr[(size_t)6 * i + k] * x[LL.col((size_t)k * lda).one(i)];
            y[i] = acc;
        }
    }
}

// The vector kernels take 16 bytes of each vector per access: two rows a
// thread in FP64, four in FP32.
template <typename T> struct VecOf;
template <> struct VecOf<float> {
    static constexpr int R = 4;
    typedef float t __attribute__((ext_vector_type(4)));
};
template <> struct VecOf<double> {
    static constexpr int R = 2;
    typedef double t __attribute__((ext_vector_type(2)));
};

// E = R - ALP*Q (pcgpme_para.F90:173) as its own pass.
template <typename T>
__global__ void __launch_bounds__(BS)
k_eupd_vec(const Scal *__restrict__ sc, int N, const T *__restrict__ r, const T *__restrict__ q,
           T *__restrict__ e) {
    if (sc->done) return;
    using VT = typename VecOf<T>::t;
    constexpr int RV = VecOf<T>::R;
    const int i = (logical_block() * BS + threadIdx.x) * RV;
    const T alp = (T)sc->ALP;
    if (i + RV - 1 < N) {
        const VT rv = *reinterpret_cast<const VT *>(r + i), qv = *reinterpret_cast<const VT *>(q + i);
        VT o;
#pragma unroll
        for (int k = 0; k < RV; ++k) o[k] = rv[k] - alp * qv[k];
        *reinterpret_cast<VT *>(e + i) = o;
    } else {
        for (int k = i; k < N && k < i + RV; ++k) e[k] = r[k] - alp * q[k];
    }
}

// X = X + ALP*P + AMU*E;  R = E - AMU*V;  RN += R*R;  C1 += R*R^
// (pcgpme_para.F90:203-211); per-block partials of RN and C1.
template <typename T>
__global__ void __launch_bounds__(BS)
k_update_vec(const Scal *__restrict__ sc, int N, T *__restrict__ x, const T *__restrict__ p,
             const T *__restrict__ e, const T *__restrict__ v, T *__restrict__ r, const T *__restrict__ rh,
             double *__restrict__ part) {
    if (sc->done) return;
    using VT = typename VecOf<T>::t;
    constexpr int RV = VecOf<T>::R;
    const T alp = (T)sc->ALP, amu = (T)sc->AMU;
    double rn = 0.0, c1 = 0.0;
    const int lb = logical_block();
    const int i = (lb * BS + threadIdx.x) * RV;
    auto row = [&](T xi, T pi, T ei, T vi, T rhi, T &xo, T &ro) {
        T t = ei;
        xo = xi + alp * pi + amu * t;
        t = t - amu * vi;
        ro = t;
        rn += (double)t * (double)t;
        c1 += (double)t * (double)rhi;
    };
    if (i + RV - 1 < N) {
        const VT xv = *reinterpret_cast<const VT *>(x + i), pv = *reinterpret_cast<const VT *>(p + i),
                 ev = *reinterpret_cast<const VT *>(e + i), vv = *reinterpret_cast<const VT *>(v + i),
                 hv = *reinterpret_cast<const VT *>(rh + i);
        VT xo, ro;
#pragma unroll
        for (int k = 0; k < RV; ++k) {
            T a, b;
            row(xv[k], pv[k], ev[k], vv[k], hv[k], a, b);
            xo[k] = a;
            ro[k] = b;
        }
        *reinterpret_cast<VT *>(x + i) = xo;
        *reinterpret_cast<VT *>(r + i) = ro;
    } else {
        for (int k = i; k < N && k < i + RV; ++k) {
            T xo, ro;
            row(x[k], p[k], e[k], v[k], rh[k], xo, ro);
            x[k] = xo;
            r[k] = ro;
        }
    }
    __shared__ double lds[BS / 64];
    const double t0 = block_sum<BS>(rn, lds);
    const double t1 = block_sum<BS>(c1, lds);
    if (threadIdx.x == 0) {
        part[lb] = t0;
        part[gridDim.x + lb] = t1;
    }
}

// P = R + BETA*(P - AMU*Q) (pcgpme_para.F90:244), in place.
template <typename T>
__global__ void __launch_bounds__(BS)
k_pupd_vec(const Scal *__restrict__ sc, int N, T *__restrict__ p, const T *__restrict__ r,
           const T *__restrict__ q) {
    if (sc->done) return;
    using VT = typename VecOf<T>::t;
    constexpr int RV = VecOf<T>::R;
    const int i = (logical_block() * BS + threadIdx.x) * RV;
    const T beta = (T)sc->BETA, amu = (T)sc->AMU;
    if (i + RV - 1 < N) {
        const VT pv = *reinterpret_cast<const VT *>(p + i), rv = *reinterpret_cast<const VT *>(r + i),
                 qv = *reinterpret_cast<const VT *>(q + i);
        VT o;
#pragma unroll
        for (int k = 0; k < RV; ++k) o[k] = rv[k] + beta * (pv[k] - amu * qv[k]);
        *reinterpret_cast<VT *>(p + i) = o;
    } else {
        for (int k = i; k < N && k < i + RV; ++k) p[k] = r[k] + beta * (p[k] - amu * q[k]);
    }
}

// Initial residual (a5, pcgpme_para.F90:133-142): wk1 holds PM*W*b, wk3 holds
// PM*A*x0 (nullptr when x0 = 0, where it is exactly zero).
template <typename T>
__global__ void __launch_bounds__(BS)
k_init(const Scal *__restrict__ sc, int N, T *__restrict__ wk1, const T *__restrict__ wk3, T *__restrict__ r,
       T *__restrict__ rh, double *__restrict__ part) {
    if (sc->done) return;
    double bn = 0.0, c1 = 0.0;
    const int lb = logical_block();
    const int i = lb * BS + threadIdx.x;
    if (i < N) {
        T t = wk1[i];
        bn += (double)t * (double)t;
        if (wk3) t = t - wk3[i];
        r[i] = t;
        c1 += (double)t * (double)t;
        wk1[i] = t;
        rh[i] = t;
    }
    __shared__ double lds[BS / 64];
    const double t0 = block_sum<BS>(bn, lds);
    const double t1 = block_sum<BS>(c1, lds);
    if (threadIdx.x == 0) {
        part[lb] = t0;
        part[gridDim.x + lb] = t1;
    }
}

// ------------------------------------------------------------ finalize --
// Each thread sums its partials k, k + FIN_BS, k + 2 FIN_BS, ... in that
// order, then the block tree adds the threads in fixed order.  The partials
// are few (one per 512 rows) and latency, not bandwidth, bounds this: every
// thread issues its loads (FIN_U per array at a time, predicated) before the
// first add, and the two-array form loads both arrays at once.
constexpr int FIN_U = 16;  // loads per array and thread issued together
template <int NA>
__device__ __forceinline__ void fin_sums(const double *__restrict__ p, int G, double *out) {
    __shared__ double lds[FIN_BS / 64];
    double s[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) s[a] = 0.0;
    for (int k0 = 0; k0 < G; k0 += FIN_U * FIN_BS) {
        double v[NA][FIN_U];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u) {
                const int k = k0 + (int)threadIdx.x + u * FIN_BS;
                v[a][u] = k < G ? p[(size_t)a * G + k] : 0.0;
            }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int u = 0; u < FIN_U; ++u)
                if (k0 + (int)threadIdx.x + u * FIN_BS < G) s[a] += v[a][u];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) out[a] = block_sum<FIN_BS>(s[a], lds);
}
__device__ __forceinline__ double fin_sum(const double *__restrict__ p, int G) {
    double r;
    fin_sums<1>(p, G, &r);
    return r;
}

__global__ void k_reset(Scal *sc, double eps, int itermax, int beta_prod) {
    if (threadIdx.x != 0) return;
    Scal s = {};
    s.EPS = eps;
    s.itermax = itermax;
    s.beta_prod = beta_prod;
    *sc = s;
}

// After k_init: BN, C1.  A zero right-hand side with a zero residual returns
// at once with X unchanged (the reference divides 0/0 there).
__global__ void __launch_bounds__(FIN_BS)
k_fin_init(Scal *sc, const double *__restrict__ part, int G) {
    if (sc->done) return;
    double r[2];
    fin_sums<2>(part, G, r);
    const double bn = r[0], c1 = r[1];
    if (threadIdx.x == 0) {
        sc->BN = bn;
        sc->C1 = c1;
        sc->ERRP = 0.0;
        sc->K = 0;
        if (bn == 0.0 && c1 == 0.0) {
            sc->ERR = 0.0;
            sc->kout = 0;
            sc->ier = 0;
            sc->done = 1;
        }
    }
}

// ALP = C1/RQ (pcgpme_para.F90:167).  The scalars are read before the sums
// (their loads overlap the partials').
__global__ void __launch_bounds__(FIN_BS)
k_fin_alpha(Scal *sc, const double *__restrict__ part, int G) {
    if (sc->done) return;
    const double C1 = sc->C1;
    const double rq = fin_sum(part, G);
    if (threadIdx.x == 0) {
        sc->RQ = rq;
        sc->ALP = C1 / rq;
    }
}

// AMU = (E,V)/(V,V) (pcgpme_para.F90:197).
__global__ void __launch_bounds__(FIN_BS)
k_fin_omega(Scal *sc, const double *__restrict__ part, int G) {
    if (sc->done) return;
    double r[2];
    fin_sums<2>(part, G, r);
    if (threadIdx.x == 0) {
        sc->SS = r[1];
        sc->AMU = r[0] / r[1];
    }
}

// ERR and the exits in the reference's order (pcgpme_para.F90:215-251).
__global__ void __launch_bounds__(FIN_BS)
k_fin_beta(Scal *sc, const double *__restrict__ part, int G, int K) {
    if (sc->done) return;
    const double BN = sc->BN, ERRP = sc->ERRP, EPS = sc->EPS, AMU = sc->AMU, RQ = sc->RQ;
    const int beta_prod = sc->beta_prod, itermax = sc->itermax;
    double r[2];
    fin_sums<2>(part, G, r);
    const double rn = r[0], c1 = r[1];
    if (threadIdx.x != 0) return;
    sc->RN = rn;
    sc->C1 = c1;
    sc->K = K;
    const double err = sqrt(rn / BN);
    sc->ERR = err;
    int stop = 0;
    if (!(err == err) || err > 1.7976931348623157e308) {
        sc->ier = 4001; // CLOWNS_ENONFINITE
        stop = 1;
    } else if (K >= 10 && fabs(err - ERRP) <= err * 1.0e-6) {
        stop = 1;
    } else if (err <= EPS) {
        stop = 1;
    } else {
        sc->ERRP = err;
        if (AMU == 0.0 || c1 == 0.0) {
            stop = 1;
        } else {
            // Same quantity, two roundings: pcgpme_para.F90:241 divides
            // twice, pcgpme_2DH.f90:241 divides by the product.
            sc->BETA = beta_prod ? c1 / (AMU * RQ) : c1 / AMU / RQ;
            if (K >= itermax) {
                sc->ier = 3000;
                sc->kout = K + 1;
                sc->done = 1;
                return;
            }
        }
    }
    if (stop) {
        sc->kout = K;
        sc->done = 1;
    }
}

// Mixed precision, outer step: ERR64 = sqrt(RN/BN) of the true preconditioned
// residual; RN0 = ||r|| scales the FP32 inner right-hand side.
__global__ void __launch_bounds__(FIN_BS)
k_fin_true(Scal *outer, const double *__restrict__ part, int G) {
    const double rn = fin_sum(part, G);
    if (threadIdx.x == 0) {
        outer->RN = rn;
        outer->RN0 = sqrt(rn);
        outer->ERR = sqrt(rn / outer->BN);
    }
}

// Mixed: FP32 inner right-hand side = r/||r||.
__global__ void __launch_bounds__(BS)
k_rhs32(int N, const Scal *__restrict__ outer, const double *__restrict__ r,
        float *__restrict__ r32) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    const double s = outer->RN0 > 0.0 ? 1.0 / outer->RN0 : 0.0;
    r32[i] = (float)(r[i] * s);
}

// Mixed: x += ||r|| * delta (FP64 accumulate of the FP32 correction).
__global__ void __launch_bounds__(BS)
k_correct(int N, const Scal *__restrict__ outer, const float *__restrict__ d,
          double *__restrict__ x) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    x[i] = x[i] + outer->RN0 * (double)d[i];
}

// ------------------------------------------------------------- adjoint --
// A^T in the reference layout, from the unscaled caller-layout operator.
// The 7-point pattern is structurally symmetric (a face is open for both of
// its cells, src2LOWNS/1sub-matrix.F90:41-45, 1sub-momentum.F90:496-512), so
// row j of A^T keeps row j's slots: A^T(j, LL(j,k)) = A(LL(j,k), j), found by
// searching row LL(j,k)'s slots for j.  A missing partner sets *bad.
__global__ void __launch_bounds__(BS)
k_transpose(int N, int lda, int NL, const double *__restrict__ AL,
            const int *__restrict__ LL, const double *__restrict__ D,
            double *__restrict__ ALt, int *__restrict__ LLt,
            double *__restrict__ Dt, int *__restrict__ bad) {
    const int j = blockIdx.x * BS + threadIdx.x;
    if (j >= N) return;
    for (int k = 0; k < NL; ++k) {
        const int i = LL[(size_t)k * lda + j];
        double val = 0.0;
        int idx = 0;
        if (i >= 1 && i <= N) {
            int found = 0;
            for (int kk = 0; kk < NL; ++kk) {
                if (LL[(size_t)kk * lda + (i - 1)] == j + 1) {
                    val = AL[(size_t)kk * lda + (i - 1)];
                    found = 1;
                    break;
                }
            }
            if (!found) atomicOr(bad, 1);
            idx = i;
        } else if (i != 0) {
            atomicOr(bad, 2);
        }
        ALt[(size_t)k * lda + j] = val;
        LLt[(size_t)k * lda + j] = idx;
    }
    Dt[j] = D[j];
}

// dL/dAL(j,k) = -lambda_j * x(LL(j,k)),  dL/dD_j = -lambda_j * x_j, in the
// caller's unscaled layout (LL = 0 slots get 0).
__global__ void __launch_bounds__(BS)
k_grad_coef(int N, int lda, int lda_out, int NL, const int *__restrict__ LL,
            const double *__restrict__ lam, const double *__restrict__ x,
            double *__restrict__ gAL, double *__restrict__ gD) {
    const int j = blockIdx.x * BS + threadIdx.x;
    if (j >= N) return;
    const double l = lam[j];
    for (int k = 0; k < NL; ++k) {
        const int i = LL[(size_t)k * lda + j];
        if (gAL) gAL[(size_t)k * lda_out + j] = (i >= 1 && i <= N) ? -l * x[i - 1] : 0.0;
    }
    if (gD) gD[j] = -l * x[j];
}

// -------------------------------------------------- PCGPME_WILL output --
// The output pass of PCGPME_WILL (src2LOWNS/pcgpme_2DH.f90:255-268), serial
// in the reference: row i zeroes |X(i)| <= 1e-14, then walks its caller-order
// neighbours LL(i,1..NL) (skipping itself, stopping at the first 0) and
// snaps X(i) to a neighbour within 1e-14; B(i) receives X(i) at every
// neighbour that differs from it and otherwise keeps the final residual.
// Row i reads the already filtered X of neighbours j < i and the raw X of
// neighbours j > i.  On the device the filtered vector is the fixed point of
// Y <- F(Y), F(Y)_i = row i's pass with neighbours j < i read from Y: the
// dependence is on lower rows only, so the fixed point is unique and equals
// the serial result; the rounds below reach it after at most (longest chain
// of snaps) + 1 rounds, each touching only the rows whose inputs changed.
constexpr double WILL_LIMIT = 1e-14;

__device__ __forceinline__ double will_row(int i, int lda, int NL, const int *__restrict__ LL,
                                           const double *__restrict__ X, const double *__restrict__ Y,
                                           bool snap_y, bool &assigned) {
    double xi = X[i];
    if (fabs(xi) <= WILL_LIMIT) xi = 0.0;
    assigned = false;
    for (int k = 0; k < NL; ++k) {
        const int j = LL[(size_t)k * lda + i];
        if (j == i + 1) continue;
        if (j == 0) break;
        double xj;
        if (j - 1 < i) {
            xj = Y[j - 1];
            if (snap_y && fabs(xj) <= WILL_LIMIT) xj = 0.0;
        } else {
            xj = X[j - 1];
        }
        if (xi == xj) continue;
        if (fabs(xi - xj) <= WILL_LIMIT) xi = xj;
        assigned = true;
    }
    return xi;
}

// Start of the rounds: Y = X with the zero limit applied (the value of a row
// that snaps to no neighbour).
__global__ void __launch_bounds__(BS)
k_will_init(int N, const double *__restrict__ X, double *__restrict__ Y) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    const double x = X[i];
    Y[i] = fabs(x) <= WILL_LIMIT ? 0.0 : x;
}

// Readers of each row as CSR (off[0..N], readers[off[j] .. off[j+1])): the
// rows m > j whose pass reads row j (j+1 in a slot before m's first zero).
__global__ void __launch_bounds__(BS)
k_will_rdeg(int N, int lda, int NL, const int *__restrict__ LL, int *__restrict__ deg) {
    const int m = blockIdx.x * BS + threadIdx.x;
    if (m >= N) return;
    for (int k = 0; k < NL; ++k) {
        const int j = LL[(size_t)k * lda + m];
        if (j == m + 1) continue;
        if (j == 0) break;
        if (j - 1 < m) atomicAdd(&deg[j - 1], 1);
    }
}

// Per-block sums of deg over blocks of WILL_SCAN rows (then scanned by
// k_cls_scan_blocks).
constexpr int WILL_SCAN = 2048;
__global__ void __launch_bounds__(BS)
k_will_rsum(int N, const int *__restrict__ deg, int *__restrict__ bsum) {
    const int base = blockIdx.x * WILL_SCAN;
    int n = 0;
    for (int u = threadIdx.x; u < WILL_SCAN; u += BS)
        if (base + u < N) n += deg[base + u];
    __shared__ int lds[BS];
    lds[threadIdx.x] = n;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = lds[0];
}

// off = exclusive scan of deg: block start from the scanned block sums, then
// a scan of each 256-row chunk in LDS.  off[N] = the total.
__global__ void __launch_bounds__(BS)
k_will_roff(int N, const int *__restrict__ deg, const int *__restrict__ bstart, int *__restrict__ off) {
    const int base = blockIdx.x * WILL_SCAN;
    __shared__ int lds[BS];
    int run = bstart[blockIdx.x];
    for (int u = 0; u < WILL_SCAN; u += BS) {
        const int i = base + u + threadIdx.x;
        const int d = i < N ? deg[i] : 0;
        lds[threadIdx.x] = d;
        __syncthreads();
        for (int o = 1; o < BS; o <<= 1) {
            const int v = threadIdx.x >= o ? lds[threadIdx.x - o] : 0;
            __syncthreads();
            lds[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < N) off[i] = run + lds[threadIdx.x] - d;
        run += lds[BS - 1];
        __syncthreads();
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) off[N] = bstart[gridDim.x];
}

__global__ void __launch_bounds__(BS)
k_will_rfill(int N, int lda, int NL, const int *__restrict__ LL, const int *__restrict__ off,
             int *__restrict__ fillc, int *__restrict__ readers) {
    const int m = blockIdx.x * BS + threadIdx.x;
    if (m >= N) return;
    for (int k = 0; k < NL; ++k) {
        const int j = LL[(size_t)k * lda + m];
        if (j == m + 1) continue;
        if (j == 0) break;
        if (j - 1 < m) readers[off[j - 1] + atomicAdd(&fillc[j - 1], 1)] = m;
    }
}

// One round, in place over a candidate list: round 0 evaluates every row,
// round r > 0 the rows pushed in round r-1.  A row whose value changes
// stores it and pushes its readers (deduplicated by `stamp`) onto the next
// list.  Rows read the current Y of their lower neighbours, this round's or
// the last one's: every row is evaluated again after the last change of each
// row it reads, so when a round pushes nothing, Y = F(Y): the unique fixed
// point.  cnt[r] = length of round r's list.
__global__ void __launch_bounds__(BS)
k_will_front(int N, int lda, int NL, const int *__restrict__ LL, const double *__restrict__ X,
             double *Y, const int *__restrict__ off, const int *__restrict__ readers,
             const int *__restrict__ cand_in, int *__restrict__ cand_out, int *__restrict__ stamp,
             int *__restrict__ cnt, int r) {
    const int n = r == 0 ? N : cnt[r];
    for (int t = blockIdx.x * BS + threadIdx.x; t < n; t += gridDim.x * BS) {
        const int i = r == 0 ? t : cand_in[t];
        bool as;
        const double y = will_row(i, lda, NL, LL, X, Y, false, as);
        if (__double_as_longlong(y) == __double_as_longlong(Y[i])) continue;
        Y[i] = y;
        for (int e = off[i]; e < off[i + 1]; ++e) {
            const int m = readers[e];
            if (atomicMax(&stamp[m], r + 1) < r + 1) cand_out[atomicAdd(&cnt[r + 1], 1)] = m;
        }
    }
}

// Outputs from the fixed point Y: X' (optional) and B.
__global__ void __launch_bounds__(BS)
k_will_out(int N, int lda, int NL, const int *__restrict__ LL, const double *__restrict__ X,
           const double *__restrict__ Y, const double *__restrict__ R, double *__restrict__ Xout,
           double *__restrict__ B) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i >= N) return;
    bool as;
    const double y = will_row(i, lda, NL, LL, X, Y, false, as);
    if (Xout) Xout[i] = y;
    B[i] = as ? y : R[i];
}

} // namespace clowns
