// pcg_solver.hip -- host runtime of the gfx950 PCGPME solve and its C-ABI.
//
// The reference (src2LOWNS/pcgpme_para.F90) allocates its work arrays on the
// stack per call and rescales/re-sorts the matrix on every call.  Here a
// solver object owns grow-only device buffers; the scaled, sorted operator
// stays resident in HBM between solves (the SMAC pass loop of
// src2LOWNS/1sub-momentum.F90:42-60 changes only the right-hand side), and a
// solve runs entirely on one HIP stream.  Convergence is decided on the
// device (k_fin_beta); the host only polls a pinned copy of the scalars every
// few iterations, so iterations run back to back without host round trips.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <sched.h>
#include <system_error>
#include <thread>
#include <vector>

#include "clowns_pcg.h"
#include "pcg_kernels.hpp"
#include "assembly_kernels.hpp"
#include "classify_kernels.hpp"
#include "vof_kernels.hpp"

using namespace clowns;

#define HIPCHK(expr)                                                         \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            fprintf(stderr, "clowns_pcg: %s failed: %s (%s:%d)\n", #expr,    \
                    hipGetErrorString(e_), __FILE__, __LINE__);              \
            return e_ == hipErrorOutOfMemory ? CLOWNS_ENOMEM : CLOWNS_EHIP;  \
        }                                                                    \
    } while (0)

namespace {

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;  // memory the buffer does not own (never freed or grown)
    int ensure(size_t want) {
        if (want <= n) return CLOWNS_OK;
        if (borrowed) return CLOWNS_EBADARG;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e == hipErrorOutOfMemory ? CLOWNS_ENOMEM : CLOWNS_EHIP;
        }
        n = want;
        return CLOWNS_OK;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
        borrowed = false;
    }
    void borrow(T *q, size_t cap) {
        p = q;
        n = cap;
        borrowed = true;
    }
};

// Vectors of one precision for one Bi-CGSTAB run.
template <typename T> struct VecSet {
    DevBuf<T> r, rh, p0, q, e, v, y, t, x;
    int padN = -1;  // the row count whose ghost rows [N, N+64) were last zeroed
    int ensure(size_t n) {
        DevBuf<T> *all[] = {&r, &rh, &p0, &q, &e, &v, &y, &t, &x};
        for (auto *b : all) {
            int rc = b->ensure(n);
            if (rc) return rc;
        }
        return CLOWNS_OK;
    }
    void release() {
        DevBuf<T> *all[] = {&r, &rh, &p0, &q, &e, &v, &y, &t, &x};
        for (auto *b : all) b->release();
        padN = -1;
    }
};

} // namespace

// In-solve kernel durations (clowns_solver_kernel_profile): while armed,
// every launch of the solve's kernels goes through hipExtLaunchKernelGGL
// with a pair of HIP events that the runtime stamps with that dispatch's own
// start and end (the interval a kernel trace reports, without the event
// packets' own time), tagged with the kernel class, the Bi-CGSTAB run and
// its iteration; the launches a run queued past its last iteration (they
// return at once) are left out of the sums.  Disarmed (the default) it
// records nothing and the launches are plain ones.
enum KClass { KC_SPMV, KC_PM, KC_PM2, KC_PMDOT, KC_PUPD, KC_EUPD, KC_UPDATE, KC_FIN, KC_RESID, KC_N };
struct KProf {
    bool on = false;
    std::vector<hipEvent_t> pool;  // two per record
    struct Rec {
        int cls, run, K;
    };
    std::vector<Rec> rec;
    std::vector<int> kout;  // per run: its last working iteration
    int run = -1;           // run being enqueued (-1: outside run_bicgstab)
    int K = 0;              // its iteration being enqueued
    int begin(int cls) {
        const size_t i = rec.size();
        while (pool.size() < 2 * (i + 1)) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            pool.push_back(e);
        }
        rec.push_back({cls, run, K});
        return (int)i;
    }
    void release() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
        rec.clear();
        kout.clear();
    }
};
// The event pair a KScope armed for the next klaunch on this thread.
struct KSlot {
    hipEvent_t start = nullptr, stop = nullptr;
};
thread_local KSlot t_kslot;

struct KScope {
    KProf *p;
    int i = -1;
    KScope(KProf *prof, int cls) : p(prof) {
        if (p && (i = p->begin(cls)) >= 0) t_kslot = {p->pool[2 * i], p->pool[2 * i + 1]};
    }
    ~KScope() {
        if (p && t_kslot.start) {  // no launch took the events: drop the record
            if (i >= 0) p->rec[i].cls = -1;
            t_kslot = {};
        }
    }
};

// Launch f<<<g, b, 0, s>>>(a...); when a KScope armed an event pair, through
// hipExtLaunchKernelGGL so the events carry the dispatch's start and end.
template <typename... P, typename... A>
inline void klaunch(void (*f)(P...), dim3 g, dim3 b, hipStream_t s, A... a) {
    if (t_kslot.start) {
        const KSlot k = t_kslot;
        t_kslot = {};
        hipExtLaunchKernelGGL(f, g, b, 0, s, k.start, k.stop, 0, a...);
    } else {
        f<<<g, b, 0, s>>>(a...);
    }
}

// Run-time switches (environment variables), read once when a solver is
// created and kept in that solver: creating another solver never changes the
// kernel paths of a live one.
struct Knobs {
    // Pairwise fusion of single-term PM stages (k_pm2_v2) saves one pass over
    // the intermediate vector but recomputes the first stage at every gather
    // point (three dependent loads).  Which form is faster depends on the
    // operator: measured on MI355X, fused is 9% (FP64) and 12% (FP32) faster
    // per solve at 2.45M rows, but at 13.1M rows (512x512 planes) the fused
    // FP32 pair takes 83 us against 2 x 31 us.  So each solver times both
    // forms on its operator the first time it solves at a given size
    // (tune_fuse) and keeps the faster.  CLOWNS_FUSE_PM=0/1 forces both
    // precisions off/on (-1: timed).
    int fuse_pm = -1;
    // Non-temporal loads/stores of the streamed arrays in the PM stages and
    // the SpMV (the ST kernels).  Measured on MI355X, at 13.1M rows FP64 SpMV
    // 224 -> 196 us and PM stage 56 -> 51 us, but FP32 PM stage 28 -> 35 us;
    // at 2.45M rows (operator inside the 256 MB last-level cache) every
    // kernel slower.  tune_stream times both per operator and precision;
    // CLOWNS_STREAM forces them: 0 neither, 1 both, 2 the PM stages only, 3
    // the SpMV only (-1: timed).  As a mask: bit 0 PM stages, bit 1 SpMV.
    int stream = -1;
    // 8-bit coded neighbour indices (Ix8) in place of the 16-bit codes.  Half
    // the index bytes, but more decoding and a scattered 32-bit read per wide
    // slot: measured on MI355X, FP64 solve 2.6% faster at 13.1M rows and 2%
    // slower at 2.45M rows (operator inside the last-level cache), so
    // tune_stream decides per operator.  CLOWNS_LL8 = 0 off, 1 on, -1
    // (default) timed.
    int ll8 = -1;
    // 16-bit coded neighbour indices.  Measured on MI355X: single PM stage 5%
    // faster at 2.45M rows (FP64 10.05 vs 10.61 us, FP32 6.72 vs 7.10 us),
    // equal at 13.1M rows where the +-plane gathers, not the index stream,
    // set the pace.  CLOWNS_LL16 = 0 reads the 32-bit LL (and builds no
    // 8-bit codes either).
    bool ll16 = true;
    // Four rows per thread in the FP32 single-term PM stages (k_pm_v4): with
    // half-precision coefficients two rows load only 4 bytes of them.
    // Measured on MI355X: mixed solve 199.9 -> 195.3 ms at 13.1M rows (single
    // stages), unchanged at 2.45M rows (fused pairs).  CLOWNS_V4_32 = 0 keeps
    // two rows.
    bool v4_32 = true;
    // The FP32 SpMV from row-major copies of its coefficients and 16-bit
    // codes (k_spmv_rows32): three 16-byte and three 8-byte loads for two
    // rows instead of twelve smaller ones.  Measured on MI355X (mixed solve):
    // 27.9 -> 27.0 ms at 2.45M rows, 196 -> 195 ms at 13.1M rows.  With it the
    // FP32 kernels keep 16-bit codes.  CLOWNS_SPMV_ROWS = 0 keeps the
    // column-major operator.
    bool rows32 = true;
    // k_spmv_rows32 in the two-chunk form (k_spmv_rows32_p2), with 8- or
    // 16-bit codes: both chunks' pattern ids, codes and x in flight before
    // the first chunk's gathers.  Bit-identical.  Measured on MI355X at 13.1M
    // rows, 8-bit codes: 66.4 -> 58.7 us (cached), 75.5 -> 69.3 us
    // (streamed).  CLOWNS_SPMV32_P2 = 0 keeps one chunk.
    bool spmv32_p2 = true;
    // The FP32 inner solve of the mixed form applies the PM preconditioner
    // with half-precision coefficients (k_to_f16): the outer FP64 loop
    // measures the true residual with the exact FP64 preconditioner, so this
    // changes only how much each refinement gains (by about |M~ - M| |M~^-1|,
    // some 1e-3 against the inner tolerance 1e-2), not the accuracy of the
    // result; the inner operator keeps FP32.  CLOWNS_PM_HALF = 0 keeps FP32
    // coefficients.
    bool pm_half = true;
    // Block order of the FP64 gather kernels for an assembled operator (see
    // ensure_block_order).  CLOWNS_BLOCK_ORDER = 0 switches it off.
    bool block_order = true;
    // Row patterns (k_pat_assign): the SpMV of both precisions reads a
    // one-byte pattern id instead of the coefficients of rows that repeat a
    // table row.  Measured on MI355X at 13.1M rows (95% of rows on
    // patterns): FP32 SpMV 108 -> 65 us, FP64 SpMV (streamed) 198 -> 143 us;
    // at 2.45M rows FP64 SpMV 36 -> 20 us (cached).  Results are
    // bit-identical.  CLOWNS_PATTERNS = 0 switches them off.
    bool patterns = true;
    // FP64 PM stages of an operator with row patterns in the two-chunk form
    // with the pattern table (k_pm_p2): a wavefront's two chunks of loads
    // in flight before its gathers, a pattern row's coefficient from LDS.
    // Bit-identical to k_pm_v2.  tune_stream times it with the other forms
    // (bit 3); CLOWNS_PM_PAT = 0 off, 1 on, -1 (default) timed.  (The FP32
    // four-row stages with the table -- their four coefficients loaded only
    // when a row has its own -- were slower: two inner iterations 1375
    // against 1217 us at 13.1M rows; they keep the coefficient columns.)
    int pm_pat = -1;
    // CLOWNS_VERBOSE = 1: report per-operator tuning decisions on stderr.
    bool verbose = false;
};

struct clowns_solver {
    Knobs kn;
    KProf prof;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // Resident operator (reference layout after PCGPME's scale + sort).
    int N = 0, NL = 0, lda = 0;  // lda: device column stride (multiple of 64)
    int lda_user = 0;             // column stride of the caller's arrays
    double parm = 5.0;
    DevBuf<double> AL, D, W;
    DevBuf<int> LL;
    DevBuf<short> LL16;    // LL coded in 16 bits (k_ll16_build), valid if ll16_ok
    DevBuf<int2> base16;
    bool ll16_ok = false;
    int ll16_wide = 0;     // slots read from the 32-bit LL (LL16_WIDE)
    DevBuf<int> order;     // row-block order of the FP64 gather kernels (set_block_order)
    int order_G = 0;       // its length; used while it equals the operator's block count
    DevBuf<int> order_first;  // assembled operator: first cell of each row block ...
    long order_X = 0, order_Y = 0, order_Z = 0;  // ... in this box
    bool order_pending = false;  // ... to be sorted into `order` by the next FP64 solve
    DevBuf<uint8_t> LL8;   // LL coded in 8 bits (k_ll8_build), valid if ll8_ok
    DevBuf<int4> base8;
    bool ll8_ok = false;
    int ll8_wide = 0;      // slots read from the 32-bit LL (LL8_WIDE)
    DevBuf<uint8_t> pid;   // row pattern ids (k_pat_assign), valid if pat_ok
    DevBuf<double> ptab;   // pattern table, PAT_MAX rows of PAT_W values
    DevBuf<double> pat_keys;  // sampled rows for a new table
    int npat = 0;          // rows in ptab (kept for the next operator)
    int pat_rows = 0;      // rows of the resident operator on a pattern
    bool pat_ok = false;
    DevBuf<float> AL32, D32;
    DevBuf<float> ALr32;   // row-major FP32 coefficients of the FP32 SpMV (Knobs::rows32)
    DevBuf<short> Lr16;    // row-major 16-bit codes of the same
    DevBuf<uint8_t> Lr8;   // row-major 8-bit codes of the same (with LL8)
    bool have_rows8 = false;
    DevBuf<int> nonunit;   // rows whose FP32 diagonal is not 1 (k_to_rows32)
    bool have_rows = false;
    DevBuf<_Float16> ALh;  // half-precision PM coefficients of the FP32 inner solve (Knobs::pm_half)
    bool have32 = false;
    bool have_h = false;   // ALh built with the FP32 copy
    // Transposed operator (adjoint), scaled + sorted the same way.
    DevBuf<double> ALt, Dt, Wt;
    DevBuf<int> LLt;
    bool have_t = false;
    DevBuf<short> LLt16;   // its 16- and 8-bit coded LL (build_codes_t)
    DevBuf<int2> baset16;
    DevBuf<uint8_t> LLt8;
    DevBuf<int4> baset8;
    bool t16_ok = false, t8_ok = false;
    // Caller-layout copy of the unscaled operator (gradients).
    DevBuf<double> AL0, D0;
    DevBuf<int> LL0;
    bool raw_valid = false;

    VecSet<double> v64;
    VecSet<float> v32;
    DevBuf<double> bs;     // scaled right-hand side
    int bs_padN = -1;      // row count whose ghost rows of bs were last zeroed
    DevBuf<double> stage_b, stage_x;  // clowns_pcgpme host staging
    struct Stager *stager = nullptr;  // pinned chunks for the drop-in's host arrays (ensure_stager)
    struct Stager *stager_out = nullptr;  // ... and for its outputs copied during the solve
    DevBuf<double> part;   // per-block partial sums
    DevBuf<int> flag;      // device error flag
    DevBuf<int> will_flags;  // PCGPME_WILL output pass: per-round change flags / counts
    DevBuf<int> will_cand;   // PCGPME_WILL output pass: candidate lists (2N) + stamps (N)
    DevBuf<int> will_rev;    // PCGPME_WILL output pass: readers of each row (CSR)
    bool will_rev_valid = false;  // will_rev holds the readers of the resident LL0
    DevBuf<int> cls_scratch;  // classification: extents, change flag, block counts
    DevBuf<double> pc_b, pc_x, pc_part;  // pressure-correction loop: RHS, xe, DMAX partials
    DevBuf<int> vof_state;     // VOF update: quvwf_modify state per cell
    DevBuf<double> vof_term;   // VOF update: quvwf_modify correction per cell
    DevBuf<int> vof_cnt;       // VOF update: flags and counters (VC_*)
    DevBuf<Scal> sc;       // [0] inner/FP64 run, [1] outer (mixed)
    Scal *h_sc = nullptr;  // pinned mirror: [0] inner/FP64 run, [1] outer, [2], [3] poll slots
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    hipEvent_t evp[2] = {nullptr, nullptr};  // ends of the two queued iteration batches

    // mixed: smallest relative tolerance of an inner solve.  Short inner
    // solves restart Bi-CGSTAB from the true residual more often, which on
    // these operators costs fewer total iterations: measured on MI355X,
    // tau_min = 1e-2 against 1e-3: 150 vs 164 inner iterations at 2.45M rows
    // (27.6 vs 30.6 ms), 233 vs 252 at 13.1M rows (224 vs 239 ms), 156 vs 174
    // per time step at 26M rows; 2e-2 .. 1e-1 and 3e-4 were slower.
    double tau_min = 1e-2;
    bool cls_force_host = false;  // test hook: always take the exact host path
    int poll = 8;          // iterations between host polls of `done`
    int will_max_rounds = 4096;  // PCGPME_WILL output pass: device rounds before the host pass
    int fuse64 = -1, fuse32 = -1;      // pairwise PM fusion chosen by tune_fuse ...
    int fuse64_N = 0, fuse32_N = 0;    // ... for operators of about this many rows (tuned_size)
    int st64 = -1, st32 = -1;          // non-temporal streams chosen by tune_stream (bit 0 PM, bit 1 SpMV) ...
    int st64_N = 0, st32_N = 0;        // ... for operators of about this many rows (tuned_size)
    int st64_f = -1, st32_f = -1;      // ... under this forcing (force_key)
    int ix64 = -1, ix32 = -1;          // 8-bit (1) or 16-bit (0) index codes chosen by tune_stream ...
    int ix64_N = 0, ix32_N = 0;        // ... for operators of about this many rows
    int pc64 = -1, pc64_N = 0;         // FP64 PM stages as k_pm_p2 (tune_stream), for about this many rows
};

namespace {

std::mutex g_mu;      // the default solvers behind the drop-in entry points
std::mutex g_env_mu;  // reading the run-time switches
clowns_solver *g_default[64] = {};

int n_pad(int N) { return N + PAD; }

// Grid of a kernel taking R consecutive rows a thread: a multiple of 8
// workgroups (logical_block's XCD mapping).
struct Geo {
    int G, R;
};
inline Geo geo_of(int N, int R) {
    int G = cdiv(N, (long)BS * R);
    G = (G + 7) & ~7;
    return {G, R};
}
// Workgroups of k_pm_p2 for G two-row blocks: each does two of them.
inline int geo_half(int G) { return ((G + 1) / 2 + 7) & ~7; }

int ensure_events(clowns_solver *h) {
    if (!h->ev0) {
        HIPCHK(hipEventCreate(&h->ev0));
        HIPCHK(hipEventCreate(&h->ev1));
        HIPCHK(hipEventCreate(&h->ev2));
        HIPCHK(hipEventCreateWithFlags(&h->evp[0], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&h->evp[1], hipEventDisableTiming));
    }
    return CLOWNS_OK;
}

// Dot partials: two arrays of one double per block of the largest grid any
// kernel of a run launches, the one-row kernels' (k_init).  Every other
// form covers more rows per block (geo_of(N, R) with R >= 1 never exceeds
// geo_of(N, 1)).
size_t part_len(int N) { return (size_t)2 * geo_of(N, 1).G; }

int ensure_common(clowns_solver *h, int N) {
    int rc;
    if ((rc = h->part.ensure(part_len(N)))) return rc;
    if ((rc = h->sc.ensure(2))) return rc;
    if (!h->h_sc) HIPCHK(hipHostMalloc((void **)&h->h_sc, 4 * sizeof(Scal), hipHostMallocDefault));
    return ensure_events(h);
}

template <typename T> void fill(hipStream_t s, T *p, size_t n, T val) {
    if (n == 0) return;
    k_fill<T><<<cdiv((long)n, BS), BS, 0, s>>>(n, p, val);
}

// Operator views for one run.
template <typename T> struct Op {
    int N, NL, lda;
    const T *AL;
    const int *LL;
    const T *D;
    bool fuse = false;  // PM stages fused pairwise (k_pm2) in apply_pm
    bool st_pm = false, st_spmv = false;  // non-temporal streams (tune_stream)
    const short *LL16 = nullptr;  // 16-bit coded LL (two-row kernels), or null
    const int2 *base16 = nullptr;
    const uint8_t *LL8 = nullptr; // 8-bit coded LL (two-row kernels), or null
    const int4 *base8 = nullptr;
    bool ix8 = false;             // the kernels read LL8 (tune_stream), else LL16
    const int *order = nullptr;   // row-block order of the gather kernels, or null
    const _Float16 *ALh = nullptr;  // FP32 inner solve: half-precision PM coefficients, or null
    bool v4 = false;                // FP32 single PM stages with four rows a thread (k_pm_v4)
    bool p2_32 = false;             // FP32 row-major SpMV as k_spmv_rows32_p2 (two chunks a wavefront)
    const float *ALr = nullptr;     // FP32 SpMV: row-major coefficients and codes, or null
    const short *Lr16 = nullptr;
    const uint8_t *Lr8 = nullptr;   // ... 8-bit codes (with LL8), or null
    const int *nonunit = nullptr;   // ... device count of rows whose FP32 diagonal is not 1
    const uint8_t *pid = nullptr;   // row pattern ids and table (NL = 6), or null
    const double *ptab = nullptr;
    bool pc2 = false;               // FP64 PM stages as k_pm_p2 (with pid; tune_stream)
    KProf *prof = nullptr;          // armed kernel profile, or null
};

// Kernel-class tag of a launch for the profile: FP32 classes after FP64's.
template <typename T> int kcls(int c) { return c + (sizeof(T) == 4 ? KC_N : 0); }

// The index source of column j for the two-row kernels.
template <typename T> bool has16(const Op<T> &A) { return A.LL16 != nullptr; }
template <typename T> Ix32 ix32(const Op<T> &A, size_t off) { return Ix32{A.LL + off}; }
template <typename T> Ix16 ix16(const Op<T> &A, size_t off) { return Ix16{A.LL16 + off, A.LL + off, A.base16, A.N}; }
template <typename T> bool has8(const Op<T> &A) { return A.LL8 != nullptr && A.ix8; }
template <typename T> Ix8 ix8(const Op<T> &A, size_t off) { return Ix8{A.LL8 + off, A.LL + off, A.base8, A.N}; }

// Block order of the FP64 gather kernels for an assembled operator: the
// 512-row blocks sorted by (chunk of ORDER_LINES grid lines, plane, block),
// so that a row's z neighbours were streamed a few thousand rows earlier
// rather than a plane earlier and are still in L2.  Measured on MI355X
// (FP64 solve): 306 -> 294 ms at 13.1M rows, 45.3 -> 44.8 ms at 2.45M rows;
// 16-256 lines per chunk within 1% of each other at 13.1M rows, 2-8 lines
// slower than no order (the y neighbours then miss); the FP32 kernels gain
// nothing and keep the natural order.
constexpr int ORDER_LINES = 64;
constexpr int ORDER_MIN_ROWS = 1 << 20;

// Two-row PM launches for coefficient type C (T, or _Float16: the FP32
// inner solve's half-precision preconditioner) and the operator's index form.
template <typename T, typename C, typename IX>
void launch_pm2_ix(hipStream_t s, int G2, const Scal *sc, const Op<T> &A, const C *a0, IX x0, const C *a1, IX x1,
                   const T *in, T *out) {
    if (A.st_pm)
        klaunch(k_pm2_v2<T, IX, true, C>, G2, BS, s, sc, A.N, a0, x0, a1, x1, in, out, A.order);
    else
        klaunch(k_pm2_v2<T, IX, false, C>, G2, BS, s, sc, A.N, a0, x0, a1, x1, in, out, A.order);
}

template <typename T, typename C>
void launch_pm2_c(hipStream_t s, const Scal *sc, const Op<T> &A, const C *AL, int j, const T *in, T *out) {
    const size_t o0 = (size_t)j * A.lda, o1 = (size_t)(j + 1) * A.lda;
    const int G2 = geo_of(A.N, 2).G;
    if (has8(A))
        launch_pm2_ix<T, C>(s, G2, sc, A, AL + o0, ix8(A, o0), AL + o1, ix8(A, o1), in, out);
    else if (has16(A))
        launch_pm2_ix<T, C>(s, G2, sc, A, AL + o0, ix16(A, o0), AL + o1, ix16(A, o1), in, out);
    else
        launch_pm2_ix<T, C>(s, G2, sc, A, AL + o0, ix32(A, o0), AL + o1, ix32(A, o1), in, out);
}

// Stages j and j+1 of LDUSUB fused (k_pm2_v2).
template <typename T>
void launch_pm2(hipStream_t s, const Scal *sc, const Op<T> &A, int j, const T *in, T *out) {
    KScope ks(A.prof, kcls<T>(KC_PM2));
    if constexpr (sizeof(T) == 4) {
        if (A.ALh) {
            launch_pm2_c<T, _Float16>(s, sc, A, A.ALh, j, in, out);
            return;
        }
    }
    launch_pm2_c<T, T>(s, sc, A, A.AL, j, in, out);
}

template <typename T, int NT, typename IX, bool ST, typename C>
void launch_pm_v2(hipStream_t s, int G2, const Scal *sc, int N, const C *a0, IX x0, const C *a1, IX x1,
                  const T *in, T *out, int dot, const T *w, double *part, const int *order) {
    switch (dot) {
    case 0: klaunch(k_pm_v2<T, NT, 0, IX, ST, C>, G2, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part, order); break;
    case 1: klaunch(k_pm_v2<T, NT, 1, IX, ST, C>, G2, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part, order); break;
    case 2: klaunch(k_pm_v2<T, NT, 2, IX, ST, C>, G2, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part, order); break;
    default: klaunch(k_pm_v2<T, NT, 3, IX, ST, C>, G2, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part, order);
    }
}

template <int NT, typename IX, bool ST, typename C>
void launch_pm_v4(hipStream_t s, int G4, const Scal *sc, int N, const C *a0, IX x0, const C *a1, IX x1,
                  const float *in, float *out, int dot, const float *w, double *part) {
    switch (dot) {
    case 0: klaunch(k_pm_v4<NT, 0, IX, ST, C>, G4, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part); break;
    case 1: klaunch(k_pm_v4<NT, 1, IX, ST, C>, G4, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part); break;
    case 2: klaunch(k_pm_v4<NT, 2, IX, ST, C>, G4, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part); break;
    default: klaunch(k_pm_v4<NT, 3, IX, ST, C>, G4, BS, s, sc, N, a0, x0, a1, x1, in, out, w, part); break;
    }
}

template <int NT, typename IX, bool ST>
void launch_pm_p2(hipStream_t s, int G2, const Scal *sc, const Op<double> &A, const double *a0, IX x0,
                  const double *a1, IX x1, int col, const double *in, double *out, int dot, const double *w,
                  double *part) {
    const int W = geo_half(G2);
    switch (dot) {
    case 0: klaunch(k_pm_p2<NT, 0, IX, ST>, W, BS, s, sc, A.N, G2, a0, x0, a1, x1, col, in, out, w, part, A.order, A.pid, A.ptab); break;
    case 1: klaunch(k_pm_p2<NT, 1, IX, ST>, W, BS, s, sc, A.N, G2, a0, x0, a1, x1, col, in, out, w, part, A.order, A.pid, A.ptab); break;
    case 2: klaunch(k_pm_p2<NT, 2, IX, ST>, W, BS, s, sc, A.N, G2, a0, x0, a1, x1, col, in, out, w, part, A.order, A.pid, A.ptab); break;
    default: klaunch(k_pm_p2<NT, 3, IX, ST>, W, BS, s, sc, A.N, G2, a0, x0, a1, x1, col, in, out, w, part, A.order, A.pid, A.ptab);
    }
}

// One PM stage in the form for T and the operator; returns the grid (the
// number of dot partials a dot-carrying stage wrote).  col: the stage's
// coefficient column.
template <typename T, int NT, typename C, typename IX>
int launch_pm_stage_ix(hipStream_t s, int G2, const Scal *sc, const Op<T> &A, const C *a0, IX x0, const C *a1,
                        IX x1, int col, const T *in, T *out, int dot, const T *w, double *part) {
    if constexpr (sizeof(T) == 4) {
        // the dot-carrying stages always take four rows a thread; the single
        // stages unless CLOWNS_V4_32 = 0 (bit-identical either way)
        if (dot != 0 || A.v4) {
            const int G4 = geo_of(A.N, 4).G;
            if (A.st_pm)
                launch_pm_v4<NT, IX, true, C>(s, G4, sc, A.N, a0, x0, a1, x1, in, out, dot, w, part);
            else
                launch_pm_v4<NT, IX, false, C>(s, G4, sc, A.N, a0, x0, a1, x1, in, out, dot, w, part);
            return G4;
        }
    }
    if constexpr (std::is_same<T, double>::value && std::is_same<C, double>::value) {
        if (A.pc2 && A.pid) {
            if (A.st_pm)
                launch_pm_p2<NT, IX, true>(s, G2, sc, A, a0, x0, a1, x1, col, in, out, dot, w, part);
            else
                launch_pm_p2<NT, IX, false>(s, G2, sc, A, a0, x0, a1, x1, col, in, out, dot, w, part);
            return G2;
        }
    }
    if (A.st_pm)
        launch_pm_v2<T, NT, IX, true, C>(s, G2, sc, A.N, a0, x0, a1, x1, in, out, dot, w, part, A.order);
    else
        launch_pm_v2<T, NT, IX, false, C>(s, G2, sc, A.N, a0, x0, a1, x1, in, out, dot, w, part, A.order);
    return G2;
}

template <typename T, int NT, typename C>
int launch_pm_stage_c(hipStream_t s, const Scal *sc, const Op<T> &A, const C *AL, int j, const T *in, T *out,
                      int dot, const T *w, double *part) {
    const size_t o0 = (size_t)j * A.lda, o1 = (size_t)(j + 1) * A.lda;
    const C *a0 = AL + o0, *a1 = NT == 2 ? AL + o1 : nullptr;
    const int G2 = geo_of(A.N, 2).G;
    if (has8(A)) {
        const Ix8 x0 = ix8(A, o0), x1 = NT == 2 ? ix8(A, o1) : x0;
        return launch_pm_stage_ix<T, NT, C>(s, G2, sc, A, a0, x0, a1, x1, j, in, out, dot, w, part);
    }
    if (has16(A)) {
        const Ix16 x0 = ix16(A, o0), x1 = NT == 2 ? ix16(A, o1) : x0;
        return launch_pm_stage_ix<T, NT, C>(s, G2, sc, A, a0, x0, a1, x1, j, in, out, dot, w, part);
    }
    const Ix32 x0 = ix32(A, o0), x1 = NT == 2 ? ix32(A, o1) : x0;
    return launch_pm_stage_ix<T, NT, C>(s, G2, sc, A, a0, x0, a1, x1, j, in, out, dot, w, part);
}

// One stage of LDUSUB (NT = 2: the two-term stage 5/6); returns the grid
// whose partials a dot-carrying stage wrote.
template <typename T, int NT>
int launch_pm_stage(hipStream_t s, const Scal *sc, const Op<T> &A, int j, const T *in, T *out, int dot,
                    const T *w, double *part) {
    KScope ks(A.prof, kcls<T>(NT == 2 || dot != 0 ? KC_PMDOT : KC_PM));
    if constexpr (sizeof(T) == 4) {
        if (A.ALh) return launch_pm_stage_c<T, NT, _Float16>(s, sc, A, A.ALh, j, in, out, dot, w, part);
    }
    return launch_pm_stage_c<T, NT, T>(s, sc, A, A.AL, j, in, out, dot, w, part);
}

// LDUSUB (pcgpme_para.F90:263-320): out = PM*in through scratch y, t.
// `in` may be `out`; neither may be y or t except in == y.
template <typename T>
int apply_pm(hipStream_t s, const Scal *sc, const Op<T> &A, const T *in, T *out, T *y, T *t, int dot, const T *w,
             double *part) {
    if (A.fuse) {
        // stages (1,2) and (3,4) fused pairwise; 5/6 (two-term) carries the dots
        launch_pm2<T>(s, sc, A, 0, in, t);
        if (A.NL == 6) {
            launch_pm2<T>(s, sc, A, 2, t, y);
            return launch_pm_stage<T, 2>(s, sc, A, 4, y, out, dot, w, part);
        }
        // NL = 4: stage 3 alone, then stage 4 with the dots
        launch_pm_stage<T, 1>(s, sc, A, 2, t, y, 0, w, part);
        return launch_pm_stage<T, 1>(s, sc, A, 3, y, out, dot, w, part);
    }
    if (A.NL == 6) {
        launch_pm_stage<T, 1>(s, sc, A, 0, in, t, 0, w, part);
        launch_pm_stage<T, 1>(s, sc, A, 1, t, y, 0, w, part);
        launch_pm_stage<T, 1>(s, sc, A, 2, y, t, 0, w, part);
        launch_pm_stage<T, 1>(s, sc, A, 3, t, y, 0, w, part);
        return launch_pm_stage<T, 2>(s, sc, A, 4, y, out, dot, w, part);
    }
    launch_pm_stage<T, 1>(s, sc, A, 0, in, t, 0, w, part);
    launch_pm_stage<T, 1>(s, sc, A, 1, t, y, 0, w, part);
    launch_pm_stage<T, 1>(s, sc, A, 2, y, t, 0, w, part);
    return launch_pm_stage<T, 1>(s, sc, A, 3, t, out, dot, w, part);
}

// Time the first four PM stages of an LDUSUB application as two fused pairs
// and as four single stages, on the operator, as the solve runs them
// (alternating vectors; HIP events on the solver's stream; the two forms
// interleaved three times, the best of each kept), and keep the faster form
// for operators of this size.  NL = 4 operators time stages 1-2 only (their
// apply fuses one pair).  V's y, t, e are scratch here; the run's scalars are
// reset (kernels skip work once `done` is set).
// A tuning decision holds for operators within 1/16 of the size it was
// timed at: a time-stepped run whose row count moves by a few rows per step
// (the free surface) does not time again.
inline bool tuned_size(int N, int decN) { return decN > 0 && std::abs(N - decN) <= decN / 16; }

template <typename T> bool tune_fuse(clowns_solver *h, const Op<T> &A, VecSet<T> &V) {
    if (h->kn.fuse_pm >= 0) return h->kn.fuse_pm == 1;
    int &dec = sizeof(T) == 4 ? h->fuse32 : h->fuse64;
    int &decN = sizeof(T) == 4 ? h->fuse32_N : h->fuse64_N;
    if (dec >= 0 && tuned_size(A.N, decN)) return dec == 1;
    hipStream_t s = h->stream;
    Scal *sc = h->sc.p;
    const int npairs = A.NL == 6 ? 2 : 1;
    k_reset<<<1, 64, 0, s>>>(sc, 0.0, 1, 0);
    auto fused = [&]() {
        launch_pm2<T>(s, sc, A, 0, V.y.p, V.t.p);
        if (npairs == 2) launch_pm2<T>(s, sc, A, 2, V.t.p, V.e.p);
    };
    auto single = [&]() {
        launch_pm_stage<T, 1>(s, sc, A, 0, V.y.p, V.t.p, 0, nullptr, h->part.p);
        launch_pm_stage<T, 1>(s, sc, A, 1, V.t.p, V.e.p, 0, nullptr, h->part.p);
        if (npairs == 2) {
            launch_pm_stage<T, 1>(s, sc, A, 2, V.e.p, V.t.p, 0, nullptr, h->part.p);
            launch_pm_stage<T, 1>(s, sc, A, 3, V.t.p, V.e.p, 0, nullptr, h->part.p);
        }
    };
    float best[2] = {1e30f, 1e30f};
    fused();
    single();
    for (int round = 0; round < 3; ++round) {
        for (int form = 0; form < 2; ++form) {
            float ms = 0.f;
            if (hipEventRecord(h->ev0, s) != hipSuccess) return false;
            form == 0 ? fused() : single();
            if (hipEventRecord(h->ev1, s) != hipSuccess || hipEventSynchronize(h->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess)
                return false;
            best[form] = std::min(best[form], ms);
        }
    }
    dec = best[0] < best[1] ? 1 : 0;
    decN = A.N;
    if (h->kn.verbose)
        fprintf(stderr, "clowns_pcg: N=%d %s PM stages 1-%d: fused pairs %.2f us, singles %.2f us -> %s\n", A.N,
                sizeof(T) == 4 ? "FP32" : "FP64", 2 * npairs, 1e3f * best[0], 1e3f * best[1],
                dec ? "fused" : "single");
    return dec == 1;
}

template <typename T, int NL, typename IX, bool ST>
void launch_spmv_v2_st(hipStream_t s, const Scal *sc, const Op<T> &A, IX ix, int mode, const T *x, const T *v1,
                       T *y) {
    const int G2 = geo_of(A.N, 2).G;
    if constexpr (NL == 6) {
        if (A.pid) {
            if (mode == 0)
                klaunch(k_spmv_v2<T, 6, 0, IX, ST, true>, G2, BS, s, sc, A.N, A.lda, A.AL, ix, A.D, x, v1, y, A.order,
                        A.pid, A.ptab);
            else
                klaunch(k_spmv_v2<T, 6, 3, IX, ST, true>, G2, BS, s, sc, A.N, A.lda, A.AL, ix, A.D, x, v1, y, A.order,
                        A.pid, A.ptab);
            return;
        }
    }
    if (mode == 0)
        klaunch(k_spmv_v2<T, NL, 0, IX, ST>, G2, BS, s, sc, A.N, A.lda, A.AL, ix, A.D, x, v1, y, A.order, nullptr,
                nullptr);
    else
        klaunch(k_spmv_v2<T, NL, 3, IX, ST>, G2, BS, s, sc, A.N, A.lda, A.AL, ix, A.D, x, v1, y, A.order, nullptr,
                nullptr);
}

template <typename IX, typename CT, bool ST>
void launch_rows32_st(hipStream_t s, int G2, const Scal *sc, const Op<float> &A, IX ix, const CT *Lr,
                      const float *x, float *y) {
    if (A.p2_32) {
        const int W = geo_half(G2);
        if (A.pid)
            klaunch(k_spmv_rows32_p2<IX, CT, ST, true>, W, BS, s, sc, A.N, A.lda, A.ALr, Lr, ix, A.D, A.nonunit, x,
                    y, A.pid, A.ptab);
        else
            klaunch(k_spmv_rows32_p2<IX, CT, ST, false>, W, BS, s, sc, A.N, A.lda, A.ALr, Lr, ix, A.D, A.nonunit, x,
                    y, nullptr, nullptr);
        return;
    }
    if (A.pid)
        klaunch(k_spmv_rows32<IX, CT, ST, true>, G2, BS, s, sc, A.N, A.lda, A.ALr, Lr, ix, A.D, A.nonunit, x, y, A.pid,
                A.ptab);
    else
        klaunch(k_spmv_rows32<IX, CT, ST, false>, G2, BS, s, sc, A.N, A.lda, A.ALr, Lr, ix, A.D, A.nonunit, x, y, nullptr,
                nullptr);
}

template <typename IX, typename CT>
void launch_rows32(hipStream_t s, const Scal *sc, const Op<float> &A, IX ix, const CT *Lr, const float *x,
                   float *y) {
    const int G2 = geo_of(A.N, 2).G;
    if (A.st_spmv)
        launch_rows32_st<IX, CT, true>(s, G2, sc, A, ix, Lr, x, y);
    else
        launch_rows32_st<IX, CT, false>(s, G2, sc, A, ix, Lr, x, y);
}

template <typename T, int NL, typename IX>
void launch_spmv_v2(hipStream_t s, const Scal *sc, const Op<T> &A, IX ix, int mode, const T *x, const T *v1,
                    T *y) {
    if constexpr (sizeof(T) == 4 && NL == 6 && !std::is_same<IX, Ix32>::value) {
        if (mode == 0 && A.ALr) {
            if constexpr (std::is_same<IX, Ix8>::value) {
                if (A.Lr8) return launch_rows32(s, sc, A, ix, A.Lr8, x, y);
            } else {
                return launch_rows32(s, sc, A, ix, A.Lr16, x, y);
            }
        }
    }
    if (A.st_spmv)
        launch_spmv_v2_st<T, NL, IX, true>(s, sc, A, ix, mode, x, v1, y);
    else
        launch_spmv_v2_st<T, NL, IX, false>(s, sc, A, ix, mode, x, v1, y);
}

// AXSUB: y = A x (mode 0) or the true residual y = v1 - A x (mode 3).
template <typename T>
void spmv(hipStream_t s, const Scal *sc, const Op<T> &A, int mode, const T *x, const T *v1, T *y) {
    KScope ks(A.prof, kcls<T>(mode == 3 ? KC_RESID : KC_SPMV));
    if (A.NL == 6) {
        if (has8(A))
            launch_spmv_v2<T, 6, Ix8>(s, sc, A, ix8(A, 0), mode, x, v1, y);
        else if (has16(A))
            launch_spmv_v2<T, 6, Ix16>(s, sc, A, ix16(A, 0), mode, x, v1, y);
        else
            launch_spmv_v2<T, 6, Ix32>(s, sc, A, ix32(A, 0), mode, x, v1, y);
    } else {
        if (has8(A))
            launch_spmv_v2<T, 4, Ix8>(s, sc, A, ix8(A, 0), mode, x, v1, y);
        else if (has16(A))
            launch_spmv_v2<T, 4, Ix16>(s, sc, A, ix16(A, 0), mode, x, v1, y);
        else
            launch_spmv_v2<T, 4, Ix32>(s, sc, A, ix32(A, 0), mode, x, v1, y);
    }
}

// Kernel forms as bits: bit 0 streamed PM stages, bit 1 streamed SpMV, bit 2
// 8-bit index codes (only where the operator has them), bit 3 the FP64 PM
// stages as k_pm_p2 (only FP64 operators with row patterns).
constexpr int N_FORMS = 16;
template <typename T> bool can_ix8(const Op<T> &A) { return A.LL8 != nullptr; }
template <typename T> bool can_pc2(const Op<T> &A) { return sizeof(T) == 8 && A.pid != nullptr && A.NL == 6; }
template <typename T> void set_forms(Op<T> &A, int f) {
    A.st_pm = (f & 1) != 0;
    A.st_spmv = (f & 2) != 0;
    A.ix8 = (f & 4) != 0 && can_ix8(A);
    A.pc2 = (f & 8) != 0 && can_pc2(A);
}
// The environment's forcing, which the stored decisions are keyed on.
inline int force_key(const clowns_solver *h) {
    return (h->kn.stream + 1) | ((h->kn.ll8 + 1) << 2) | ((h->kn.pm_pat + 1) << 4);
}
// Bits fixed for operator A (mask) and their values.
template <typename T> void fixed_forms(const clowns_solver *h, const Op<T> &A, int &mask, int &val) {
    mask = val = 0;
    if (h->kn.stream >= 0) mask |= 3, val |= h->kn.stream;
    if (!can_ix8(A) || h->kn.ll8 >= 0) mask |= 4, val |= (can_ix8(A) && h->kn.ll8 == 1) ? 4 : 0;
    if (!can_pc2(A) || h->kn.pm_pat >= 0) mask |= 8, val |= (can_pc2(A) && h->kn.pm_pat == 1) ? 8 : 0;
}
// The stored decisions, if they hold for A: forms in *f.  The index-code
// decision is kept apart from the streamed forms', so an operator without
// 8-bit codes (the adjoint's) reuses the streamed forms of its forward
// operator.
template <typename T> bool stored_forms(const clowns_solver *h, const Op<T> &A, int *f) {
    int mask, val;
    fixed_forms(h, A, mask, val);
    const bool f32 = sizeof(T) == 4;
    const int st = f32 ? h->st32 : h->st64, stN = f32 ? h->st32_N : h->st64_N;
    const int stF = f32 ? h->st32_f : h->st64_f;
    const int ix = f32 ? h->ix32 : h->ix64, ixN = f32 ? h->ix32_N : h->ix64_N;
    const bool key = stF == force_key(h);
    const bool st_ok = (mask & 3) == 3 || (st >= 0 && tuned_size(A.N, stN) && key);
    const bool ix_ok = (mask & 4) != 0 || (ix >= 0 && tuned_size(A.N, ixN) && key);
    const bool pc_ok = (mask & 8) != 0 || (h->pc64 >= 0 && tuned_size(A.N, h->pc64_N) && key);
    if (!st_ok || !ix_ok || !pc_ok) return false;
    *f = ((mask & 3) == 3 ? (val & 3) : st) | ((mask & 4) ? (val & 4) : 4 * ix) |
         ((mask & 8) ? (val & 8) : 8 * h->pc64);
    return true;
}

// Chooses, once per size and precision, the non-temporal (ST) or cached form
// of the PM stages and of the SpMV, and the 8- or 16-bit index codes, for
// this operator (A.fuse already decided).  What the forms leave in the
// last-level cache decides between them, so each combination runs a
// steady-state block of its own: two untimed, then three timed sequences of
// two iterations' kernels, the combinations in turn, the whole twice; the
// fastest mean wins.  (Interleaving single applications of different forms
// favours the streamed form, which finds data the cached form left behind.)
template <typename T> void tune_stream(clowns_solver *h, Op<T> &A, VecSet<T> &V) {
    int mask, fval, have;
    fixed_forms(h, A, mask, fval);
    if (stored_forms(h, A, &have)) {
        set_forms(A, have);
        return;
    }
    const bool f32 = sizeof(T) == 4;
    int &st = f32 ? h->st32 : h->st64;
    int &stN = f32 ? h->st32_N : h->st64_N;
    int &stF = f32 ? h->st32_f : h->st64_f;
    int &ix = f32 ? h->ix32 : h->ix64;
    int &ixN = f32 ? h->ix32_N : h->ix64_N;
    hipStream_t s = h->stream;
    Scal *sc = h->sc.p;
    k_reset<<<1, 64, 0, s>>>(sc, 0.0, 1, 0);
    // two Bi-CGSTAB iterations' kernels in run_bicgstab's order, without the
    // one-block finalizations (the vectors are scratch here: the solve
    // initializes every one of them after tuning)
    const int Gv = geo_of(A.N, VecOf<T>::R).G;
    double *part = h->part.p;
    auto seq = [&](int form) {
        Op<T> B = A;
        set_forms(B, form);
        for (int it = 0; it < 2; ++it) {
            k_pupd_vec<T><<<Gv, BS, 0, s>>>(sc, A.N, V.p0.p, V.r.p, V.q.p);
            spmv<T>(s, sc, B, 0, V.p0.p, nullptr, V.y.p);
            apply_pm<T>(s, sc, B, V.y.p, V.q.p, V.y.p, V.t.p, 1, V.rh.p, part);
            k_eupd_vec<T><<<Gv, BS, 0, s>>>(sc, A.N, V.r.p, V.q.p, V.e.p);
            spmv<T>(s, sc, B, 0, V.e.p, nullptr, V.y.p);
            apply_pm<T>(s, sc, B, V.y.p, V.v.p, V.y.p, V.t.p, 2, V.e.p, part);
            k_update_vec<T><<<Gv, BS, 0, s>>>(sc, A.N, V.x.p, V.p0.p, V.e.p, V.v.p, V.r.p, V.rh.p, part);
        }
    };
    float best[N_FORMS];
    for (float &b : best) b = 1e30f;
    bool ok = true;
    for (int pass = 0; pass < 2 && ok; ++pass)
        for (int f = 0; f < N_FORMS && ok; ++f) {
            if ((f & mask) != fval) continue;
            seq(f);
            seq(f);
            float ms = 0.f;
            ok = hipEventRecord(h->ev0, s) == hipSuccess;
            if (!ok) break;
            for (int r = 0; r < 3; ++r) seq(f);
            ok = hipEventRecord(h->ev1, s) == hipSuccess && hipEventSynchronize(h->ev1) == hipSuccess &&
                 hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess;
            best[f] = std::min(best[f], ms / 3.f);
        }
    if (!ok) {
        set_forms(A, fval);
        return;
    }
    int f_best = fval;
    for (int f = 0; f < N_FORMS; ++f)
        if (best[f] < best[f_best]) f_best = f;
    if ((mask & 3) != 3) {
        st = f_best & 3;
        stN = A.N;
    }
    if (!(mask & 4)) {
        ix = (f_best >> 2) & 1;
        ixN = A.N;
    }
    if (!(mask & 8)) {
        h->pc64 = (f_best >> 3) & 1;
        h->pc64_N = A.N;
    }
    stF = force_key(h);
    if (h->kn.verbose) {
        fprintf(stderr, "clowns_pcg: N=%d %s two iterations (us):", A.N, f32 ? "FP32" : "FP64");
        for (int f = 0; f < N_FORMS; ++f)
            if (best[f] < 1e29f)
                fprintf(stderr, " [%s%s%s%s] %.1f", (f & 4) ? "Ix8 " : "", (f & 8) ? "pat " : "",
                        (f & 1) ? "PM-st " : "PM ", (f & 2) ? "SpMV-st" : "SpMV", 1e3f * best[f]);
        fprintf(stderr, " -> %s, PM %s%s, SpMV %s\n", (f_best & 4) ? "8-bit" : "16-bit",
                (f_best & 1) ? "streamed" : "cached", (f_best & 8) ? " with the pattern table" : "",
                (f_best & 2) ? "streamed" : "cached");
    }
    set_forms(A, f_best);
}

// The decisions already taken for this operator (time_kernel reports the
// forms the solver runs).
template <typename T> void known_stream(const clowns_solver *h, Op<T> &A) {
    int f;
    if (stored_forms(h, A, &f)) set_forms(A, f);
}

// How many iterations to enqueue next.  The device decides convergence
// (k_fin_beta sets `done`, after which every kernel returns at once), so the
// host only chooses how far ahead of the last status it has seen to run:
// too short and the GPU idles while the host waits, too long and the
// iterations queued past convergence each cost a dozen empty launches.  From
// the residual at the last two statuses the host predicts the iteration at
// which sqrt(RN/BN) reaches EPS (Bi-CGSTAB converges about geometrically)
// and queues up to it; without a prediction the batch doubles, up to 32.
struct PollPlan {
    int n = 8;                          // last batch length
    int k0 = 0, k1 = 0;                 // iterations of the last two statuses seen
    double e0 = 0.0, e1 = 0.0;          // their ERR
    // status after iteration k: ERR err, target eps
    void seen(int k, double err) {
        if (k1 > 0 && k - k1 < 4) return;  // too close to the last one for a slope
        k0 = k1;
        e0 = e1;
        k1 = k;
        e1 = err;
    }
    // length of the batch that starts at iteration `next` (statuses seen up
    // to k1; iterations k1+1 .. next-1 already queued)
    int next_len(int next, double eps) {
        int len = std::min(32, 2 * n);
        if (k0 > 0 && e0 > 0.0 && e1 > 0.0 && e1 < e0 && eps > 0.0) {
            const double rate = std::log(e0 / e1) / (double)(k1 - k0);
            const double left = e1 > eps ? std::log(e1 / eps) / rate : 0.0;
            const double stop = (double)k1 + left;  // predicted last iteration
            const double want = std::ceil(stop - (double)next + 1.0);
            len = !(want < 32.0) ? 32 : (want <= 2.0 ? 2 : (int)want);  // NaN: 32
        }
        n = len;
        return len;
    }
};

// The PCGPME iteration (pcgpme_para.F90:148-247) on already initialised
// vectors: V.p0 = V.r = V.rh = r0, V.x = x0, scalars BN/C1 set by k_fin_init.
// Batches of iterations are queued one ahead of the status the host reads
// (two pinned status slots), so the stream never drains while the host
// looks.  Returns after the device has set `done`; the final scalars are in
// h->h_sc[sc_index].  eps: the run's stopping value (for the batch plan only).
template <typename T>
int run_bicgstab(clowns_solver *h, const Op<T> &A, VecSet<T> &V, Scal *sc, int itermax, int sc_index,
                 double eps) {
    hipStream_t s = h->stream;
    double *part = h->part.p;
    const int Gv = geo_of(A.N, VecOf<T>::R).G;
    KProf *pf = A.prof;
    if (pf) pf->run = (int)pf->kout.size();
    auto iteration = [&](int K) {
        if (pf) pf->K = K;
        // Q = PM*A*P (:149-151), after P = R + BETA*(P - AMU*Q) (:244) for K > 1.
        if (K > 1) {
            KScope ks(pf, kcls<T>(KC_PUPD));
            klaunch(k_pupd_vec<T>, Gv, BS, s, sc, A.N, V.p0.p, V.r.p, V.q.p);
        }
        spmv<T>(s, sc, A, 0, V.p0.p, nullptr, V.y.p);
        const int Gq = apply_pm<T>(s, sc, A, V.y.p, V.q.p, V.y.p, V.t.p, 1, V.rh.p, part);
        {
            KScope ks(pf, kcls<T>(KC_FIN));
            klaunch(k_fin_alpha, 1, FIN_BS, s, sc, part, Gq);
        }
        // V = PM*A*E with E = R - ALP*Q (:173-181).
        {
            KScope ks(pf, kcls<T>(KC_EUPD));
            klaunch(k_eupd_vec<T>, Gv, BS, s, sc, A.N, V.r.p, V.q.p, V.e.p);
        }
        spmv<T>(s, sc, A, 0, V.e.p, nullptr, V.y.p);
        const int Gw = apply_pm<T>(s, sc, A, V.y.p, V.v.p, V.y.p, V.t.p, 2, V.e.p, part);
        {
            KScope ks(pf, kcls<T>(KC_FIN));
            klaunch(k_fin_omega, 1, FIN_BS, s, sc, part, Gw);
        }
        {
            KScope ks(pf, kcls<T>(KC_UPDATE));
            klaunch(k_update_vec<T>, Gv, BS, s, sc, A.N, V.x.p, V.p0.p, V.e.p, V.v.p, V.r.p, V.rh.p, part);
        }
        KScope ks(pf, kcls<T>(KC_FIN));
        klaunch(k_fin_beta, 1, FIN_BS, s, sc, part, Gv, K);
    };
    Scal *slot[2] = {h->h_sc + 2, h->h_sc + 3};
    PollPlan plan;
    plan.n = std::max(1, h->poll) / 2;
    int K = 1;
    auto enqueue = [&](int len, int si) -> int {
        const int kend = std::min(itermax, K + len - 1);
        for (; K <= kend; ++K) iteration(K);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(slot[si], sc, sizeof(Scal), hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(h->evp[si], s));
        return CLOWNS_OK;
    };
    int rc, cur = 0;
    if ((rc = enqueue(plan.next_len(K, eps), cur))) return rc;
    for (;;) {
        const bool ahead = K <= itermax;
        if (ahead && (rc = enqueue(plan.next_len(K, eps), cur ^ 1))) return rc;
        HIPCHK(hipEventSynchronize(h->evp[cur]));
        const Scal hs = *slot[cur];
        if (hs.done) {
            // a batch queued behind the converged one only returns at once
            h->h_sc[sc_index] = hs;
            if (ahead) {
                HIPCHK(hipEventSynchronize(h->evp[cur ^ 1]));
                h->h_sc[sc_index] = *slot[cur ^ 1];
            }
            break;
        }
        if (!ahead) {
            // unreachable: the device sets done at itermax
            h->h_sc[sc_index] = hs;
            break;
        }
        plan.seen(hs.K, hs.ERR);
        cur ^= 1;
    }
    if (pf) {
        pf->kout.push_back(h->h_sc[sc_index].kout);
        pf->run = -1;
        pf->K = 0;
    }
    return CLOWNS_OK;
}

int check_nl(int NL) { return (NL == 6 || NL == 4) ? CLOWNS_OK : CLOWNS_EBADNL; }

int build_ll16(clowns_solver *h);
int build_ll8(clowns_solver *h);
int build_patterns(clowns_solver *h);
int ensure_block_order(clowns_solver *h);

// Load AL/LL/D (reference layout, stride lda) from `kind` memory and scale +
// sort in place on the device.
} // namespace

// ------------------------------------------------------- host staging --
// The drop-in entry points (clowns_pcgpme, clowns_pcgpme_will) take the
// caller's arrays in pageable host memory (Fortran allocatables), which the
// runtime copies at 28.6 GB/s to the device and 18 GB/s back (measured on
// MI355X, 1.2 GB).  Through two pinned chunks that host threads fill (or
// drain) while the DMA of the other chunk runs, both directions come near
// the pinned DMA rate (57 GB/s; 16 host threads copy into pinned memory at
// 66 GB/s).
struct Stager {
    static constexpr size_t CHUNK = (size_t)32 << 20;  // bytes per pinned chunk
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};  // a DMA from or into pin[b] was queued
    int next = 0;                   // chunk counter (buffer = next & 1)
    char *pdst = nullptr;           // device-to-host chunk waiting in pin[pb]
    size_t plen = 0;
    int pb = -1;
    bool pooled = true;             // host copies over the host threads (else the calling thread's)
    hipStream_t cs = nullptr;       // a copy stream of its own (default priority), or null
    hipEvent_t ready = nullptr;     // with cs: what cs waits for before its first copy
};

namespace {

// Fork-join work over host threads kept for the life of the process (as
// many as the process may run on, at most 16; the calling thread takes the
// first part).  Jobs come one at a time from the drop-in entry points, which
// hold g_mu.
class HostPool {
public:
    explicit HostPool(int n) {
        // as many threads as the system gives (the calling thread is part 0)
        int made = 1;
        try {
            for (; made < std::max(1, n); ++made) th_.emplace_back([this, made] { loop(made); });
        } catch (const std::system_error &) {
        }
        nt_ = made;
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    // f(a, b) over [0, n) in nt_ contiguous parts (multiples of `align`)
    void run(size_t n, size_t align, const std::function<void(size_t, size_t)> &f) {
        {
            std::lock_guard<std::mutex> l(mu_);
            f_ = &f;
            n_ = n;
            align_ = align;
            pending_ = nt_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0, f, n, align);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return pending_ == 0; });
    }
    void copy(char *dst, const char *src, size_t n) {
        run(n, 64, [&](size_t a, size_t b) { std::memcpy(dst + a, src + a, b - a); });
    }

private:
    void part(int t, const std::function<void(size_t, size_t)> &f, size_t n, size_t align) const {
        const size_t per = ((n + nt_ - 1) / nt_ + align - 1) / align * align;
        const size_t a = std::min(n, per * (size_t)t), b = std::min(n, a + per);
        if (b > a) f(a, b);
    }
    void loop(int t) {
        unsigned long seen = 0;
        for (;;) {
            const std::function<void(size_t, size_t)> *f;
            size_t n, align;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                f = f_;
                n = n_;
                align = align_;
            }
            part(t, *f, n, align);
            std::lock_guard<std::mutex> l(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    int nt_ = 1;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t, size_t)> *f_ = nullptr;
    size_t n_ = 0, align_ = 1;
    int pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
};

HostPool &host_pool() {
    static HostPool pool([] {
        cpu_set_t set;
        const int n = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : 8;
        return std::min(16, std::max(1, n));
    }());
    return pool;
}

// Whether any of x[0, n) is not zero, over the host threads.
bool any_nonzero(const double *x, size_t n) {
    std::atomic<bool> found(false);
    host_pool().run(n, 1, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i += 4096) {
            if (found.load(std::memory_order_relaxed)) return;
            const size_t e = std::min(b, i + 4096);
            for (size_t j = i; j < e; ++j)
                if (x[j] != 0.0) {
                    found.store(true, std::memory_order_relaxed);
                    return;
                }
        }
    });
    return found.load();
}

void free_stager(Stager *st) {
    for (int b = 0; b < 2; ++b) {
        if (st->busy[b]) (void)hipEventSynchronize(st->ev[b]);
        if (st->pin[b]) (void)hipHostFree(st->pin[b]);
        if (st->ev[b]) (void)hipEventDestroy(st->ev[b]);
    }
    if (st->ready) (void)hipEventDestroy(st->ready);
    if (st->cs) (void)hipStreamDestroy(st->cs);
    delete st;
}

// The stager in `slot`, made on first use.  pooled: host copies over the
// host threads; own_stream: a copy stream of its own (for copies that run
// beside the solver's stream).
int ensure_stager(Stager *&slot, bool pooled, bool own_stream) {
    if (slot) return CLOWNS_OK;
    Stager *st = new (std::nothrow) Stager();
    if (!st) return CLOWNS_ENOMEM;
    st->pooled = pooled;
    bool ok = true;
    for (int b = 0; b < 2 && ok; ++b)
        ok = hipHostMalloc((void **)&st->pin[b], Stager::CHUNK) == hipSuccess &&
             hipEventCreateWithFlags(&st->ev[b], hipEventDisableTiming) == hipSuccess;
    if (ok && own_stream)
        ok = hipStreamCreateWithFlags(&st->cs, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&st->ready, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        free_stager(st);
        return CLOWNS_ENOMEM;
    }
    slot = st;
    return CLOWNS_OK;
}

void release_stager(clowns_solver *h) {
    for (Stager **slot : {&h->stager, &h->stager_out})
        if (*slot) {
            free_stager(*slot);
            *slot = nullptr;
        }
}

void stage_copy(const Stager *st, char *dst, const char *src, size_t n) {
    if (st->pooled)
        host_pool().copy(dst, src, n);
    else
        std::memcpy(dst, src, n);
}

// Drains the device-to-host chunk still in a pinned buffer.
int stage_finish(Stager *st) {
    if (st->pb < 0) return CLOWNS_OK;
    HIPCHK(hipEventSynchronize(st->ev[st->pb]));
    stage_copy(st, st->pdst, st->pin[st->pb], st->plen);
    st->pb = -1;
    return CLOWNS_OK;
}

// Queues src[0, n) (host) -> dst (device) on s through the pinned chunks.
int stage_h2d(Stager *st, hipStream_t s, void *dst, const void *src, size_t n) {
    int rc;
    if ((rc = stage_finish(st))) return rc;
    for (size_t off = 0; off < n; off += Stager::CHUNK) {
        const size_t len = std::min(Stager::CHUNK, n - off);
        const int b = st->next++ & 1;
        if (st->busy[b]) HIPCHK(hipEventSynchronize(st->ev[b]));
        stage_copy(st, st->pin[b], (const char *)src + off, len);
        HIPCHK(hipMemcpyAsync((char *)dst + off, st->pin[b], len, hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(st->ev[b], s));
        st->busy[b] = true;
    }
    return CLOWNS_OK;
}

// Queues src[0, n) (device, after the stream's earlier work) -> dst (host);
// each chunk is copied out of its pinned buffer while the next one's DMA
// runs, the last one by stage_finish.
int stage_d2h(Stager *st, hipStream_t s, void *dst, const void *src, size_t n) {
    for (size_t off = 0; off < n; off += Stager::CHUNK) {
        const size_t len = std::min(Stager::CHUNK, n - off);
        const int b = st->next++ & 1;  // the pending chunk, if any, is in pin[b ^ 1]
        if (st->busy[b]) HIPCHK(hipEventSynchronize(st->ev[b]));
        HIPCHK(hipMemcpyAsync(st->pin[b], (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(st->ev[b], s));
        st->busy[b] = true;
        int rc;
        if ((rc = stage_finish(st))) return rc;
        st->pdst = (char *)dst + off;
        st->plen = len;
        st->pb = b;
    }
    return CLOWNS_OK;
}

// NL columns of N elements of size es: host column stride ldh, device ldd.
int stage_h2d_cols(Stager *st, hipStream_t s, void *dst, size_t ldd, const void *src, size_t ldh, size_t N,
                   int NL, size_t es) {
    for (int k = 0; k < NL; ++k) {
        int rc = stage_h2d(st, s, (char *)dst + (size_t)k * ldd * es, (const char *)src + (size_t)k * ldh * es, N * es);
        if (rc) return rc;
    }
    return CLOWNS_OK;
}
int stage_d2h_cols(Stager *st, hipStream_t s, void *dst, size_t ldh, const void *src, size_t ldd, size_t N,
                   int NL, size_t es) {
    for (int k = 0; k < NL; ++k) {
        int rc = stage_d2h(st, s, (char *)dst + (size_t)k * ldh * es, (const char *)src + (size_t)k * ldd * es, N * es);
        if (rc) return rc;
    }
    return CLOWNS_OK;
}

int load_matrix(clowns_solver *h, const double *AL, int NL, int lda, const int *LL,
                const double *D, int N, double parm, hipMemcpyKind kind, bool keep_raw,
                Stager *st = nullptr) {
    int rc;
    if ((rc = check_nl(NL))) return rc;
    if (N < 1 || lda < N || !AL || !LL || !D) return CLOWNS_EBADARG;
    // The device copy gets its own 64-aligned stride (aligned vector loads).
    const int ldu = lda;
    lda = (N + 63) & ~63;
    const size_t nm = (size_t)lda * NL;
    if ((rc = h->AL.ensure(nm)) || (rc = h->LL.ensure(nm)) || (rc = h->D.ensure(N)) ||
        (rc = h->W.ensure(N)))
        return rc;
    if ((rc = ensure_common(h, N))) return rc;
    hipStream_t s = h->stream;
    if (st && kind == hipMemcpyHostToDevice) {
        if ((rc = stage_h2d_cols(st, s, h->AL.p, lda, AL, ldu, N, NL, sizeof(double))) ||
            (rc = stage_h2d_cols(st, s, h->LL.p, lda, LL, ldu, N, NL, sizeof(int))) ||
            (rc = stage_h2d(st, s, h->D.p, D, (size_t)N * sizeof(double))))
            return rc;
    } else {
        HIPCHK(hipMemcpy2DAsync(h->AL.p, (size_t)lda * sizeof(double), AL, (size_t)ldu * sizeof(double),
                                (size_t)N * sizeof(double), NL, kind, s));
        HIPCHK(hipMemcpy2DAsync(h->LL.p, (size_t)lda * sizeof(int), LL, (size_t)ldu * sizeof(int),
                                (size_t)N * sizeof(int), NL, kind, s));
        HIPCHK(hipMemcpyAsync(h->D.p, D, (size_t)N * sizeof(double), kind, s));
    }
    if (keep_raw) {
        if ((rc = h->AL0.ensure(nm)) || (rc = h->LL0.ensure(nm)) || (rc = h->D0.ensure(N)))
            return rc;
        HIPCHK(hipMemcpyAsync(h->AL0.p, h->AL.p, nm * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(h->LL0.p, h->LL.p, nm * sizeof(int), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(h->D0.p, h->D.p, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    // Reject out-of-range neighbour indices before any gather uses them.
    if ((rc = h->flag.ensure(1))) return rc;
    fill<int>(s, h->flag.p, 1, 0);
    k_check_ll<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, h->LL.p, h->flag.p);
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, h->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hbad) {
        h->N = 0;
        return CLOWNS_EBADARG;
    }
    double P1 = parm, P2;
    if (parm < 1.0) P1 = 2.0 * 2.0;
    P2 = P1 + 1.0;
    if (NL == 6)
        k_scale_sort<6><<<cdiv(N, BS), BS, 0, s>>>(N, lda, h->AL.p, h->LL.p, h->D.p, h->W.p, P1, P2);
    else
        k_scale_sort<4><<<cdiv(N, BS), BS, 0, s>>>(N, lda, h->AL.p, h->LL.p, h->D.p, h->W.p, P1, P2);
    HIPCHK(hipGetLastError());
    h->N = N;
    h->NL = NL;
    h->lda = lda;
    h->lda_user = ldu;
    h->parm = parm;
    h->have32 = false;
    h->have_t = false;
    h->raw_valid = keep_raw;
    h->will_rev_valid = false;
    h->order_G = 0;  // a caller's operator: natural order unless set_block_order
    h->order_pending = false;
    int rc2 = build_ll16(h);
    return rc2 ? rc2 : build_patterns(h);
}

int ensure_f32(clowns_solver *h) {
    if (h->have32) return CLOWNS_OK;
    int rc;
    const size_t nm = (size_t)h->lda * h->NL;
    if ((rc = h->AL32.ensure(nm)) || (rc = h->D32.ensure(h->N))) return rc;
    hipStream_t s = h->stream;
    k_to_f32<<<cdiv(h->N, BS), BS, 0, s>>>(h->N, h->lda, h->NL, h->AL.p, h->D.p, h->AL32.p, h->D32.p);
    if (h->kn.pm_half) {
        if ((rc = h->ALh.ensure(nm))) return rc;
        k_to_f16<<<cdiv(h->N, BS), BS, 0, s>>>(h->N, h->lda, h->NL, h->AL.p, h->ALh.p);
    }
    h->have_h = h->kn.pm_half;
    h->have_rows = h->have_rows8 = false;
    if (h->kn.rows32 && h->NL == 6 && h->ll16_ok) {
        const size_t n6 = (size_t)6 * h->N;
        if ((rc = h->ALr32.ensure(n6)) || (rc = h->Lr16.ensure(n6)) || (rc = h->nonunit.ensure(1))) return rc;
        if (h->ll8_ok && (rc = h->Lr8.ensure(n6))) return rc;
        fill<int>(s, h->nonunit.p, 1, 0);
        k_to_rows32<<<cdiv(h->N, BS), BS, 0, s>>>(h->N, h->lda, h->AL.p, h->LL16.p, h->ll8_ok ? h->LL8.p : nullptr,
                                                  h->D32.p, h->ALr32.p, h->Lr16.p, h->ll8_ok ? h->Lr8.p : nullptr,
                                                  h->nonunit.p);
        h->have_rows = true;
        h->have_rows8 = h->ll8_ok;
    }
    HIPCHK(hipGetLastError());
    h->have32 = true;
    return CLOWNS_OK;
}

// The 16-bit form of the resident sorted LL, used by the two-row kernels
// when every offset fits (always, for box grids up to 16383 cells a line).
int build_ll16(clowns_solver *h) {
    h->ll16_ok = false;
    h->ll16_wide = 0;
    if (!h->kn.ll16) return CLOWNS_OK;
    const int N = h->N, nblk = cdiv(N, LL16_ROWS);
    int rc;
    if ((rc = h->LL16.ensure((size_t)h->lda * h->NL)) || (rc = h->base16.ensure(nblk)) || (rc = h->flag.ensure(2)))
        return rc;
    hipStream_t s = h->stream;
    fill<int>(s, h->flag.p, 2, 0);
    k_ll16_build<<<nblk, BS, 0, s>>>(N, h->lda, h->NL, h->LL.p, h->LL16.p, h->base16.p, h->flag.p,
                                     h->flag.p + 1);
    HIPCHK(hipGetLastError());
    int res[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(res, h->flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    h->ll16_ok = res[0] == 0;
    h->ll16_wide = res[1];
    return build_ll8(h);
}

// The 8-bit form (Ix8), built beside the 16-bit one (which stays the
// fallback: CLOWNS_LL8 = 0).
int build_ll8(clowns_solver *h) {
    h->ll8_ok = false;
    h->ll8_wide = 0;
    if (h->kn.ll8 == 0 || !h->ll16_ok) return CLOWNS_OK;
    // not needed where both precisions were timed at this size and kept the
    // 16-bit codes: an assembly per time step need not rebuild them
    const bool k64 = h->ix64 == 0 && tuned_size(h->N, h->ix64_N) && h->st64_f == force_key(h);
    const bool k32 = h->ix32 == 0 && tuned_size(h->N, h->ix32_N) && h->st32_f == force_key(h);
    if (h->kn.ll8 < 0 && k64 && k32) return CLOWNS_OK;
    const int N = h->N, nblk = cdiv(N, LL8_ROWS);
    int rc;
    if ((rc = h->LL8.ensure((size_t)h->lda * h->NL)) || (rc = h->base8.ensure(nblk)) || (rc = h->flag.ensure(2)))
        return rc;
    hipStream_t s = h->stream;
    fill<int>(s, h->flag.p, 2, 0);
    k_ll8_build<<<nblk, BS, 0, s>>>(N, h->lda, h->NL, h->LL.p, h->LL8.p, h->base8.p, h->flag.p, h->flag.p + 1);
    HIPCHK(hipGetLastError());
    int res[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(res, h->flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    h->ll8_ok = res[0] == 0;
    h->ll8_wide = res[1];
    if (h->kn.verbose)
        fprintf(stderr, "clowns_pcg: N=%d 8-bit neighbour codes: %d of %d slots wide\n", N, res[1], N * h->NL);
    return CLOWNS_OK;
}

// Row patterns of the resident (scaled, sorted) operator: the table of the
// last operator is kept, and a new operator first takes its ids from it (a
// time-stepped run reassembles an operator of the same structure every
// step); a new table is built when fewer than 4 in 5 rows match: the rows
// of a 32768-row sample that repeat most often (at least 1 in 200 of the
// sample), counted on the host.  Patterns are used while at least half the
// rows are on one.
constexpr int PAT_MIN_ROWS = 1 << 18;  // below this a solve is launch-bound
constexpr int PAT_SAMPLE = 1 << 15;

int assign_patterns(clowns_solver *h, int *nnone) {
    hipStream_t s = h->stream;
    const int N = h->N;
    fill<int>(s, h->flag.p, 1, 0);
    const int npad = n_pad(N) + 4;
    k_pat_assign<<<cdiv(npad, BS), BS, 0, s>>>(N, npad, h->lda, h->AL.p, h->D.p, h->ptab.p, h->npat, h->pid.p,
                                               h->flag.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(nnone, h->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

int build_patterns(clowns_solver *h) {
    h->pat_ok = false;
    h->pat_rows = 0;
    const int N = h->N;
    if (!h->kn.patterns || h->NL != 6 || N < PAT_MIN_ROWS) return CLOWNS_OK;
    int rc;
    if ((rc = h->pid.ensure((size_t)n_pad(N) + 4)) || (rc = h->ptab.ensure(PAT_MAX * PAT_W)) ||
        (rc = h->flag.ensure(1)))
        return rc;
    hipStream_t s = h->stream;
    int nnone = N;
    if (h->npat > 0 && (rc = assign_patterns(h, &nnone))) return rc;
    if (h->npat == 0 || nnone > N / 5) {
        const int S = std::min(N, PAT_SAMPLE), stride = N / S;
        if ((rc = h->pat_keys.ensure((size_t)S * PAT_W))) return rc;
        k_pat_sample<<<cdiv(S, BS), BS, 0, s>>>(N, h->lda, h->AL.p, h->D.p, S, stride, h->pat_keys.p);
        HIPCHK(hipGetLastError());
        try {  // host containers: an allocation failure is reported, not thrown across the C-ABI
            std::vector<double> keys((size_t)S * PAT_W);
            HIPCHK(hipMemcpyAsync(keys.data(), h->pat_keys.p, keys.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            typedef std::array<uint64_t, PAT_W> Key;
            std::map<Key, int> count;
            for (int r = 0; r < S; ++r) {
                Key k;
                std::memcpy(k.data(), &keys[(size_t)r * PAT_W], sizeof(Key));
                ++count[k];
            }
            std::vector<std::pair<int, Key>> top;
            for (const auto &kv : count)
                if (kv.second >= std::max(2, S / 200)) top.push_back({kv.second, kv.first});
            // most frequent first, ties by value: the same table on every run
            std::sort(top.begin(), top.end(), [](const auto &a, const auto &b) {
                return a.first != b.first ? a.first > b.first : a.second < b.second;
            });
            h->npat = (int)std::min<size_t>(top.size(), PAT_MAX);
            std::vector<double> tab((size_t)PAT_MAX * PAT_W, 0.0);
            for (int p = 0; p < h->npat; ++p) std::memcpy(&tab[(size_t)p * PAT_W], top[p].second.data(), sizeof(Key));
            HIPCHK(hipMemcpyAsync(h->ptab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s));
            if ((rc = assign_patterns(h, &nnone))) return rc;
        } catch (const std::bad_alloc &) {
            return CLOWNS_ENOMEM;
        }
    }
    h->pat_rows = N - nnone;
    h->pat_ok = h->npat > 0 && nnone <= N / 2;
    if (h->kn.verbose)
        fprintf(stderr, "clowns_pcg: N=%d row patterns: %d in the table, %d rows on them%s\n", N, h->npat,
                h->pat_rows, h->pat_ok ? "" : " (not used)");
    return CLOWNS_OK;
}

// Views of the resident operator.
Op<double> op64(const clowns_solver *h) {
    Op<double> A{h->N, h->NL, h->lda, h->AL.p, h->LL.p, h->D.p};
    if (h->ll16_ok) {
        A.LL16 = h->LL16.p;
        A.base16 = h->base16.p;
    }
    if (h->ll8_ok) {
        A.LL8 = h->LL8.p;
        A.base8 = h->base8.p;
    }
    if (h->order_G > 0 && h->order_G == geo_of(h->N, 2).G) A.order = h->order.p;
    if (h->pat_ok) {
        A.pid = h->pid.p;
        A.ptab = h->ptab.p;
    }
    return A;
}
Op<float> op32(const clowns_solver *h) {
    // natural block order: the FP32 kernels gained nothing from a block
    // order (their vectors stay in L2 either way)
    Op<float> A{h->N, h->NL, h->lda, h->AL32.p, h->LL.p, h->D32.p};
    if (h->have_h) A.ALh = h->ALh.p;
    if (h->ll16_ok) {
        A.LL16 = h->LL16.p;
        A.base16 = h->base16.p;
    }
    if (h->ll8_ok) {
        A.LL8 = h->LL8.p;
        A.base8 = h->base8.p;
    }
    A.v4 = h->kn.v4_32;
    A.p2_32 = h->kn.spmv32_p2;
    if (h->pat_ok) {
        A.pid = h->pid.p;
        A.ptab = h->ptab.p;
    }
    if (h->have_rows) {
        // the row-major SpMV: codes of the width the FP32 kernels run with
        A.ALr = h->ALr32.p;
        A.Lr16 = h->Lr16.p;
        A.Lr8 = h->have_rows8 ? h->Lr8.p : nullptr;
        A.nonunit = h->nonunit.p;
        if (!h->have_rows8) {
            A.LL8 = nullptr;
            A.base8 = nullptr;
        }
    }
    return A;
}

// Ghost rows [N, N+64) of every vector of a run read as zero, as the
// reference's zeroed work arrays do (pcgpme_para.F90:59-75): a gather through
// a padding slot (LL = 0 -> N+1+mod(I,64), :80) then adds AL * 0 even where a
// caller's AL is not 0 there.  Rows past N keep whatever a larger earlier
// solve left, so the ghost rows are cleared whenever N changes.  VEC_SLACK
// entries more follow the ghost rows: the FP32 SpMV gathers a row pair's two
// neighbours with one 8-byte load (k_spmv_rows32), which through the last
// ghost row reads the entry after it (its value is not used: the second row
// of that pair has another neighbour and loads it on its own).
constexpr int VEC_SLACK = 4;
template <typename T> int ensure_vecs(clowns_solver *h, VecSet<T> &V, int N) {
    int rc = V.ensure((size_t)n_pad(N) + VEC_SLACK);
    if (rc) return rc;
    if (V.padN != N) {
        T *all[] = {V.r.p, V.rh.p, V.p0.p, V.q.p, V.e.p, V.v.p, V.y.p, V.t.p, V.x.p};
        for (T *b : all) fill<T>(h->stream, b + N, PAD, T(0));
        V.padN = N;
    }
    return CLOWNS_OK;
}

// The scaled right-hand side of a run (the mixed solve gathers it in its
// first PM application): N + 64 entries, ghost rows zero.
int ensure_bs(clowns_solver *h, int N) {
    int rc = h->bs.ensure((size_t)n_pad(N));
    if (rc) return rc;
    if (h->bs_padN != N) {
        fill<double>(h->stream, h->bs.p + N, PAD, 0.0);
        h->bs_padN = N;
    }
    return CLOWNS_OK;
}

// FP64 PCGPME on the resident operator `A` with row scales `Wsc`: exactly the
// reference's sequence (setup :115-145, iteration :148-247).
// beta_prod selects PCGPME_WILL's rounding of BETA (pcgpme_2DH.f90:241).
int solve_fp64(clowns_solver *h, const Op<double> &A_in, const double *Wsc, const double *d_b,
               double *d_x, double eps, int itermax, bool warm, clowns_info *info,
               bool beta_prod = false) {
    hipStream_t s = h->stream;
    int rc;
    VecSet<double> &V = h->v64;
    if ((rc = ensure_vecs(h, V, A_in.N)) || (rc = ensure_bs(h, A_in.N))) return rc;
    Op<double> A = A_in;
    A.fuse = tune_fuse(h, A, V);
    tune_stream(h, A, V);
    if (h->prof.on) A.prof = &h->prof;
    const int N = A.N;
    const int G1 = geo_of(N, 1).G;
    Scal *sc = h->sc.p;
    double *part = h->part.p;
    HIPCHK(hipEventRecord(h->ev0, s));
    k_reset<<<1, 64, 0, s>>>(sc, eps, itermax, beta_prod ? 1 : 0);
    // WK1 = W*B; X = x0.
    k_scale_b<<<cdiv(N, BS), BS, 0, s>>>(N, d_b, Wsc, V.p0.p, warm ? d_x : nullptr, V.x.p);
    // WK1 = PM*WK1 (:123).
    apply_pm<double>(s, sc, A, V.p0.p, V.p0.p, V.y.p, V.t.p, 0, nullptr, part);
    const double *wk3 = nullptr;
    if (warm) {
        // WK3 = PM*A*X (:126-127).
        spmv<double>(s, sc, A, 0, V.x.p, nullptr, V.y.p);
        apply_pm<double>(s, sc, A, V.y.p, V.q.p, V.y.p, V.t.p, 0, nullptr, part);
        wk3 = V.q.p;
    }
    k_init<double><<<G1, BS, 0, s>>>(sc, N, V.p0.p, wk3, V.r.p, V.rh.p, part);
    k_fin_init<<<1, FIN_BS, 0, s>>>(sc, part, G1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(h->ev2, s));
    if ((rc = run_bicgstab<double>(h, A, V, sc, itermax, 0, eps))) return rc;
    HIPCHK(hipMemcpyAsync(d_x, V.x.p, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    const Scal &hs = h->h_sc[0];
    if (info) {
        float ms = 0.f, ms0 = 0.f;
        (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
        (void)hipEventElapsedTime(&ms0, h->ev0, h->ev2);
        info->iters = hs.kout;
        info->ier = hs.ier;
        info->refinements = 0;
        info->err = hs.ERR;
        info->ms = ms;
        info->setup_ms = ms0;
    }
    return CLOWNS_OK;
}

// Mixed precision: FP32 inner PM-Bi-CGSTAB on (M A) d = r, FP64 refinement
// x += d with the true preconditioned residual r = M (W b - W A x) recomputed
// in FP64 after each inner solve; the outer stopping value is the
// reference's sqrt(RN/BN) evaluated on that true residual.
int solve_mixed(clowns_solver *h, const Op<double> &A_in, const Op<float> &A32_in, const double *Wsc,
                const double *d_b, double *d_x, double eps, int itermax, bool warm,
                clowns_info *info) {
    hipStream_t s = h->stream;
    int rc;
    VecSet<double> &V = h->v64;
    VecSet<float> &F = h->v32;
    if ((rc = ensure_vecs(h, V, A_in.N)) || (rc = ensure_vecs(h, F, A_in.N)) || (rc = ensure_bs(h, A_in.N)))
        return rc;
    Op<double> A = A_in;
    Op<float> A32 = A32_in;
    A.fuse = tune_fuse(h, A, V);
    A32.fuse = tune_fuse(h, A32, F);
    tune_stream(h, A, V);
    tune_stream(h, A32, F);
    if (h->prof.on) A.prof = A32.prof = &h->prof;
    const int N = A.N;
    const int G1 = geo_of(N, 1).G;
    Scal *sc = h->sc.p, *outer = h->sc.p + 1;
    double *part = h->part.p;
    HIPCHK(hipEventRecord(h->ev0, s));
    k_reset<<<1, 64, 0, s>>>(outer, eps, itermax, 0);
    k_reset<<<1, 64, 0, s>>>(sc, eps, itermax, 0);
    // bs = W*b ; BN = ||PM*bs||^2.
    k_scale_b<<<cdiv(N, BS), BS, 0, s>>>(N, d_b, Wsc, h->bs.p, warm ? d_x : nullptr, V.x.p);
    const int Gb = apply_pm<double>(s, sc, A, h->bs.p, V.r.p, V.y.p, V.t.p, 3, nullptr, part);
    k_fin_true<<<1, FIN_BS, 0, s>>>(outer, part, Gb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->h_sc + 1, outer, sizeof(Scal), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double BN = h->h_sc[1].RN;
    HIPCHK(hipEventRecord(h->ev2, s));
    double err = 1.0;
    int total_inner = 0, refinements = 0, ier = CLOWNS_NOCONV;
    if (BN == 0.0 && !warm) {
        err = 0.0;
        ier = CLOWNS_OK;
    } else {
        // outer->BN for k_fin_true; cold start: r = PM*bs, so ERR = 1.
        Scal &init = h->h_sc[1];
        init.BN = BN;
        init.ERR = std::sqrt(init.RN / BN);
        HIPCHK(hipMemcpyAsync(outer, &init, sizeof(Scal), hipMemcpyHostToDevice, s));
        if (warm) {
            // r = PM*(bs - A x).
            spmv<double>(s, sc, A, 3, V.x.p, h->bs.p, V.y.p);
            const int Gr = apply_pm<double>(s, sc, A, V.y.p, V.r.p, V.y.p, V.t.p, 3, nullptr, part);
            k_fin_true<<<1, FIN_BS, 0, s>>>(outer, part, Gr);
        }
        const double tau_min = h->tau_min;
        for (int it = 0; it < 60; ++it) {
            HIPCHK(hipMemcpyAsync(h->h_sc + 1, outer, sizeof(Scal), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            err = h->h_sc[1].ERR;
            if (!(err == err) || std::isinf(err)) {
                ier = CLOWNS_ENONFINITE;
                break;
            }
            if (err <= eps) {
                ier = CLOWNS_OK;
                break;
            }
            if (total_inner >= itermax) break;
            // Inner solve of (M A) d = r/||r|| to relative tolerance tau.
            double tau = std::max(tau_min, 0.3 * eps / err);
            k_reset<<<1, 64, 0, s>>>(sc, tau, std::max(1, itermax - total_inner), 0);
            k_rhs32<<<cdiv(N, BS), BS, 0, s>>>(N, outer, V.r.p, F.p0.p);
            fill<float>(s, F.x.p, (size_t)N, 0.f);
            k_init<float><<<G1, BS, 0, s>>>(sc, N, F.p0.p, nullptr, F.r.p, F.rh.p, part);
            k_fin_init<<<1, FIN_BS, 0, s>>>(sc, part, G1);
            if ((rc = run_bicgstab<float>(h, A32, F, sc, std::max(1, itermax - total_inner), 0, tau)))
                return rc;
            total_inner += std::max(0, std::min(h->h_sc[0].kout, itermax));
            ++refinements;
            // x += ||r|| d ; r = PM*(bs - A x) ; ERR.
            k_correct<<<cdiv(N, BS), BS, 0, s>>>(N, outer, F.x.p, V.x.p);
            k_reset<<<1, 64, 0, s>>>(sc, eps, itermax, 0);
            spmv<double>(s, sc, A, 3, V.x.p, h->bs.p, V.y.p);
            const int Gr = apply_pm<double>(s, sc, A, V.y.p, V.r.p, V.y.p, V.t.p, 3, nullptr, part);
            k_fin_true<<<1, FIN_BS, 0, s>>>(outer, part, Gr);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipMemcpyAsync(d_x, V.x.p, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info) {
        float ms = 0.f, ms0 = 0.f;
        (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
        (void)hipEventElapsedTime(&ms0, h->ev0, h->ev2);
        info->iters = total_inner;
        info->ier = ier;
        info->refinements = refinements;
        info->err = err;
        info->ms = ms;
        info->setup_ms = ms0;
    }
    return CLOWNS_OK;
}

int set_device(clowns_solver *h) {
    HIPCHK(hipSetDevice(h->device));
    return CLOWNS_OK;
}

clowns_solver *default_solver(int dev) {
    if (dev < 0 || dev >= 64) return nullptr;
    if (!g_default[dev]) {
        clowns_solver *h = nullptr;
        if (clowns_solver_create(&h, dev, nullptr) != CLOWNS_OK) return nullptr;
        g_default[dev] = h;
    }
    return g_default[dev];
}

} // namespace

// ======================================================================= C-ABI

extern "C" int clowns_solver_create(clowns_solver **out, int device, void *stream) {
    if (!out) return CLOWNS_EBADARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CLOWNS_EHIP;
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return CLOWNS_EHIP;
    }
    if (device >= ndev) return CLOWNS_EBADARG;
    clowns_solver *h = new clowns_solver();
    h->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete h;
        return CLOWNS_EHIP;
    }
    // NULL is the device's null stream (torch's default stream), so work is
    // ordered with the caller's default-stream work.
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
    {
        // getenv under a lock of its own (clowns_pcgpme creates its default
        // solver while holding g_mu); the values stay with this solver
        std::lock_guard<std::mutex> lock(g_env_mu);
        auto flag = [](const char *name, bool &v) {
            if (const char *e = getenv(name)) v = atoi(e) != 0;
        };
        auto tri = [](const char *name, int &v) {
            if (const char *e = getenv(name)) v = atoi(e) < 0 ? -1 : (atoi(e) != 0 ? 1 : 0);
        };
        Knobs &k = h->kn;
        tri("CLOWNS_FUSE_PM", k.fuse_pm);
        if (const char *e = getenv("CLOWNS_STREAM")) {
            const int v = atoi(e);
            k.stream = v < 0 ? -1 : (v == 0 ? 0 : (v == 1 ? 3 : (v == 2 ? 1 : 2)));
        }
        tri("CLOWNS_LL8", k.ll8);
        tri("CLOWNS_PM_PAT", k.pm_pat);
        flag("CLOWNS_LL16", k.ll16);
        flag("CLOWNS_V4_32", k.v4_32);
        flag("CLOWNS_SPMV_ROWS", k.rows32);
        flag("CLOWNS_SPMV32_P2", k.spmv32_p2);
        flag("CLOWNS_PM_HALF", k.pm_half);
        flag("CLOWNS_BLOCK_ORDER", k.block_order);
        flag("CLOWNS_VERBOSE", k.verbose);
        flag("CLOWNS_PATTERNS", k.patterns);
        flag("CLOWNS_CLASSIFY_HOST", h->cls_force_host);
        if (const char *e = getenv("CLOWNS_POLL")) h->poll = std::max(1, atoi(e));
        if (const char *e = getenv("CLOWNS_TAU_MIN")) h->tau_min = std::max(1e-8, atof(e));
        if (const char *e = getenv("CLOWNS_WILL_MAX_ROUNDS")) h->will_max_rounds = std::max(1, atoi(e));
    }
    *out = h;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_destroy(clowns_solver *h) {
    if (!h) return CLOWNS_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->AL.release(); h->D.release(); h->W.release(); h->LL.release(); h->LL16.release(); h->base16.release(); h->LL8.release(); h->base8.release(); h->order.release(); h->order_first.release();
    h->AL32.release(); h->D32.release(); h->ALh.release(); h->ALr32.release(); h->Lr16.release(); h->Lr8.release(); h->nonunit.release(); h->pid.release(); h->ptab.release(); h->pat_keys.release();
    h->ALt.release(); h->Dt.release(); h->Wt.release(); h->LLt.release();
    h->LLt16.release(); h->baset16.release(); h->LLt8.release(); h->baset8.release();
    h->AL0.release(); h->D0.release(); h->LL0.release();
    release_stager(h);
    h->v64.release(); h->v32.release(); h->bs.release(); h->stage_b.release(); h->stage_x.release(); h->part.release(); h->flag.release(); h->will_flags.release(); h->will_cand.release(); h->will_rev.release(); h->cls_scratch.release(); h->pc_b.release(); h->pc_x.release(); h->pc_part.release(); h->sc.release();
    if (h->h_sc) (void)hipHostFree(h->h_sc);
    if (h->ev0) { (void)hipEventDestroy(h->ev0); (void)hipEventDestroy(h->ev1); (void)hipEventDestroy(h->ev2); }
    for (hipEvent_t e : h->evp) if (e) (void)hipEventDestroy(e);
    h->prof.release();
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_set_stream(clowns_solver *h, void *stream) {
    if (!h) return CLOWNS_EBADARG;
    if (h->own_stream) (void)hipStreamDestroy(h->stream);
    h->own_stream = false;
    h->stream = (hipStream_t)stream;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_set_matrix(clowns_solver *h, const double *d_AL, int32_t NL,
                                        int32_t lda, const int32_t *d_LL, const double *d_D,
                                        int32_t N, double parm) {
    if (!h) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    rc = load_matrix(h, d_AL, NL, lda, d_LL, d_D, N, parm, hipMemcpyDeviceToDevice, true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return CLOWNS_OK;
}

extern "C" int clowns_solver_solve(clowns_solver *h, const double *d_b, double *d_x, double eps,
                                   int32_t maxiter, int32_t flags, clowns_info *info) {
    if (!h || !d_b || !d_x || h->N < 1) return CLOWNS_EBADARG;
    if (flags & ~(CLOWNS_MIXED | CLOWNS_WARM_START)) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    if (eps <= 1e-15) eps = 1e-9;                                        // :51
    if (maxiter <= 0) maxiter = (int)(10.0 * (std::sqrt((double)h->N) + 1.0)); // :52
    if (!(flags & CLOWNS_MIXED) && (rc = ensure_block_order(h))) return rc;
    Op<double> A = op64(h);
    const bool warm = (flags & CLOWNS_WARM_START) != 0;
    if (flags & CLOWNS_MIXED) {
        if ((rc = ensure_f32(h))) return rc;
        Op<float> A32 = op32(h);
        return solve_mixed(h, A, A32, h->W.p, d_b, d_x, eps, maxiter, warm, info);
    }
    return solve_fp64(h, A, h->W.p, d_b, d_x, eps, maxiter, warm, info);
}

// ----------------------------------------------------- allocation-free solve
// The work memory of clowns_solver_solve_noalloc: a file-scope device array
// (every GPU the library runs on holds its own copy until the process ends)
// for the runs' vectors (nine FP64 and nine FP32 ones of N + 64 entries and
// their slack, each 256-byte aligned), the scaled right-hand side, the dot
// partials and the run scalars, for up to NOALLOC_MAX_ROWS rows: 58 MiB.
constexpr int NOALLOC_MAX_ROWS = 1 << 19;
constexpr size_t NA_VEC = (size_t)NOALLOC_MAX_ROWS + PAD + 64;             // entries per vector
constexpr size_t NA_G1 = (((size_t)NOALLOC_MAX_ROWS + BS - 1) / BS + 7) / 8 * 8;  // largest grid
constexpr size_t NA_V64 = 0;                                                 // 9 FP64 vectors + bs
constexpr size_t NA_V32 = NA_V64 + 10 * NA_VEC * sizeof(double);             // 9 FP32 vectors
constexpr size_t NA_PART = NA_V32 + 9 * NA_VEC * sizeof(float);              // 2 x NA_G1 partials
constexpr size_t NA_SC = NA_PART + 2 * NA_G1 * sizeof(double);               // 2 run scalars
constexpr size_t NA_BYTES = NA_SC + 2 * sizeof(Scal);
static_assert(NA_VEC * sizeof(float) % 256 == 0 && NA_V32 % 256 == 0 && NA_PART % 256 == 0 && NA_SC % 256 == 0,
              "vectors start 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
std::mutex g_scratch_mu;
unsigned char *g_scratch[64] = {};  // device address of `scratch` on each GPU

// The address of `scratch` on the solver's (current) device, looked up once
// per device: in host code the symbol names no device memory.
int scratch_of(int dev, unsigned char **out) {
    if (dev < 0 || dev >= 64) return CLOWNS_EBADARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return CLOWNS_OK;
}

template <typename T> void borrow_vecs(VecSet<T> &V, unsigned char *base) {
    DevBuf<T> *all[] = {&V.r, &V.rh, &V.p0, &V.q, &V.e, &V.v, &V.y, &V.t, &V.x};
    for (int k = 0; k < 9; ++k) all[k]->borrow((T *)(base + (size_t)k * NA_VEC * sizeof(T)), NA_VEC);
    V.padN = -1;  // the array keeps the last call's contents: ghost rows cleared again
}
} // namespace

extern "C" int clowns_solver_solve_noalloc(clowns_solver *h, const double *d_b, double *d_x, double eps,
                                           int32_t maxiter, int32_t flags, clowns_info *info) {
    if (!h || !d_b || !d_x || h->N < 1) return CLOWNS_EBADARG;
    if (flags & ~(CLOWNS_MIXED | CLOWNS_WARM_START)) return CLOWNS_EBADARG;
    if (h->N > NOALLOC_MAX_ROWS) {
        fprintf(stderr, "clowns_pcg: clowns_solver_solve_noalloc: %d rows exceed its maximum of %d\n", h->N,
                NOALLOC_MAX_ROWS);
        return CLOWNS_EBADARG;
    }
    if ((flags & CLOWNS_MIXED) && !h->have32) {
        fprintf(stderr, "clowns_pcg: clowns_solver_solve_noalloc: the operator's FP32 copies are not built "
                        "(one clowns_solver_solve with CLOWNS_MIXED builds them)\n");
        return CLOWNS_EBADARG;
    }
    // the pinned status mirror and the events come with the operator
    if (!h->h_sc || !h->ev0 || !h->evp[0]) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    unsigned char *base = nullptr;
    if ((rc = scratch_of(h->device, &base))) return rc;
    if (eps <= 1e-15) eps = 1e-9;                                                // :51
    if (maxiter <= 0) maxiter = (int)(10.0 * (std::sqrt((double)h->N) + 1.0));  // :52
    // the run's work buffers point into the array for this call
    VecSet<double> v64;
    VecSet<float> v32;
    DevBuf<double> bs, part;
    DevBuf<Scal> sc;
    borrow_vecs(v64, base + NA_V64);
    bs.borrow((double *)(base + NA_V64 + 9 * NA_VEC * sizeof(double)), NA_VEC);
    borrow_vecs(v32, base + NA_V32);
    part.borrow((double *)(base + NA_PART), 2 * NA_G1);
    sc.borrow((Scal *)(base + NA_SC), 2);
    std::swap(h->v64, v64);
    std::swap(h->v32, v32);
    std::swap(h->bs, bs);
    std::swap(h->part, part);
    std::swap(h->sc, sc);
    const int bs_padN = h->bs_padN;
    const bool prof_on = h->prof.on;
    h->bs_padN = -1;
    h->prof.on = false;  // its events would be created on demand
    // the grid's block order is taken if it was built (its build allocates)
    Op<double> A = op64(h);
    if (flags & CLOWNS_MIXED)
        rc = solve_mixed(h, A, op32(h), h->W.p, d_b, d_x, eps, maxiter, (flags & CLOWNS_WARM_START) != 0, info);
    else
        rc = solve_fp64(h, A, h->W.p, d_b, d_x, eps, maxiter, (flags & CLOWNS_WARM_START) != 0, info);
    std::swap(h->v64, v64);
    std::swap(h->v32, v32);
    std::swap(h->bs, bs);
    std::swap(h->part, part);
    std::swap(h->sc, sc);
    h->bs_padN = bs_padN;
    h->prof.on = prof_on;
    return rc;
}

extern "C" int clowns_solver_get_scaled(clowns_solver *h, double *d_AL, int32_t *d_LL,
                                        double *d_D, double *d_W, double *d_B) {
    if (!h || h->N < 1) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const size_t N = h->N;
    if (d_AL)
        HIPCHK(hipMemcpy2DAsync(d_AL, (size_t)h->lda_user * sizeof(double), h->AL.p, (size_t)h->lda * sizeof(double),
                                N * sizeof(double), h->NL, hipMemcpyDeviceToDevice, s));
    if (d_LL)
        HIPCHK(hipMemcpy2DAsync(d_LL, (size_t)h->lda_user * sizeof(int), h->LL.p, (size_t)h->lda * sizeof(int),
                                N * sizeof(int), h->NL, hipMemcpyDeviceToDevice, s));
    if (d_D) HIPCHK(hipMemcpyAsync(d_D, h->D.p, (size_t)h->N * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (d_W) HIPCHK(hipMemcpyAsync(d_W, h->W.p, (size_t)h->N * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (d_B) {
        if (h->v64.r.n < (size_t)h->N) return CLOWNS_EBADARG;
        HIPCHK(hipMemcpyAsync(d_B, h->v64.r.p, (size_t)h->N * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

extern "C" void clowns_pcgpme(double *AL, const int32_t *NL_, const int32_t *NA_, int32_t *LL,
                              double *D, const int32_t *N_, const int32_t *N2_, double *B,
                              double *X, double *EPS, const double *PARM, int32_t *ITER,
                              int32_t *IER) {
    if (!IER) return;
    if (!AL || !NL_ || !NA_ || !LL || !D || !N_ || !N2_ || !B || !X || !EPS || !PARM || !ITER) {
        *IER = CLOWNS_EBADARG;
        return;
    }
    const int NL = *NL_, NA = *NA_, N = *N_, N2 = *N2_;
    if (check_nl(NL)) {
        *IER = CLOWNS_EBADNL;
        return;
    }
    if (N < 1 || NA < N || N2 < 64) {
        *IER = CLOWNS_EBADARG;
        return;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        *IER = CLOWNS_EHIP;
        return;
    }
    clowns_solver *h = default_solver(dev);
    if (!h) {
        *IER = CLOWNS_EHIP;
        return;
    }
    if (set_device(h)) {
        *IER = CLOWNS_EHIP;
        return;
    }
    double eps = EPS[0];
    int itermax = *ITER;
    if (eps <= 1e-15) eps = 1e-9;                                            // :51
    if (itermax <= 0) itermax = (int)(10.0 * (std::sqrt((double)N) + 1.0));  // :52
    // The caller normally passes X = 0 (1sub-momentum.F90:267-271); then the
    // reference's PM*A*X term is exactly zero and is skipped.
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const bool warm = any_nonzero(X, (size_t)N);
    hipStream_t s = h->stream;
    int rc = ensure_stager(h->stager, true, false);
    if (!rc) rc = ensure_stager(h->stager_out, false, true);
    Stager *st = h->stager, *so = h->stager_out;
    const auto t1 = clk::now();
    if (!rc) rc = load_matrix(h, AL, NL, NA, LL, D, N, *PARM, hipMemcpyHostToDevice, false, st);
    const auto t2 = clk::now();
    // The reference leaves AL, D scaled and LL sorted (pcgpme_para.F90:77-112),
    // final once the matrix is set up: a host thread copies them out on the
    // out-stager's stream while this one runs the solve (the solve only reads
    // them).  It is joined before any return.
    const size_t ld = (size_t)h->lda, n8 = (size_t)N * sizeof(double);
    std::thread out_thread;
    int rc_out = CLOWNS_OK;
    auto copy_out = [h, so, AL, LL, D, NA, N, NL, ld, n8]() -> int {
        int r = set_device(h);
        if (!r) r = stage_d2h_cols(so, so->cs, AL, NA, h->AL.p, ld, N, NL, sizeof(double));
        if (!r) r = stage_d2h_cols(so, so->cs, LL, NA, h->LL.p, ld, N, NL, sizeof(int));
        if (!r) r = stage_d2h(so, so->cs, D, h->D.p, n8);
        if (!r) r = stage_finish(so);
        if (!r && hipStreamSynchronize(so->cs) != hipSuccess) r = CLOWNS_EHIP;
        return r;
    };
    if (!rc) {
        if (hipEventRecord(so->ready, s) != hipSuccess || hipStreamWaitEvent(so->cs, so->ready, 0) != hipSuccess) {
            rc = CLOWNS_EHIP;
        } else {
            try {
                out_thread = std::thread([&rc_out, &copy_out] { rc_out = copy_out(); });
            } catch (const std::system_error &) {
                rc_out = copy_out();  // no thread to be had: the same copies, before the solve
            }
        }
    }
    if (!rc) rc = h->stage_b.ensure(N);
    if (!rc) rc = h->stage_x.ensure(N);
    if (!rc) rc = stage_h2d(st, s, h->stage_b.p, B, n8);
    if (!rc && warm) rc = stage_h2d(st, s, h->stage_x.p, X, n8);
    clowns_info info{};
    if (!rc) {
        Op<double> A = op64(h);
        rc = solve_fp64(h, A, h->W.p, h->stage_b.p, h->stage_x.p, eps, itermax, warm, &info);
    }
    const auto t3 = clk::now();
    if (!rc) {
        // B the final scaled residual, X the solution (:250-257)
        if ((rc = stage_d2h(st, s, B, h->v64.r.p, n8)) || (rc = stage_d2h(st, s, X, h->stage_x.p, n8)) ||
            (rc = stage_finish(st))) {
        } else if (hipStreamSynchronize(s) != hipSuccess) {
            rc = CLOWNS_EHIP;
        }
    }
    if (out_thread.joinable()) out_thread.join();
    if (!rc) rc = rc_out;
    if (h->kn.verbose) {
        const auto ms = [](clk::time_point a, clk::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        fprintf(stderr, "clowns_pcg: drop-in N=%d: X scan %.1f ms, matrix in + setup %.1f ms, B in + solve %.1f ms, "
                        "out %.1f ms\n", N, ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, clk::now()));
    }
    if (rc) {
        *IER = rc;
        return;
    }
    *ITER = info.iters;
    EPS[0] = info.err;
    *IER = info.ier;
}

// ================================================================ PCGPME_WILL

namespace {

// Device rounds of the output pass before the exact host pass takes over (a
// chain of snaps longer than this is the only way to reach it); the solver's
// will_max_rounds may lower it.
constexpr int WILL_MAX_ROUNDS = 4096;

// PCGPME_WILL's output pass in the reference's serial order
// (pcgpme_2DH.f90:255-268), for the capped case.
void will_host(int N, int lda, int NL, const std::vector<int> &LL, std::vector<double> &X,
               std::vector<double> &B) {
    for (int i = 0; i < N; ++i) {
        if (std::fabs(X[i]) <= WILL_LIMIT) X[i] = 0.0;
        for (int k = 0; k < NL; ++k) {
            const int j = LL[(size_t)k * lda + i];
            if (j == i + 1) continue;
            if (j == 0) break;
            if (X[i] == X[j - 1]) continue;
            if (std::fabs(X[i] - X[j - 1]) <= WILL_LIMIT) X[i] = X[j - 1];
            B[i] = X[i];
        }
    }
}

// X' and B of the output pass from the solution X and the final residual R
// (B may be R).  LL: caller order, 1-based, column stride lda, entries in
// [0, N] (checked by the caller).  reuse_rev: will_rev already holds this
// LL's readers.  *rounds: device rounds, or -1 when the host pass ran.
int will_filter(clowns_solver *h, const int *d_LL, int NL, int lda, int N, const double *d_X,
                const double *d_R, double *d_Xout, double *d_B, int *rounds, bool reuse_rev) {
    hipStream_t s = h->stream;
    int rc;
    const int nb = cdiv(N, WILL_SCAN);
    const size_t nrev = (size_t)2 * N + 1 + nb + 1 + (size_t)N * NL;
    if ((rc = ensure_vecs(h, h->v64, N)) || (rc = h->will_flags.ensure(WILL_MAX_ROUNDS + 1)) ||
        (rc = h->will_cand.ensure((size_t)3 * N)) || (rc = h->will_rev.ensure(nrev)))
        return rc;
    int *off = h->will_rev.p, *deg = off + N + 1, *bsum = deg + N, *readers = bsum + nb + 1;
    if (!reuse_rev) {
        fill<int>(s, deg, (size_t)N, 0);
        k_will_rdeg<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, d_LL, deg);
        k_will_rsum<<<nb, BS, 0, s>>>(N, deg, bsum);
        k_cls_scan_blocks<<<1, 1024, 0, s>>>(nb, bsum);
        k_will_roff<<<nb, BS, 0, s>>>(N, deg, bsum, off);
        fill<int>(s, deg, (size_t)N, 0);
        k_will_rfill<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, d_LL, off, deg, readers);
        HIPCHK(hipGetLastError());
    }
    double *Y = h->v64.y.p;
    int *cand[2] = {h->will_cand.p, h->will_cand.p + N}, *stamp = h->will_cand.p + 2 * (size_t)N;
    const int cap = std::max(1, std::min(WILL_MAX_ROUNDS, h->will_max_rounds));
    fill<int>(s, h->will_flags.p, WILL_MAX_ROUNDS + 1, 0);
    fill<int>(s, stamp, (size_t)N, 0);
    k_will_init<<<cdiv(N, BS), BS, 0, s>>>(N, d_X, Y);
    const int Gf = std::min(cdiv(N, BS), 1024);
    int launched = 0, batch = 2, r0 = -1;
    int hf[256];
    while (launched < cap) {
        const int nr = std::min(batch, cap - launched);
        for (int r = launched; r < launched + nr; ++r)
            k_will_front<<<r == 0 ? cdiv(N, BS) : Gf, BS, 0, s>>>(N, lda, NL, d_LL, d_X, Y, off, readers,
                                                                cand[r & 1], cand[(r + 1) & 1], stamp,
                                                                h->will_flags.p, r);
        HIPCHK(hipGetLastError());
        // will_flags[r + 1]: rows pushed by round r; none = fixed point
        HIPCHK(hipMemcpyAsync(hf, h->will_flags.p + launched + 1, (size_t)nr * sizeof(int),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int r = 0; r < nr && r0 < 0; ++r)
            if (hf[r] == 0) r0 = launched + r;
        launched += nr;
        if (r0 >= 0) break;
        batch = std::min(256, batch * 2);
    }
    if (r0 >= 0) {
        k_will_out<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, d_LL, d_X, Y, d_R, d_Xout, d_B);
        HIPCHK(hipGetLastError());
        if (rounds) *rounds = r0 + 1;
        return CLOWNS_OK;
    }
    try {  // host vectors: an allocation failure is reported, not thrown across the C-ABI
        std::vector<int> LL((size_t)lda * NL);
        std::vector<double> X(N), B(N);
        HIPCHK(hipMemcpy2DAsync(LL.data(), (size_t)lda * sizeof(int), d_LL, (size_t)lda * sizeof(int),
                                (size_t)N * sizeof(int), NL, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(X.data(), d_X, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(B.data(), d_R, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        will_host(N, lda, NL, LL, X, B);
        if (d_Xout)
            HIPCHK(hipMemcpyAsync(d_Xout, X.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_B, B.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    } catch (const std::bad_alloc &) {
        return CLOWNS_ENOMEM;
    }
    if (rounds) *rounds = -1;
    return CLOWNS_OK;
}

// PCGPME_WILL on the resident operator: zero start, the reference's EPS and
// ITER as given (no defaults), BETA as C1/(AMU*RQ), then the output pass over
// the caller-order LL kept by load_matrix.  d_out receives B.
int solve_will(clowns_solver *h, const double *d_b, double *d_out, double *d_xout, double eps,
               int iter, clowns_info *info, int *rounds) {
    if (!h->raw_valid) return CLOWNS_EBADARG;
    int rc;
    const int N = h->N;
    if ((rc = h->stage_x.ensure(N))) return rc;
    Op<double> A = op64(h);
    clowns_info inf{};
    if ((rc = solve_fp64(h, A, h->W.p, d_b, h->stage_x.p, eps, iter, false, &inf, true))) return rc;
    // No iteration ran: the reference falls through to IER = 3000
    // (pcgpme_2DH.f90:248-249); a zero right-hand side returns at once.
    if (iter <= 0 && !(h->h_sc[0].done && h->h_sc[0].BN == 0.0)) inf.ier = CLOWNS_NOCONV;
    if ((rc = will_filter(h, h->LL0.p, h->NL, h->lda, N, h->stage_x.p, h->v64.r.p, d_xout, d_out,
                          rounds, h->will_rev_valid)))
        return rc;
    h->will_rev_valid = true;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (info) *info = inf;
    return CLOWNS_OK;
}

} // namespace

extern "C" void clowns_pcgpme_will(const double *AL, const int32_t *NL_, const int32_t *NA_,
                                   const int32_t *LL, const double *D, const int32_t *N_,
                                   const int32_t *N2_, double *B, const double *EPS,
                                   const double *PARM, const int32_t *ITER, int32_t *IER) {
    if (!IER) return;
    if (!AL || !NL_ || !NA_ || !LL || !D || !N_ || !N2_ || !B || !EPS || !PARM || !ITER) {
        *IER = CLOWNS_EBADARG;
        return;
    }
    const int NL = *NL_, NA = *NA_, N = *N_, N2 = *N2_;
    if (check_nl(NL)) {
        *IER = CLOWNS_EBADNL;
        return;
    }
    if (N < 1 || NA < N || N2 < 64) {
        *IER = CLOWNS_EBADARG;
        return;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        *IER = CLOWNS_EHIP;
        return;
    }
    clowns_solver *h = default_solver(dev);
    if (!h || set_device(h)) {
        *IER = CLOWNS_EHIP;
        return;
    }
    hipStream_t s = h->stream;
    // AL, LL and D are inputs only (the reference scales copies,
    // pcgpme_2DH.f90:56-76); keep_raw holds the caller-order LL.
    int rc = ensure_stager(h->stager, true, false);
    Stager *st = h->stager;
    if (!rc) rc = load_matrix(h, AL, NL, NA, LL, D, N, *PARM, hipMemcpyHostToDevice, true, st);
    if (!rc) rc = h->stage_b.ensure(N);
    if (!rc) rc = stage_h2d(st, s, h->stage_b.p, B, (size_t)N * sizeof(double));
    clowns_info info{};
    if (!rc) rc = solve_will(h, h->stage_b.p, h->stage_b.p, nullptr, *EPS, *ITER, &info, nullptr);
    if (!rc && !(rc = stage_d2h(st, s, B, h->stage_b.p, (size_t)N * sizeof(double))) &&
        !(rc = stage_finish(st)) && hipStreamSynchronize(s) != hipSuccess)
        rc = CLOWNS_EHIP;
    *IER = rc ? rc : info.ier;
}

extern "C" int clowns_solver_solve_will(clowns_solver *h, const double *d_b, double *d_out,
                                        double eps, int32_t iter, clowns_info *info,
                                        int32_t *filter_rounds) {
    if (!h || !d_b || !d_out || h->N < 1) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    return solve_will(h, d_b, d_out, nullptr, eps, iter, info, filter_rounds);
}

extern "C" int clowns_will_filter(clowns_solver *h, const int32_t *d_LL, int32_t NL, int32_t lda,
                                  int32_t N, const double *d_X, const double *d_R, double *d_Xout,
                                  double *d_B, int32_t *rounds) {
    if (!h || !d_LL || !d_X || !d_R || !d_B || N < 1 || lda < N || NL < 1 || NL > 8)
        return CLOWNS_EBADARG;
    if (d_Xout == d_X) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    if ((rc = h->flag.ensure(1))) return rc;
    fill<int>(s, h->flag.p, 1, 0);
    k_check_ll<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, d_LL, h->flag.p);
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, h->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hbad) return CLOWNS_EBADARG;
    if ((rc = will_filter(h, d_LL, NL, lda, N, d_X, d_R, d_Xout, d_B, rounds, false))) return rc;
    h->will_rev_valid = false;  // will_rev now holds the caller's LL
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

extern "C" int clowns_solver_kernel_profile(clowns_solver *h, int32_t enable, double *ms, int32_t *count) {
    if (!h) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    KProf &p = h->prof;
    if (enable) {
        p.rec.clear();
        p.kout.clear();
        p.run = -1;
        p.K = 0;
        p.on = true;
        return CLOWNS_OK;
    }
    if (!p.on) return CLOWNS_EBADARG;
    p.on = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < 2 * KC_N; ++c) {
        if (ms) ms[c] = 0.0;
        if (count) count[c] = 0;
    }
    for (size_t i = 0; i < p.rec.size(); ++i) {
        const KProf::Rec &r = p.rec[i];
        if (r.cls < 0) continue;
        if (r.run >= 0 && r.run < (int)p.kout.size() && r.K > p.kout[r.run]) continue;  // returned at once
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, p.pool[2 * i], p.pool[2 * i + 1]));
        if (ms) ms[r.cls] += e;
        if (count) count[r.cls] += 1;
    }
    p.rec.clear();
    p.kout.clear();
    return CLOWNS_OK;
}

extern "C" int clowns_solver_time_kernel(clowns_solver *h, int32_t kernel, int32_t reps,
                                         double *avg_ms) {
    if (!h || h->N < 1 || reps < 1 || !avg_ms) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const int N = h->N;
    if ((rc = ensure_vecs(h, h->v64, N))) return rc;
    if (kernel == 2 || kernel == 3 || kernel == 6) {
        if ((rc = ensure_f32(h)) || (rc = ensure_vecs(h, h->v32, N))) return rc;
    }
    Scal *sc = h->sc.p;
    k_reset<<<1, 64, 0, s>>>(sc, 0.0, 1, 0);
    if ((rc = ensure_block_order(h))) return rc;
    Op<double> A = op64(h);
    Op<float> A32 = op32(h);
    known_stream(h, A);
    known_stream(h, A32);
    auto one = [&]() {
        switch (kernel) {
        case 0: spmv<double>(s, sc, A, 0, h->v64.p0.p, nullptr, h->v64.y.p); break;
        case 1: launch_pm_stage<double, 1>(s, sc, A, 0, h->v64.y.p, h->v64.t.p, 0, nullptr, h->part.p); break;
        case 2: spmv<float>(s, sc, A32, 0, h->v32.p0.p, nullptr, h->v32.y.p); break;
        case 3: launch_pm_stage<float, 1>(s, sc, A32, 0, h->v32.y.p, h->v32.t.p, 0, nullptr, h->part.p); break;
        case 4: k_copy<double><<<cdiv(N, BS), BS, 0, s>>>(N, h->v64.r.p, h->v64.e.p); break;
        case 5: launch_pm2<double>(s, sc, A, 0, h->v64.y.p, h->v64.t.p); break;
        default: launch_pm2<float>(s, sc, A32, 0, h->v32.y.p, h->v32.t.p); break;
        }
    };
    one();
    HIPCHK(hipEventRecord(h->ev0, s));
    for (int k = 0; k < reps; ++k) one();
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *avg_ms = (double)ms / reps;
    return CLOWNS_OK;
}

namespace {
// The 16- and 8-bit codes of the transposed operator's sorted LL, as
// build_ll16 / build_ll8 make them for the forward one.
int build_codes_t(clowns_solver *h) {
    h->t16_ok = h->t8_ok = false;
    if (!h->kn.ll16) return CLOWNS_OK;
    const int N = h->N;
    const size_t nm = (size_t)h->lda * h->NL;
    int rc;
    if ((rc = h->LLt16.ensure(nm)) || (rc = h->baset16.ensure(cdiv(N, LL16_ROWS))) || (rc = h->flag.ensure(4)))
        return rc;
    const bool want8 = h->kn.ll8 != 0;
    if (want8 && ((rc = h->LLt8.ensure(nm)) || (rc = h->baset8.ensure(cdiv(N, LL8_ROWS))))) return rc;
    hipStream_t s = h->stream;
    fill<int>(s, h->flag.p, 4, 0);
    k_ll16_build<<<cdiv(N, LL16_ROWS), BS, 0, s>>>(N, h->lda, h->NL, h->LLt.p, h->LLt16.p, h->baset16.p, h->flag.p,
                                                   h->flag.p + 1);
    if (want8)
        k_ll8_build<<<cdiv(N, LL8_ROWS), BS, 0, s>>>(N, h->lda, h->NL, h->LLt.p, h->LLt8.p, h->baset8.p,
                                                     h->flag.p + 2, h->flag.p + 3);
    HIPCHK(hipGetLastError());
    int res[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(res, h->flag.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    h->t16_ok = res[0] == 0;
    h->t8_ok = want8 && h->t16_ok && res[2] == 0;
    return CLOWNS_OK;
}

// Build and scale/sort A^T from the caller-layout copy kept by set_matrix.
int ensure_transpose(clowns_solver *h) {
    if (h->have_t) return CLOWNS_OK;
    if (!h->raw_valid) return CLOWNS_EBADARG;
    int rc;
    const int N = h->N, NL = h->NL, lda = h->lda;
    const size_t nm = (size_t)lda * NL;
    if ((rc = h->ALt.ensure(nm)) || (rc = h->LLt.ensure(nm)) || (rc = h->Dt.ensure(N)) ||
        (rc = h->Wt.ensure(N)) || (rc = h->flag.ensure(1)))
        return rc;
    hipStream_t s = h->stream;
    int *bad = h->flag.p;
    fill<int>(s, bad, 1, 0);
    k_transpose<<<cdiv(N, BS), BS, 0, s>>>(N, lda, NL, h->AL0.p, h->LL0.p, h->D0.p, h->ALt.p,
                                          h->LLt.p, h->Dt.p, bad);
    int hbad = 0;
    HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hbad) return CLOWNS_ESTRUCT;
    double P1 = h->parm, P2;
    if (h->parm < 1.0) P1 = 2.0 * 2.0;
    P2 = P1 + 1.0;
    if (NL == 6)
        k_scale_sort<6><<<cdiv(N, BS), BS, 0, s>>>(N, lda, h->ALt.p, h->LLt.p, h->Dt.p, h->Wt.p, P1, P2);
    else
        k_scale_sort<4><<<cdiv(N, BS), BS, 0, s>>>(N, lda, h->ALt.p, h->LLt.p, h->Dt.p, h->Wt.p, P1, P2);
    HIPCHK(hipGetLastError());
    if ((rc = build_codes_t(h))) return rc;
    h->have_t = true;
    return CLOWNS_OK;
}
} // namespace

extern "C" int clowns_solver_solve_transpose(clowns_solver *h, const double *d_g,
                                             double *d_lambda, double eps, int32_t maxiter,
                                             int32_t flags, clowns_info *info) {
    if (!h || !d_g || !d_lambda || h->N < 1) return CLOWNS_EBADARG;
    if (flags & ~CLOWNS_WARM_START) return CLOWNS_EBADARG;  // FP64 only
    int rc = set_device(h);
    if (rc) return rc;
    if ((rc = ensure_transpose(h))) return rc;
    if (eps <= 1e-15) eps = 1e-9;
    if (maxiter <= 0) maxiter = (int)(10.0 * (std::sqrt((double)h->N) + 1.0));
    if ((rc = ensure_block_order(h))) return rc;
    Op<double> At{h->N, h->NL, h->lda, h->ALt.p, h->LLt.p, h->Dt.p};
    if (h->t16_ok) {
        At.LL16 = h->LLt16.p;
        At.base16 = h->baset16.p;
    }
    if (h->t8_ok) {
        At.LL8 = h->LLt8.p;
        At.base8 = h->baset8.p;
    }
    // the transpose has the forward operator's rows: the same block order
    if (h->order_G > 0 && h->order_G == geo_of(h->N, 2).G) At.order = h->order.p;
    return solve_fp64(h, At, h->Wt.p, d_g, d_lambda, eps, maxiter,
                      (flags & CLOWNS_WARM_START) != 0, info);
}

extern "C" int clowns_grad_coefficients(clowns_solver *h, const double *d_lambda,
                                        const double *d_x, double *d_gAL, double *d_gD) {
    if (!h || !d_lambda || !d_x || h->N < 1) return CLOWNS_EBADARG;
    if (!h->raw_valid) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    k_grad_coef<<<cdiv(h->N, BS), BS, 0, h->stream>>>(h->N, h->lda, h->lda_user, h->NL, h->LL0.p,
                                                      d_lambda, d_x, d_gAL, d_gD);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return CLOWNS_OK;
}

namespace {
GridDev grid_of(const clowns_grid_dev *g) {
    GridDev d;
    d.nx = g->nx;
    d.ny = g->ny;
    d.nz = g->nz;
    d.stride = g->stride;
    d.N = g->nrows;
    d.ncell = (long)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
    d.dt = g->dt;
    d.mp = g->mp;
    d.nf = g->nf;
    d.nb = g->nb;
    d.mno = g->mno;
    d.row2cell = g->row2cell;
    d.sp_row = g->sp_row;
    d.F = g->F;
    d.fb = g->fb;
    d.a = g->a;
    d.un = g->un;
    d.dx = g->dx;
    d.xf = g->xf;
    return d;
}
bool grid_ok(const clowns_grid_dev *g) {
    return g && g->nrows >= 1 && g->nx >= 1 && g->ny >= 1 && g->nz >= 1 && g->nf && g->nb && g->mno &&
           g->row2cell && g->sp_row && g->F && g->fb && g->a && g->un && g->dx && g->xf && g->dt != 0.0 &&
           g->stride >= std::max(g->nx, std::max(g->ny, g->nz)) + 2;
}
} // namespace

extern "C" int clowns_assemble(const clowns_grid_dev *g, int32_t lda, double *d_AL, int32_t *d_LL,
                               double *d_D, double *d_B, void *stream) {
    if (!grid_ok(g) || lda < g->nrows || !d_AL || !d_LL || !d_D || !d_B) return CLOWNS_EBADARG;
    hipStream_t s = (hipStream_t)stream;
    k_assemble<false><<<cdiv(g->nrows, 256), 256, 0, s>>>(grid_of(g), lda, d_AL, d_LL, d_D, d_B, nullptr,
                                                          nullptr, nullptr, nullptr, 0.0, 0.0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

extern "C" int clowns_grad_open_ratios(clowns_solver *h, const clowns_grid_dev *g, const double *d_lambda,
                                       const double *d_x, double *d_ga, double *d_gfb) {
    if (!h || !grid_ok(g) || !d_lambda || !d_x || (!d_ga && !d_gfb)) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    const GridDev G = grid_of(g);
    hipStream_t s = h->stream;
    if (d_ga) k_grad_open_faces<<<cdiv(3 * G.ncell, 256), 256, 0, s>>>(G, d_lambda, d_x, d_ga);
    if (d_gfb) k_grad_open_cells<<<cdiv(G.ncell, 256), 256, 0, s>>>(G, d_lambda, d_x, d_gfb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

namespace {
__global__ void __launch_bounds__(BS)
k_block_first(int nb, const int *__restrict__ row2cell, int *__restrict__ out) {
    const int b = blockIdx.x * BS + threadIdx.x;
    if (b < nb) out[b] = row2cell[(size_t)b * (2 * BS)];
}

// Assembly: record the first cell of each row block (the block's place in
// the box) for the block order; no host work until an FP64 solve needs it.
int note_block_order(clowns_solver *h, const clowns_grid_dev *g) {
    h->order_G = 0;
    h->order_pending = false;
    // below ~1M rows a plane's worth of rows stays in L2 anyway, and building
    // the order (two host syncs) would only add latency to a short solve
    if (!h->kn.block_order || h->N < ORDER_MIN_ROWS) return CLOWNS_OK;
    const int nb = cdiv(h->N, 2 * BS);
    int rc;
    if ((rc = h->order_first.ensure(nb))) return rc;
    k_block_first<<<cdiv(nb, BS), BS, 0, h->stream>>>(nb, g->row2cell, h->order_first.p);
    HIPCHK(hipGetLastError());
    h->order_X = g->nx + 2;
    h->order_Y = g->ny + 2;
    h->order_Z = g->nz + 2;
    h->order_pending = true;
    return CLOWNS_OK;
}

int ensure_block_order(clowns_solver *h) {
    if (!h->order_pending) return CLOWNS_OK;
    h->order_pending = false;
    const int N = h->N, nb = cdiv(N, 2 * BS), G = geo_of(N, 2).G;
    int rc;
    if ((rc = h->order.ensure(G))) return rc;
    hipStream_t s = h->stream;
    std::vector<int> first(nb);
    HIPCHK(hipMemcpyAsync(first.data(), h->order_first.p, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // stable bucket sort of the blocks by (line chunk, plane)
    const long X = h->order_X, Y = h->order_Y, Z = h->order_Z;
    const long nchunk = (Y + ORDER_LINES - 1) / ORDER_LINES, nbucket = nchunk * Z;
    std::vector<int> bucket(nb), start(nbucket + 1, 0), ord(G);
    for (int b = 0; b < nb; ++b) {
        const long c = first[b];
        long u = (c >= 0 && c < X * Y * Z) ? ((c / X) % Y) / ORDER_LINES * Z + c / (X * Y) : 0;
        bucket[b] = (int)u;
        ++start[u + 1];
    }
    for (long u = 0; u < nbucket; ++u) start[u + 1] += start[u];
    for (int b = 0; b < nb; ++b) ord[start[bucket[b]]++] = b;
    for (int b = nb; b < G; ++b) ord[b] = b;
    HIPCHK(hipMemcpyAsync(h->order.p, ord.data(), (size_t)G * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    h->order_G = G;
    return CLOWNS_OK;
}
} // namespace

extern "C" int clowns_solver_assemble(clowns_solver *h, const clowns_grid_dev *g, double parm,
                                      double *d_B, int32_t keep_raw) {
    if (!h || !grid_ok(g) || !d_B) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    const int N = g->nrows;
    const int lda = (N + 63) & ~63;
    const size_t nm = (size_t)lda * 6;
    if ((rc = h->AL.ensure(nm)) || (rc = h->LL.ensure(nm)) || (rc = h->D.ensure(N)) ||
        (rc = h->W.ensure(N)))
        return rc;
    if (keep_raw && ((rc = h->AL0.ensure(nm)) || (rc = h->LL0.ensure(nm)) || (rc = h->D0.ensure(N))))
        return rc;
    if ((rc = ensure_common(h, N))) return rc;
    double P1 = parm, P2;
    if (parm < 1.0) P1 = 2.0 * 2.0;
    P2 = P1 + 1.0;
    hipStream_t s = h->stream;
    k_assemble<true><<<cdiv(N, 256), 256, 0, s>>>(grid_of(g), lda, h->AL.p, h->LL.p, h->D.p, d_B, h->W.p,
                                                 keep_raw ? h->AL0.p : nullptr, keep_raw ? h->LL0.p : nullptr,
                                                 keep_raw ? h->D0.p : nullptr, P1, P2);
    HIPCHK(hipGetLastError());
    h->N = N;
    h->NL = 6;
    h->lda = lda;
    h->lda_user = lda;
    h->parm = parm;
    h->have32 = false;
    h->have_t = false;
    h->raw_valid = keep_raw != 0;
    h->will_rev_valid = false;
    if ((rc = build_ll16(h))) return rc;
    if ((rc = build_patterns(h))) return rc;
    if ((rc = note_block_order(h, g))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    return CLOWNS_OK;
}

namespace {
__global__ void __launch_bounds__(256)
k_cls_ring(int nx, int ny, int nz, const int8_t *__restrict__ solid, int *__restrict__ bad) {
    const long X = nx + 2, Y = ny + 2, Z = nz + 2;
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= X * Y * Z) return;
    const long i = c % X, j = (c / X) % Y, k = c / (X * Y);
    if ((i == 0 || j == 0 || k == 0 || i == X - 1 || j == Y - 1 || k == Z - 1) && !solid[c]) atomicOr(bad, 1);
}

template <typename T> int to_host(std::vector<T> &v, const T *d, size_t n, hipStream_t s) {
    v.resize(n);
    HIPCHK(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return CLOWNS_OK;
}
template <typename T> int to_dev(T *d, const std::vector<T> &v, size_t n, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(d, v.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
    return CLOWNS_OK;
}
} // namespace

extern "C" int clowns_solver_classify(clowns_solver *h, const clowns_grid_desc *g, const int8_t *d_solid,
                                      double *d_F, const double *d_fb, const double *d_a, const double *d_dx,
                                      const double *d_xf, int8_t *d_nf, int8_t *d_nb, int32_t *d_mno,
                                      int32_t *d_row2cell, double *d_sp_row, int32_t cap, int32_t *nrows,
                                      double *mp, int32_t *host_fallback) {
    if (!h || !g || !d_solid || !d_F || !d_fb || !d_a || !d_dx || !d_xf || !d_nf || !d_nb || !d_mno ||
        !d_row2cell || !d_sp_row || !nrows || !mp)
        return CLOWNS_EBADARG;
    if (g->nx < 1 || g->ny < 1 || g->nz < 1 || g->stride < std::max(g->nx, std::max(g->ny, g->nz)) + 2)
        return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    ClsGrid G;
    G.nx = g->nx;
    G.ny = g->ny;
    G.nz = g->nz;
    G.X = g->nx + 2;
    G.Y = g->ny + 2;
    G.stride = g->stride;
    G.ncell = (long)G.X * G.Y * (g->nz + 2);
    G.solid = d_solid;
    G.F = d_F;
    G.fb = d_fb;
    G.a = d_a;
    G.dx = d_dx;
    G.xf = d_xf;
    G.nf = d_nf;
    G.nb = d_nb;
    const int nblk = cdiv(G.ncell, 256);
    const int nscan = cdiv(G.ncell, (long)SCAN_BS * SCAN_PER);
    if ((rc = h->cls_scratch.ensure(8 + (size_t)nscan + 1))) return rc;
    int *ext = h->cls_scratch.p, *changed = ext + 2, *bcount = ext + 8;
    int hv[5];
    fill<int>(s, ext, 8, 0);
    k_cls_ring<<<nblk, 256, 0, s>>>(g->nx, g->ny, g->nz, d_solid, changed);
    HIPCHK(hipMemcpyAsync(hv, changed, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hv[0]) return CLOWNS_EBADARG;  // the outer ring must be solid
    k_cls_init<<<nblk, 256, 0, s>>>(G, ext);
    HIPCHK(hipMemcpyAsync(hv, ext, 5 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int imax = std::min(g->nx, hv[0] + 1), kmax = std::min(g->nz, hv[1] + 1);
    // A negative F is clipped during the serial first pass (nfd2015.F90:45),
    // which later passes read: the exact host path handles that input.
    bool fallback = hv[4] != 0;
    if (!fallback) k_cls_clip<<<nblk, 256, 0, s>>>(G.ncell, d_nf, d_F);
    for (int pass = 0; !fallback; ++pass) {
        k_cls_surface<<<nblk, 256, 0, s>>>(G, imax, kmax);
        fill<int>(s, changed, 1, 0);
        k_cls_unsupported<<<nblk, 256, 0, s>>>(G, changed);
        HIPCHK(hipMemcpyAsync(hv, changed, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (hv[0] && pass < 10000) continue;  // nfd2015.F90:128-131: restart
        fill<int>(s, changed, 1, 0);
        k_cls_onesided<<<nblk, 256, 0, s>>>(G, changed);
        k_cls_orient<<<nblk, 256, 0, s>>>(G, changed + 1);
        HIPCHK(hipMemcpyAsync(hv, changed, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        fallback = hv[0] || hv[1] || h->cls_force_host;
        break;
    }
    try {  // host vectors: an allocation failure is reported, not thrown across the C-ABI
    // dx/xf are tiny: mp = (smallest cell size)^2 on the host
    std::vector<double> hdx;
    if ((rc = to_host(hdx, d_dx, (size_t)3 * g->stride, s))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    double hmin = 1e300;
    const int n3[3] = {g->nx, g->ny, g->nz};
    for (int l = 0; l < 3; ++l)
        for (int i = 1; i <= n3[l]; ++i) {
            const double v = hdx[(size_t)l * g->stride + i];
            if (v > 0.0 && v < hmin) hmin = v;
        }
    *mp = hmin * hmin;
    if (host_fallback) *host_fallback = fallback ? 1 : 0;
    if (fallback) {
        // an order-dependent sweep changes flags: the exact serial pass on the host
        const size_t nc = (size_t)G.ncell;
        std::vector<int8_t> sol, nf(nc), nb(nc);
        std::vector<double> F, fb, a, xf, sp(cap);
        std::vector<int32_t> mno(nc), r2c(cap);
        if ((rc = to_host(sol, d_solid, nc, s)) || (rc = to_host(F, (const double *)d_F, nc, s)) ||
            (rc = to_host(fb, d_fb, nc, s)) || (rc = to_host(a, d_a, 3 * nc, s)) ||
            (rc = to_host(xf, d_xf, (size_t)3 * g->stride, s)))
            return rc;
        HIPCHK(hipStreamSynchronize(s));
        int32_t N = 0;
        double mph = 0;
        rc = clowns_classify(g, sol.data(), F.data(), fb.data(), a.data(), hdx.data(), xf.data(), nf.data(),
                             nb.data(), mno.data(), r2c.data(), sp.data(), cap, &N, &mph);
        if (rc) return rc;
        if ((rc = to_dev(d_F, F, nc, s)) || (rc = to_dev(d_nf, nf, nc, s)) || (rc = to_dev(d_nb, nb, nc, s)) ||
            (rc = to_dev(d_mno, mno, nc, s)) || (rc = to_dev(d_row2cell, r2c, (size_t)N, s)) ||
            (rc = to_dev(d_sp_row, sp, (size_t)N, s)))
            return rc;
        HIPCHK(hipStreamSynchronize(s));
        *nrows = N;
        return CLOWNS_OK;
    }
    } catch (const std::bad_alloc &) {
        return CLOWNS_ENOMEM;
    }
    k_cls_count<<<nscan, SCAN_BS, 0, s>>>(G, imax, kmax, bcount);
    k_cls_scan_blocks<<<1, 1024, 0, s>>>(nscan, bcount);
    k_cls_number<<<nscan, SCAN_BS, 0, s>>>(G, imax, kmax, bcount, cap, d_mno, d_row2cell, d_sp_row);
    HIPCHK(hipGetLastError());
    int N = 0;
    HIPCHK(hipMemcpyAsync(&N, bcount + nscan, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (N > cap) return CLOWNS_EBADARG;
    *nrows = N;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_pressure_correction(clowns_solver *h, const clowns_grid_dev *g, double *d_un,
                                                 double *d_p, double d0, int32_t max_passes, double eps,
                                                 int32_t flags, clowns_pc_info *info) {
    if (!h || !grid_ok(g) || !d_un || !d_p || max_passes < 1) return CLOWNS_EBADARG;
    if (flags & ~CLOWNS_MIXED) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    clowns_grid_dev gg = *g;
    gg.un = d_un;
    GridDev G = grid_of(&gg);
    const int N = g->nrows;
    // the operator depends only on the flags, F and sp: assembled (fused with
    // scaling and sort) once and resident for every pass
    if ((rc = h->pc_b.ensure(N)) || (rc = h->pc_x.ensure(N))) return rc;
    if ((rc = clowns_solver_assemble(h, &gg, 5.0, h->pc_b.p, 0))) return rc;
    const int G1 = cdiv(N, 256);
    if ((rc = h->pc_part.ensure((size_t)G1 + 1))) return rc;
    double *dmax_dev = h->pc_part.p + G1;
    clowns_pc_info out{};
    HIPCHK(hipEventRecord(h->ev0, s));
    int ik = 0;
    for (;;) {
        // cal_err_of_continuity: RHS and DMAX from the current velocities
        k_rhs<<<G1, 256, 0, s>>>(G, h->pc_b.p, h->pc_part.p);
        k_max_fin<<<1, 1024, 0, s>>>(h->pc_part.p, G1, dmax_dev);
        double dmax = 0.0;
        HIPCHK(hipMemcpyAsync(&dmax, dmax_dev, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (ik == 0) out.dmax_first = dmax;
        out.dmax_last = dmax;
        if (dmax < d0) break;        // check_dmax (1sub-momentum.F90:437-441)
        if (++ik > max_passes) {     // (1sub-momentum.F90:448-454)
            out.ier = CLOWNS_NOCONV;
            break;
        }
        clowns_info si{};
        rc = clowns_solver_solve(h, h->pc_b.p, h->pc_x.p, eps, 10000, flags & CLOWNS_MIXED, &si);
        if (rc) return rc;
        out.iters += si.iters;
        if (si.ier != 0 && si.ier != CLOWNS_NOCONV) {
            out.ier = si.ier;
            break;
        }
        k_pressure_update<<<G1, 256, 0, s>>>(N, g->row2cell, h->pc_x.p, d_p);
        k_velocity_correction<<<cdiv(G.ncell, 256), 256, 0, s>>>(G, h->pc_x.p, d_un);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    out.passes = ik;
    out.ms = ms;
    if (info) *info = out;
    return CLOWNS_OK;
}

namespace {
bool vof_grid_ok(const clowns_grid_dev *g) {
    return g && g->nx >= 1 && g->ny >= 1 && g->nz >= 1 && g->nf && g->nb && g->F && g->fb && g->a && g->un &&
           g->dx && g->dt > 0.0 && g->stride >= std::max(g->nx, std::max(g->ny, g->nz)) + 2;
}
__global__ void k_vof_cnt_zero(int *__restrict__ cnt, int i0, int n) {
    if ((int)threadIdx.x < n) cnt[i0 + threadIdx.x] = 0;
}

// Buffers, the device view and the flag check shared by the VOF entries;
// on success the counters are zeroed and every flag the kernels rely on
// for their neighbour reads has been verified on the device.
int vof_setup(clowns_solver *h, const clowns_grid_dev *g, const double *d_Fp, double *d_qf, double *d_Fnew,
              VofDev &v, int n_plic = 0, const int32_t *d_plic_cells = nullptr) {
    if (!h || !vof_grid_ok(g) || !d_Fp || !d_qf || !d_Fnew) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const long nc = (long)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
    if ((rc = h->vof_state.ensure(nc)) || (rc = h->vof_term.ensure(nc)) || (rc = h->vof_cnt.ensure(VC_N)) ||
        (rc = ensure_events(h)))
        return rc;
    v.nx = g->nx;
    v.ny = g->ny;
    v.nz = g->nz;
    v.stride = g->stride;
    v.ncell = nc;
    v.nf = g->nf;
    v.nb = g->nb;
    v.F = g->F;
    v.Fp = d_Fp;
    v.fb = g->fb;
    v.a = g->a;
    v.un = g->un;
    v.dx = g->dx;
    v.qf = d_qf;
    v.Fn = d_Fnew;
    v.D = h->vof_term.p;  // quvwf_modify's terms are applied before k_vof_newf writes D
    int *cnt = h->vof_cnt.p;
    k_vof_cnt_zero<<<1, 64, 0, s>>>(cnt, 0, VC_N);
    k_vof_check<<<cdiv(nc, 256), 256, 0, s>>>(v, cnt);
    if (n_plic > 0) k_vof_check_plic<<<cdiv(n_plic, 256), 256, 0, s>>>(v, n_plic, d_plic_cells, cnt);
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, cnt + VC_BAD, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return bad ? CLOWNS_EBADARG : CLOWNS_OK;
}

// One cal_vof: mode 0 (PLIC face fluxes given per face, or the PLIC
// volumes of n_plic listed cells) or 1 (quvwf_kaidan).
int vof_run(clowns_solver *h, const clowns_grid_dev *g, const double *d_Fp, int32_t mode, const double *d_qs,
            const int8_t *d_qs_mask, int n_plic, const int32_t *d_plic_cells, const double *d_plic_vol,
            double *d_qf, double *d_Fnew, clowns_vof_info *info) {
    VofDev v;
    int rc = vof_setup(h, g, d_Fp, d_qf, d_Fnew, v, n_plic, d_plic_cells);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const long nc = v.ncell;
    int *cnt = h->vof_cnt.p;
    double dt = g->dt;
    const int G = cdiv(nc, 256), G8 = (G + 7) & ~7, Gc8 = (cdiv((long)g->nx * g->ny, 256) + 7) & ~7;
    clowns_vof_info out{};
    HIPCHK(hipEventRecord(h->ev0, s));
    // cal_quvwf with the surface stage
    if (mode == 1)
        k_vof_flux<1><<<G8, 256, 0, s>>>(v, dt, nullptr, nullptr);
    else
        k_vof_flux<0><<<G8, 256, 0, s>>>(v, dt, d_qs, d_qs_mask);
    if (n_plic > 0) k_vof_plic<<<cdiv(n_plic, 256), 256, 0, s>>>(v, dt, n_plic, d_plic_cells, d_plic_vol);
    // quvwf_modify: rounds until every surface cell is decided (each round
    // decides at least the first pending cell of the serial order)
    k_vof_state_init<<<G, 256, 0, s>>>(nc, h->vof_state.p);
    for (int round = 1;; ++round) {
        k_vof_cnt_zero<<<1, 64, 0, s>>>(cnt, VC_PENDING, 1);
        k_vof_mod_eval<<<G, 256, 0, s>>>(v, dt, round, h->vof_state.p, h->vof_term.p, cnt);
        int pending = 0;
        HIPCHK(hipMemcpyAsync(&pending, cnt + VC_PENDING, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        out.rounds = round;
        if (pending == 0) break;
        if (round >= (1 << 29)) return CLOWNS_EHIP;  // unreachable: every round decides a cell
    }
    k_vof_mod_apply<<<G, 256, 0, s>>>(v, h->vof_state.p, h->vof_term.p);
    // cal_new_f, with its retry at dt/1.1 (1sub-vof.F90:256-268)
    for (;;) {
        k_vof_cnt_zero<<<1, 64, 0, s>>>(cnt, VC_RETRY, 2);
        k_vof_newf<<<G8, 256, 0, s>>>(v, dt, cnt);
        int fl[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(fl, cnt + VC_RETRY, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (fl[1]) return CLOWNS_EVOF;  // the reference STOPs (F > F_lim on a surface cell)
        if (!fl[0]) break;
        dt = dt / 1.1;
        ++out.retries;
        if (dt < VOF_DTLIM) return CLOWNS_EVOF;
    }
    // simple_slide(0) (SLIDE0 builds)
    k_vof_slide<<<Gc8, 256, 0, s>>>(v, dt, cnt);
    HIPCHK(hipGetLastError());
    int c8[VC_N];
    HIPCHK(hipMemcpyAsync(c8, cnt, sizeof(c8), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipEventSynchronize(h->ev1));
    if (c8[VC_SOLID]) return CLOWNS_ENONFINITE;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    out.modified = c8[VC_MODIFIED];
    out.slid = c8[VC_SLID];
    out.dt = dt;
    out.ms = ms;
    if (info) *info = out;
    return CLOWNS_OK;
}
} // namespace

extern "C" int clowns_solver_vof(clowns_solver *h, const clowns_grid_dev *g, const double *d_Fp, int32_t mode,
                                 const double *d_qs, const int8_t *d_qs_mask, double *d_qf, double *d_Fnew,
                                 clowns_vof_info *info) {
    if (mode != 0 && mode != 1) return CLOWNS_EBADARG;
    if ((d_qs == nullptr) != (d_qs_mask == nullptr)) return CLOWNS_EBADARG;
    return vof_run(h, g, d_Fp, mode, d_qs, d_qs_mask, 0, nullptr, nullptr, d_qf, d_Fnew, info);
}

extern "C" int clowns_solver_vof_plic(clowns_solver *h, const clowns_grid_dev *g, const double *d_Fp,
                                      int32_t n_plic, const int32_t *d_plic_cells, const double *d_plic_vol,
                                      double *d_qf, double *d_Fnew, clowns_vof_info *info) {
    if (n_plic < 0 || (n_plic > 0 && (!d_plic_cells || !d_plic_vol))) return CLOWNS_EBADARG;
    return vof_run(h, g, d_Fp, 0, nullptr, nullptr, n_plic, d_plic_cells, d_plic_vol, d_qf, d_Fnew, info);
}

extern "C" int clowns_solver_vof_time_kernel(clowns_solver *h, const clowns_grid_dev *g, const double *d_Fp,
                                             int32_t kernel, int32_t reps, double *d_qf, double *d_Fnew,
                                             float *ms_per_launch) {
    if (kernel < 0 || kernel > 3 || reps < 1 || !ms_per_launch) return CLOWNS_EBADARG;
    VofDev v;
    int rc = vof_setup(h, g, d_Fp, d_qf, d_Fnew, v);
    if (rc) return rc;
    hipStream_t s = h->stream;
    const int G8 = (cdiv(v.ncell, 256) + 7) & ~7, Gc8 = (cdiv((long)v.nx * v.ny, 256) + 7) & ~7;
    auto launch = [&]() {
        switch (kernel) {
        case 0: k_vof_flux<1><<<G8, 256, 0, s>>>(v, g->dt, nullptr, nullptr); break;
        case 1: k_vof_newf<<<G8, 256, 0, s>>>(v, g->dt, h->vof_cnt.p); break;
        case 2: k_vof_slide<<<Gc8, 256, 0, s>>>(v, g->dt, h->vof_cnt.p); break;
        default: k_copy<double><<<cdiv(v.ncell, BS), BS, 0, s>>>((int)v.ncell, g->F, d_Fnew); break;
        }
    };
    launch();  // warm-up
    HIPCHK(hipEventRecord(h->ev0, s));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHK(hipEventRecord(h->ev1, s));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *ms_per_launch = ms / reps;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_tuning(clowns_solver *h, int32_t out[8]) {
    if (!h || !out || h->N < 1) return CLOWNS_EBADARG;
    const int N = h->N;
    const int fz = h->kn.fuse_pm;
    out[0] = fz >= 0 ? fz : (tuned_size(N, h->fuse64_N) ? h->fuse64 : -1);
    out[1] = fz >= 0 ? fz : (tuned_size(N, h->fuse32_N) ? h->fuse32 : -1);
    int f[2] = {-1, -1};
    if (!stored_forms(h, op64(h), &f[0])) f[0] = -1;
    if (!stored_forms(h, op32(h), &f[1])) f[1] = -1;
    out[2] = f[0] < 0 ? -1 : (f[0] & 3);
    out[3] = f[1] < 0 ? -1 : (f[1] & 3);
    out[4] = f[0] < 0 ? -1 : (f[0] & 4 ? 8 : (h->ll16_ok ? 16 : 32));
    out[5] = f[1] < 0 ? -1 : (f[1] & 4 ? 8 : (h->ll16_ok ? 16 : 32));
    out[6] = f[0] < 0 ? -1 : (f[0] & 8 ? 1 : 0);
    out[7] = f[1] < 0 ? -1
                      : (!h->have_rows ? 0 : (h->kn.spmv32_p2 ? 2 : 1)) |
                            (h->pat_ok ? 4 : 0);
    return CLOWNS_OK;
}

extern "C" int clowns_solver_set_block_order(clowns_solver *h, const int32_t *d_order, int32_t G) {
    if (!h || h->N < 1 || G < 0) return CLOWNS_EBADARG;
    int rc = set_device(h);
    if (rc) return rc;
    h->order_pending = false;  // the caller's order replaces the grid's
    if (!d_order || G == 0) {
        h->order_G = 0;
        return CLOWNS_OK;
    }
    if (G != geo_of(h->N, 2).G) return CLOWNS_EBADARG;
    // a permutation of 0..G-1, checked on the host before any kernel uses it
    std::vector<int> ho(G);
    HIPCHK(hipMemcpy(ho.data(), d_order, (size_t)G * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<char> seen(G, 0);
    for (int v : ho) {
        if (v < 0 || v >= G || seen[v]) return CLOWNS_EBADARG;
        seen[v] = 1;
    }
    if ((rc = h->order.ensure(G))) return rc;
    HIPCHK(hipMemcpy(h->order.p, ho.data(), (size_t)G * sizeof(int), hipMemcpyHostToDevice));
    h->order_G = G;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_patterns(clowns_solver *h, int32_t *npat, int32_t *rows) {
    if (!h || h->N < 1) return CLOWNS_EBADARG;
    if (npat) *npat = h->pat_ok ? h->npat : 0;
    if (rows) *rows = h->pat_ok ? h->pat_rows : 0;
    return CLOWNS_OK;
}

extern "C" int clowns_solver_index_bits(clowns_solver *h) {
    if (!h || h->N < 1) return -CLOWNS_EBADARG;
    int f;
    if (h->ll8_ok && (h->kn.ll8 == 1 || (stored_forms(h, op64(h), &f) && (f & 4)))) return 8;
    return h->ll16_ok ? 16 : 32;
}

extern "C" const char *clowns_version(void) {
    return "clowns_pcg gfx950 (PM-Bi-CGSTAB, FP64 + mixed FP32/FP64)";
}
