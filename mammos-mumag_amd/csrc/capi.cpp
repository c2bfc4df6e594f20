// C ABI of libmumag_hip.so: context creation, the energy/gradient evaluation,
// the Jacobi PCG (FP64, or FP32 with FP64 reliable updates), the truncated-CG
// preconditioner and the nonlinear-CG minimiser.
//
// The minimiser is a host-side control loop over device kernels; each iteration
// returns only the few scalars the reference's control flow branches on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mumag_hip.h"
#include "amg.h"
#include "assemble.h"
#include "common.h"
#include "mg.h"
#include "mlp.h"
#include "structured.h"

using namespace mm;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
}  // namespace

#define HIPC(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail(MM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

struct KernelTimer {
    std::vector<hipEvent_t> ev;  // start/stop pairs, one per launch of a solve
    int64_t launches = 0;
    double ms = 0.0;
    int64_t bytes = 0;           // algorithmic bytes of one launch
};

struct mm_ctx {
    int device = 0;
    hipStream_t s = nullptr;
    int precision = MM_PREC_FP64;
    bool structured = false;
    int64_t N = 0, NE = 0, NEm = 0, Nm = 0;
    double V = 0.0;
    int pcg_batch = 16;
    double ru_delta = 0.1;   // mixed precision: residual reduction between reliable updates
    bool ru_delta_set = false;  // given by the caller (mm_options.inner_tol)
    // MM_PREC_MIXED with eval_precision MM_PREC_MIXED: the energy evaluation's C m, D m and
    // G u in FP32 (structured kernels; the element kernels are FP64 either way)
    bool eval32 = false;
    int mg_built = 0, mg_built_max = 0;   // levels built by mm_multigrid and the bound they were built for
    int64_t k_nnz = 0, k_entries = 0;
    int64_t pc_entries = 0;   // entries of the assembled C (general meshes)
    size_t bytes = 0;
    std::vector<void *> allocs;

    double *meas = nullptr, *Dex = nullptr, *invDK = nullptr;
    float *invDK32 = nullptr;
    SellDev<double> K;
    SellDev<float> K32;
    ElemDev E;
    PairCDev PC;   // assembled C (general meshes) for the applies without D v
    GridDev G;
    TcgBlock<float> tb32;    // truncated CG on the magnet block (structured grids)
    bool ncg_block = false;  // the NCG's reductions visit the magnet block only (this call)
    TcgBlock<double> tb64;
    PoissonGrid<double> P;
    PoissonGrid<float> P32;
    MatCoef *mats = nullptr;
    int n_mats = 0;
    DevScalars *dsc = nullptr;
    DevScalars *hsc = nullptr;  // pinned host mirror (written by the device, k_publish)
    unsigned long long *hseq = nullptr, seq = 0;   // the mirror's sequence number (same allocation)
    Reducer red;

    // workspace (N)
    double *b = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr;
    float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *p32b = nullptr, *q32 = nullptr, *dx32 = nullptr;
    double *partsA = nullptr, *partsB = nullptr;  // partials of the fused PCG kernels
    double *ncg_parts = nullptr;                  // partials of the NCG scan (NCG_NV slots)
    // workspace (3N)
    double *gC = nullptr, *gd = nullptr;
    double *tp = nullptr, *tr = nullptr, *tz = nullptr, *td = nullptr, *tCd = nullptr, *tHd = nullptr;
    double *X = nullptr, *X0 = nullptr, *X1 = nullptr, *Gp = nullptr, *G1 = nullptr, *Dd = nullptr,
           *Z1 = nullptr, *gtmp = nullptr;

    // event timing of the PCG kernels (K apply and vector update)
    bool timing = false;
    KernelTimer t_spmv, t_upd, t_mg;   // t_mg: a level-0 smoothing step of the V-cycle
    // counts since the timing was enabled: true residual passes and corrections of the
    // mixed PCG (host-counted: they always run), V-cycles (device-counted) at enable time
    int64_t n_ru = 0, n_corr = 0;
    long long vcyc0 = 0;

    // multigrid preconditioner of the FP32 inner solve (structured grids)
    std::vector<double> axis[3];        // node coordinates per axis (host)
    // smoothed-aggregation AMG of the general meshes' inner solve (mm_amg): K on the host
    // (general contexts), the levels (level 0's matrix is K32), the coarsest inverse, the
    // first smoothing step's weights of level 0 and its residual vector
    CsrHost hK;
    int amg_nl = 0, amg_built = 0, amg_built_max = 0, amg_nc = 0;
    AmgLevelDev amg[kAmgMaxLevels];
    float *amg_inv = nullptr, *amg_invw0 = nullptr, *amg_r0 = nullptr;
    int mg_nl = 0;                      // levels (0: Jacobi)
    int mg_batch = 8;                   // most inner iterations between host checks with multigrid
    double mg_rate2 = 0.05;             // running r.r reduction per multigrid-PCG iteration
    int mg_deg = 3, mg_cdeg = 12;       // smoothing degree, coarsest-level polynomial degree
    int mg_deg0 = 4;                    // smoothing degree on the finest level
    MgLevelDev mg[kMgMaxLevels];
    float *mg_b[kMgMaxLevels] = {}, *mg_xa[kMgMaxLevels] = {}, *mg_xb[kMgMaxLevels] = {},
          *mg_r[kMgMaxLevels] = {};
    float mg_om[kMgMaxDeg] = {}, mg_omc[kMgMaxDeg] = {}, mg_om0[kMgMaxDeg] = {};
    int mg_tail = 0;                    // levels >= mg_tail run in the one-workgroup tail kernel
    MgTailArgs tail;

    // learned warm start (structured grids): network, features, predictions
    FeatGrid F;
    mm_mlp *ws_net = nullptr;
    double ws_scale = 0.0;
    uint16_t *ws_X = nullptr;        // [N][kMlpIn] features
    float *ws_pred = nullptr, *ws_prev = nullptr;
    bool ws_have_prev = false;
    double *ws_red = nullptr;        // [0..2] mbar sums, [4] num, [5] unused, [6] den, [7] gamma
    int64_t ws_calls = 0;
    // secant start of the line search's later potential solves (mm_secant_start): u at the
    // line search's first evaluation; [0] num, [2] den, [3] gamma of the Galerkin step
    bool sec_on = false, sec_next = false;
    double *u_sec = nullptr, *sec_red = nullptr;
    int64_t sec_calls = 0;
    // MM_WS_MEASURED: host-timed solves with and without the warm start (running means)
    int ws_policy = 0;
    int64_t ws_solves = 0, ws_n[2] = {0, 0};
    int64_t ws_every = 16, ws_next = 15;   // re-measurement interval and the next one
    bool ws_probe = false;                 // the current solve re-measures the other form
    // learned preconditioner (mm_lp_setup): FP64 flexible-PCG vectors, scaled residual,
    // features and network output
    double *lp_r = nullptr, *lp_z = nullptr, *lp_p = nullptr, *lp_q = nullptr, *lp_rold = nullptr,
           *lp_w = nullptr;
    uint16_t *lp_X = nullptr;
    float *lp_y = nullptr;
    double ws_mean[2] = {0.0, 0.0};

    Launch L() const {
        Launch l;
        l.s = s;
        l.red = red;
        l.dsc = dsc;
        l.mats = mats;
        return l;
    }

    template <typename T>
    int alloc(T **ptr, size_t count) {
        void *p_ = nullptr;
        hipError_t e = hipMalloc(&p_, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(MM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(p_);
        bytes += count * sizeof(T);
        *ptr = (T *)p_;
        // the context's stream does not wait for the null stream: the zeroing completes
        // here, before any kernel of a later call can write the buffer
        e = hipMemset(p_, 0, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) return fail(MM_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return 0;
    }
    template <typename T>
    int upload(T **ptr, const std::vector<T> &v) {
        int rc = alloc(ptr, v.size());
        if (rc) return rc;
        if (!v.empty()) HIPC(hipMemcpy(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return 0;
    }
    ~mm_ctx() {
        if (s) (void)hipStreamSynchronize(s);
        for (void *a : allocs) (void)hipFree(a);
        if (hsc) (void)hipHostFree(hsc);
        for (hipEvent_t e : t_spmv.ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : t_upd.ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : t_mg.ev) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }
};

// The per-node MLP: FP32 master parameters with Adam moments, the bf16 copies the
// kernels read, and the training workspace for batches of up to max_batch nodes.
struct mm_mlp {
    mm_ctx *ctx = nullptr;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr, *w3t = nullptr, *w2n = nullptr, *w3n = nullptr;
    int64_t max_batch = 0, ld_max = 0;
    uint16_t *h[3] = {nullptr, nullptr, nullptr}, *dz[3] = {nullptr, nullptr, nullptr};
    float *y = nullptr, *dy = nullptr;
    double *loss = nullptr;
    int step = 0;
    std::vector<void *> allocs;
    MlpDev dev() const {
        MlpDev w;
        w.w1t = w1t;
        w.w2t = w2t;
        w.w3t = w3t;
        w.w2n = w2n;
        w.w3n = w3n;
        w.b1 = p + kMlpOffB1;
        w.b2 = p + kMlpOffB2;
        w.b3 = p + kMlpOffB3;
        w.w4 = p + kMlpOffW4;
        w.b4 = p + kMlpOffB4;
        return w;
    }
    template <typename T>
    int alloc(T **ptr, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(MM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *ptr = (T *)q;
        e = hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);   // see mm_ctx::alloc
        if (e != hipSuccess) return fail(MM_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return 0;
    }
    ~mm_mlp() {
        if (ctx && ctx->s) (void)hipStreamSynchronize(ctx->s);
        for (void *a : allocs) (void)hipFree(a);
    }
};

namespace {

// The device scalars into the host mirror: a one-wave kernel writes them into pinned host
// memory and then a new sequence number behind a system-scope release; the host spins on
// the number (a stream synchronize after a device-to-host copy costs ~5 us more per round
// trip), polling the stream now and then so that a failed launch returns its error.
int sync_scalars(mm_ctx *c) {
    const unsigned long long seq = ++c->seq;
    launch_publish(c->L(), c->dsc, c->hsc, c->hseq, seq);
    for (unsigned k = 1;; ++k) {
        if (__atomic_load_n(c->hseq, __ATOMIC_ACQUIRE) == seq) return 0;
        if ((k & 0x3fffu) == 0) {
            const hipError_t e = hipStreamQuery(c->s);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return fail(MM_E_HIP, std::string("stream: ") + hipGetErrorString(e));
            // the stream has drained: the number is there unless the launch failed
            if (__atomic_load_n(c->hseq, __ATOMIC_ACQUIRE) == seq) return 0;
            const hipError_t le = hipGetLastError();
            return fail(MM_E_HIP, std::string("scalar publish did not complete: ") + hipGetErrorString(le));
        }
    }
}

// ---------------------------------------------------------------- operators

void op_K(mm_ctx *c, const double *x, double *y, double *dot, const int *skip) {
    if (c->structured)
        launch_grid_poisson(c->L(), c->P, x, y, dot, skip);
    else
        launch_sell_spmv(c->L(), c->K, x, y, dot, skip);
}
// y = C v and optionally b = Dx vx + Dy vy + Dz vz.  Structured kernels overwrite the
// magnet entries only (the others stay zero); the element kernels accumulate.
void op_C(mm_ctx *c, const double *v, double *y, double *b, const int *skip) {
    Launch L = c->L();
    if (c->structured) {
        launch_grid_exani(L, c->G, c->eval32, v, y, b, skip);
    } else {
        launch_pair_c(L, c->PC, v, y, b, skip);   // writes every node of y (and b)
    }
}
void op_grad(mm_ctx *c, const double *u, double *g) {
    Launch L = c->L();
    if (c->structured) {
        launch_grid_grad(L, c->G, c->eval32, u, g);
    } else {
        launch_ebe_grad(L, c->E, u, g, 1.0 / c->V);   // writes every node of g
    }
}

// The PCG kernels are timed on a sample: iteration i is bracketed by an event pair when
// i % stride == phase (mid-batch, away from the host synchronisation), so that the
// events barely perturb the timed solve: 1 in 16 with Jacobi, every other iteration with
// multigrid (whose solves take a few iterations each).  Pair index i / stride; pairs
// past the pool are not recorded.
struct TimerSample {
    int stride, phase;
};
TimerSample timer_sample(const mm_ctx *c) { return c->mg_nl > 0 ? TimerSample{2, 1} : TimerSample{16, 7}; }
// the launch of iteration i, with the timer's event pair attached when i is sampled
Launch timed(mm_ctx *c, KernelTimer &t, int i) {
    Launch l = c->L();
    const TimerSample ts = timer_sample(c);
    if (!c->timing || i % ts.stride != ts.phase) return l;
    const int idx = i / ts.stride;
    if (idx >= (int)t.ev.size() / 2) return l;
    l.t0 = t.ev[2 * idx];
    l.t1 = t.ev[2 * idx + 1];
    return l;
}
// add the durations of the sampled iterations in [from, from + count) (after the stream
// has synchronised)
int collect(const mm_ctx *c, KernelTimer &t, int from, int count) {
    const int pool = (int)t.ev.size() / 2;
    const TimerSample ts = timer_sample(c);
    for (int i = from; i < from + count; ++i) {
        if (i % ts.stride != ts.phase || i / ts.stride >= pool) continue;
        const int j = i / ts.stride;
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, t.ev[2 * j], t.ev[2 * j + 1]));
        t.ms += ms;
        t.launches++;
    }
    return 0;
}

// One fused PCG iteration (2 kernels).  p alternates between two buffers.
template <typename T>
void pcg_iteration(mm_ctx *c, int i, int mode, double tol, double atol, int maxiter, T *z, T *const pb[2],
                   T *q, T *x, T *r, const T *invD, const SellDev<T> &A, const PoissonGrid<T> &P, int *nB) {
    PcgIterArgs a;
    a.parts_in = c->partsB;
    a.nparts_in = *nB;
    a.parts_out = c->partsA;
    a.iter = i;
    a.mode = mode;
    a.tol = tol;
    a.atol = atol;
    a.maxiter = maxiter;
    const T *pold = pb[i & 1];
    T *pnew = pb[(i + 1) & 1];
    const Launch LA = timed(c, c->t_spmv, i);
    const int nA = c->structured ? launch_ka_grid<T>(LA, P, a, z, pold, pnew, q)
                                 : launch_ka_sell<T>(LA, A, a, z, pold, pnew, q);
    PcgIterArgs bb = a;
    bb.parts_in = c->partsA;
    bb.nparts_in = nA;
    bb.parts_out = c->partsB;
    *nB = launch_kb<T>(timed(c, c->t_upd, i), c->N, bb, x, r, invD ? z : nullptr, pnew, q, invD);
}

// jax.scipy.sparse.linalg.cg semantics, FP64; x is updated in place.  Batches of
// pcg_batch iterations are queued; the host reads the stop flag between batches and
// launches past convergence are no-ops.
int pcg64(mm_ctx *c, const double *b, double *x, double tol, double atol, int maxiter, int *iters) {
    Launch L = c->L();
    const int64_t n = c->N;
    op_K(c, x, c->q, nullptr, nullptr);
    int nB = launch_pcg0<double>(L, n, b, c->q, c->r, c->z, c->invDK, c->partsB);
    double *const pb[2] = {c->p, c->p2};
    int i = 0, rc = 0;
    for (;;) {
        for (int k = 0; k < c->pcg_batch; ++k, ++i)
            pcg_iteration<double>(c, i, i == 0 ? PCG_FIRST : PCG_NORMAL, tol, atol, maxiter, c->z, pb, c->q, x,
                                  c->r, c->invDK, c->K, c->P, &nB);
        if ((rc = sync_scalars(c))) return rc;
        if (c->hsc->pcg_done) break;
    }
    const int it = c->hsc->pcg_iter;
    if (c->timing) {
        if ((rc = collect(c, c->t_spmv, 0, it)) || (rc = collect(c, c->t_upd, 0, it))) return rc;
    }
    if (iters) *iters = it;
    if (it >= maxiter && c->hsc->pcg_rr > c->hsc->pcg_bound) return MM_W_MAXITER;
    return 0;
}

// Multigrid V-cycle M r on the FP32 vectors of the inner solve (mg.hip); skipped entirely
// once *skip is set.  Level 0 ping-pongs between z and alt; the returned buffer (one of
// the two) holds M r.  first_done: z already holds the first smoothing step
// mg[0].invw * r (the kernel that produced r wrote it).  With parts, the last level-0
// smoothing step also writes the partials (r.Mr, r.r) there; *nparts is their count.
float *mg_vcycle(mm_ctx *c, const float *r, float *z, float *alt, const int *skip, bool first_done,
                 double *parts = nullptr, int *nparts = nullptr, int it = -1) {
    Launch L = c->L();
    const int nl = c->mg_nl, cd = c->mg_cdeg;
    float *cur[kMgMaxLevels], *oth[kMgMaxLevels];
    const float *B[kMgMaxLevels];
    for (int l = 0; l < nl; ++l) {
        B[l] = l == 0 ? r : c->mg_b[l];
        cur[l] = l == 0 ? z : c->mg_xa[l];
        oth[l] = l == 0 ? alt : c->mg_xb[l];
    }
    if (!first_done) launch_mg_first(L, c->mg[0].P, r, z, nl > 1 ? c->mg_om0[0] : c->mg_omc[0], skip);
    // the last level-0 smoothing step carries the dot when the stencil grid has a
    // partial slot per block
    const bool fuse = parts && mg_stencil_fits(c->mg[0].P);
    int np = 0;
    auto smooth = [&](int l, float w, bool last0, bool sample = false) {
        // sample: the PCG iteration's timed level-0 step (HIP events on its dispatch)
        const Launch LS = sample && it >= 0 ? timed(c, c->t_mg, it) : L;
        const int nb = launch_mg_smooth(LS, c->mg[l].P, B[l], cur[l], oth[l], w, skip,
                                        last0 && fuse ? parts : nullptr);
        if (last0 && fuse) np = nb;
        std::swap(cur[l], oth[l]);
    };
    // two steps in one launch on the coarse levels (l >= 1: launch-bound kernels)
    auto smooth2 = [&](int l, float w1, float w2) {
        launch_mg_smooth2(L, c->mg[l].P, B[l], cur[l], oth[l], w1, w2, skip);
        std::swap(cur[l], oth[l]);
    };
    // levels below top run one kernel per step; from top on, the tail kernel (or, with no
    // tail, the coarsest level's polynomial)
    const bool tail = c->mg_tail < nl;
    const int top = tail ? c->mg_tail : nl - 1;
    // pre-smoothing (the first step of every level is already in cur[l]), residual, restriction
    // (level 0 smooths with its own degree and weights)
    for (int l = 0; l < top; ++l) {
        const int s = l == 0 ? c->mg_deg0 : c->mg_deg;
        const float *om = l == 0 ? c->mg_om0 : c->mg_om;
        int k = 1;
        for (; l > 0 && k + 1 < s; k += 2) smooth2(l, om[k], om[k + 1]);
        for (; k < s; ++k) smooth(l, om[k], false);
        launch_mg_resid(L, c->mg[l].P, B[l], cur[l], c->mg_r[l], skip);
        launch_mg_restrict(L, c->mg[l], c->mg[l + 1], c->mg_r[l], c->mg_b[l + 1], cur[l + 1], skip);
    }
    if (tail) {
        launch_mg_tail(L, c->tail, skip);   // reads mg_b[top], writes mg_xa[top]
        cur[top] = c->mg_xa[top];
    } else {
        for (int k = 1; k < cd; ++k) smooth(nl - 1, c->mg_omc[k], nl == 1 && k == cd - 1);
    }
    for (int l = top - 1; l >= 0; --l) {
        launch_mg_prolong(L, c->mg[l], c->mg[l + 1], cur[l + 1], cur[l], skip);
        const int s = l == 0 ? c->mg_deg0 : c->mg_deg;
        const float *om = l == 0 ? c->mg_om0 : c->mg_om;
        int k = s - 1;
        for (; l > 0 && k >= 1; k -= 2) smooth2(l, om[k], om[k - 1]);
        for (; k >= 0; --k) smooth(l, om[k], l == 0 && k == 0, l == 0 && k == s - 1 && k > 0);
    }
    if (parts && np == 0) np = launch_dot2<float>(L, c->N, r, cur[0], parts);
    if (nparts) *nparts = np;
    return cur[0];
}

// AMG V-cycle M r on the FP32 vectors of the inner solve (amg.hip), the CPU restatement's
// order (oracle/amg.py): on each level but the last, the first smoothing step from zero
// (on level 0 formed by the kernel that produced r when first_done), the other steps, the
// residual restricted with the next level's first step fused, the coarse correction
// added, the steps in reverse order; the last level by its dense inverse.  Level 0
// ping-pongs between z and alt; the returned buffer holds M r.
float *amg_vcycle(mm_ctx *c, const float *r, float *z, float *alt, const int *skip, bool first_done,
                  double *parts = nullptr, int *nparts = nullptr, int it = -1) {
    Launch L = c->L();
    const int nl = c->amg_nl;
    float *cur[kAmgMaxLevels], *oth[kAmgMaxLevels];
    const float *B[kAmgMaxLevels];
    for (int l = 0; l < nl; ++l) {
        B[l] = l == 0 ? r : c->amg[l].b;
        cur[l] = l == 0 ? z : c->amg[l].xa;
        oth[l] = l == 0 ? alt : c->amg[l].xb;
    }
    const AmgLevelDev *A = c->amg;
    // level 0 in SELL-64, the coarser levels in CSR
    auto smooth = [&](const Launch &LS, int l, float w) {
        if (l == 0) launch_amg_smooth(LS, A[0].A, A[0].dinv, B[0], cur[0], oth[0], w, skip);
        else launch_amg_smooth(LS, A[l].Ac, A[l].dinv, B[l], cur[l], oth[l], w, skip);
        std::swap(cur[l], oth[l]);
    };
    if (!first_done) launch_amg_smooth(L, A[0].A, A[0].dinv, r, nullptr, z, A[0].om[0], skip);
    for (int l = 0; l + 1 < nl; ++l) {
        for (int k = 1; k < A[l].deg; ++k) smooth(L, l, A[l].om[k]);
        float *res = l == 0 ? c->amg_r0 : A[l].r;
        if (l == 0) launch_amg_resid(L, A[0].A, B[0], cur[0], res, skip);
        else launch_amg_resid(L, A[l].Ac, B[l], cur[l], res, skip);
        const bool smooth_next = l + 2 < nl;   // the last level is solved, not smoothed
        launch_amg_restrict(L, A[l].R, res, A[l + 1].b, smooth_next ? cur[l + 1] : nullptr, A[l + 1].dinv,
                            A[l + 1].om[0], skip);
    }
    launch_amg_dense(L, c->amg_inv, c->amg_nc, B[nl - 1], cur[nl - 1], skip);
    for (int l = nl - 2; l >= 0; --l) {
        launch_amg_prolong(L, A[l].P, cur[l + 1], cur[l], skip);
        for (int k = A[l].deg - 1; k >= 0; --k)   // the PCG iteration's timed level-0 step
            smooth(l == 0 && k == A[l].deg - 1 && it >= 0 ? timed(c, c->t_mg, it) : L, l, A[l].om[k]);
    }
    int np = 0;
    if (parts) np = launch_dot2<float>(L, c->N, r, cur[0], parts);
    if (nparts) *nparts = np;
    return cur[0];
}

// Mixed precision with reliable updates: the Jacobi-PCG recursion (alpha, beta, r, z, p
// and the correction dx) runs in FP32 with FP64 dot products; whenever the FP32
// recursive residual has dropped by ru_delta since the last replacement, or claims
// convergence, x += dx in FP64 and the residual is replaced by the true FP64 residual
// b - K x; with Jacobi the search direction continues as p = z + beta p with gamma from
// the true residual, with multigrid the recursion restarts from p = M r.  Stops when the
// true residual satisfies r.r <= max(tol^2 b.b, atol^2), the FP64 path's bound.
// With multigrid the segment's target is the inner iteration's own stop bound, tested
// after each KB before the V-cycle (k_pcg_check), and the V-cycle of a batch's last
// iteration is launched after the host has read the flags, only when the segment goes on:
// no V-cycle runs for a residual that is about to be replaced or has converged.
int pcg_mixed(mm_ctx *c, const double *b, double *x, double tol, double atol, int maxiter, int *iters) {
    Launch L = c->L();
    const int64_t n = c->N;
    const double delta2 = c->ru_delta * c->ru_delta;
    const bool mg = c->mg_nl > 0 || c->amg_nl > 0;
    const double seg2 = mg ? delta2 : 0.0;
    // with multigrid, every kernel that forms z from r writes the V-cycle's first step
    const float *invD = c->mg_nl > 0 ? c->mg[0].invw : c->amg_nl > 0 ? c->amg_invw0 : c->invDK32;
    // r = b - K x in FP64 (one pass on the structured grid), its FP32 copy and z
    const bool fused = c->structured && grid_ru_residual_fits(c->P);
    auto true_residual = [&](bool init) {
        c->n_ru++;
        if (fused) {
            launch_grid_ru_residual(L, c->P, x, b, c->r32, c->z32, c->p32, invD, tol, atol, maxiter, init, seg2);
        } else {
            op_K(c, x, c->q, nullptr, nullptr);
            launch_ru_residual(L, n, b, c->q, c->r32, c->z32, c->p32, invD, tol, atol, maxiter, init, seg2);
        }
    };
    true_residual(true);
    launch_fill_zero32(L, n, c->dx32, nullptr);
    int rc = sync_scalars(c);
    if (rc) return rc;
    float *const pb[2] = {c->p32, c->p32b};
    int i = 0, nB = 0, mode = PCG_RESTART;
    int *const done = &c->dsc->pcg_done;
    // z = M r; the V-cycle's result buffer becomes z32 (the other one its scratch)
    auto precondition = [&](int it) {
        if (c->amg_nl > 0) {
            float *res = amg_vcycle(c, c->r32, c->z32, c->amg[0].xb, done, true, c->partsB, &nB, it);
            if (res != c->z32) std::swap(c->z32, c->amg[0].xb);
            return;
        }
        float *res = mg_vcycle(c, c->r32, c->z32, c->mg_xb[0], done, true, c->partsB, &nB, it);
        if (res != c->z32) std::swap(c->z32, c->mg_xb[0]);
    };
    if (mg && !c->hsc->ir_done) {   // for the initial true residual
        precondition(-1);
        mode = PCG_RESTART_PARTS;
    }
    while (!c->hsc->ir_done) {
        // FP32 iterations until the recursive residual has dropped by ru_delta since the
        // last true residual (with multigrid tested on the device after every KB, else
        // between batches) or claims convergence
        const double rmax2 = c->hsc->ir_rmax2;
        const int seg0 = i;
        // multigrid: each batch is sized to reach the segment's target at the measured
        // convergence rate (rr per iteration, a running value over segments), so that few
        // iterations are queued past the stop and few host checks are needed
        const double target2 = std::max(delta2 * rmax2, c->hsc->ir_bound);
        double rr_ref = rmax2;   // r.r at iteration j_ref (the replaced residual at seg0)
        int j_ref = seg0;
        for (;;) {
            int batch = c->pcg_batch;
            if (mg) {
                const double need = std::log(target2 / rr_ref) / std::log(c->mg_rate2);
                batch = (int)std::ceil(std::max(need, 0.0));
                batch = std::min(std::max(batch, 1), c->mg_batch);
            }
            for (int k = 0; k < batch; ++k, ++i) {
                pcg_iteration<float>(c, i, mode, tol, atol, maxiter, c->z32, pb, c->q32, c->dx32, c->r32, invD,
                                     c->K32, c->P32, &nB);
                if (mg) {
                    launch_pcg_check(L, c->partsB, nB, i, maxiter);
                    if (k + 1 < batch) precondition(i);
                }
                mode = PCG_NORMAL;
            }
            if ((rc = sync_scalars(c))) return rc;
            // pcg_rr is the r.r entering iteration jr: the last KA's, or at a stop the
            // stopped index's
            const int jr = c->hsc->pcg_done ? c->hsc->pcg_iter : i - 1;
            if (mg && jr > j_ref && c->hsc->pcg_rr > 0.0 && rr_ref > 0.0) {
                const double r2 = std::pow(c->hsc->pcg_rr / rr_ref, 1.0 / (jr - j_ref));
                c->mg_rate2 = std::min(std::max(0.5 * c->mg_rate2 + 0.5 * r2, 1e-6), 0.9);
                rr_ref = c->hsc->pcg_rr;
                j_ref = jr;
            }
            if (c->hsc->pcg_done || c->hsc->pcg_rr < delta2 * rmax2) break;
            if (mg) precondition(i - 1);   // the batch's last iteration goes on
        }
        // the stopped iteration index; the next segment continues from it with p kept
        // (the kernels queued after it in the batch did nothing)
        if (c->hsc->pcg_done) i = c->hsc->pcg_iter;
        if (c->timing) {
            const int used = i - seg0;  // event pairs are indexed by the global iteration number
            // (the segment's last iteration ran no V-cycle)
            if ((rc = collect(c, c->t_spmv, seg0, used)) || (rc = collect(c, c->t_upd, seg0, used)) ||
                (mg && (rc = collect(c, c->t_mg, seg0, used - 1))))
                return rc;
        }
        // reliable update: fold dx into x, replace r by the true residual
        launch_add_correction(L, n, x, c->dx32);
        c->n_corr++;
        true_residual(false);
        if ((rc = sync_scalars(c))) return rc;
        if (c->hsc->ir_done || i >= maxiter) break;
        mode = PCG_KEEP;
        if (mg) {
            // z = M r for the replaced residual and a restart (p = z).  The multigrid
            // recursion runs far below the FP32 residual's attainable accuracy between
            // replacements, so the true residual's r.z can exceed the recursive one by
            // orders of magnitude: continuing the direction (beta = gamma_true /
            // gamma_recursive) would wreck the conjugacy.
            precondition(-1);
            mode = PCG_RESTART_PARTS;
        }
    }
    if (iters) *iters = i;
    if (!c->hsc->ir_done) return MM_W_MAXITER;
    return 0;
}

int pcg(mm_ctx *c, const double *b, double *x, double tol, double atol, int maxiter, int *iters) {
    if (c->precision == MM_PREC_MIXED) return pcg_mixed(c, b, x, tol, atol, maxiter, iters);
    return pcg64(c, b, x, tol, atol, maxiter, iters);
}

// Learned warm start of the potential solve for b = D m: the network's prediction for m
// gives a direction d (the change of the prediction since the previous solve, or on the
// first solve the prediction minus u); u moves by the Galerkin step gamma d that
// minimises the K-energy norm of the error along d, gamma = d.(b - K u) / d.K d.  This is
// the first step of a flexible PCG whose first preconditioner application is the
// network; the Jacobi PCG that follows ends at the same tolerance, so the answer is the
// plain method's.  No host synchronisation.
void learned_warm_start(mm_ctx *c, const double *m, const double *b, double *u) {
    Launch L = c->L();
    const int64_t n = c->N;
    for (int a = 0; a < 3; ++a)
        launch_reduce(L, RED_MH, n, m, c->meas, nullptr, nullptr, a == 0, a == 1, a == 2, c->ws_red + a);
    launch_mlp_features(L, c->F, m, c->ws_red, c->ws_X);
    launch_mlp_forward(L, c->ws_net->dev(), n, c->ws_X, c->ws_pred);
    double *d = c->z, *Ku = c->q, *Kd = c->r;   // PCG workspace, free before the solve
    launch_mlp_direction(L, c->F, c->ws_pred, c->ws_have_prev ? c->ws_prev : nullptr, u, c->ws_scale, d);
    op_K(c, u, Ku, nullptr, nullptr);
    op_K(c, d, Kd, nullptr, nullptr);
    launch_reduce(L, RED_YZD, n, Ku, b, d, d, 0, 0, 0, c->ws_red + 4);
    launch_reduce(L, RED_DOT, n, d, Kd, nullptr, nullptr, 0, 0, 0, c->ws_red + 6);
    launch_mlp_galerkin(L, n, c->ws_red + 4, c->ws_red + 6, d, u, c->ws_red + 7);
    std::swap(c->ws_pred, c->ws_prev);
    c->ws_have_prev = true;
    c->ws_calls++;
}

// MM_WS_EVERY: every solve is warm-started.  MM_WS_MEASURED: each solve takes the form
// (warm-started or plain) with the lower mean wall time of the solves that used it, the
// warm start plus the PCG, measured on the host; the first 3 solves of each form
// alternate; then the other form is re-measured on solve 15 and after that at an
// interval that doubles (16, 32, ..., 256 solves) while it stays the slower one and
// returns to 16 when it wins.
constexpr int64_t kWsEvery = 16, kWsEveryMax = 256;
bool ws_choose(mm_ctx *c) {
    c->ws_probe = false;
    if (c->ws_policy != MM_WS_MEASURED) return true;
    const int64_t k = c->ws_solves;
    if (c->ws_n[0] < 3 || c->ws_n[1] < 3) return (k & 1) == 0;
    const bool best = c->ws_mean[1] < c->ws_mean[0];
    c->ws_probe = k >= c->ws_next;
    return c->ws_probe ? !best : best;
}
void ws_record(mm_ctx *c, bool ws, double seconds) {
    const int a = ws ? 1 : 0;
    c->ws_n[a]++;
    const double w = c->ws_n[a] < 8 ? 1.0 / c->ws_n[a] : 0.125;   // mean, then a running mean
    c->ws_mean[a] += w * (seconds - c->ws_mean[a]);
    if (c->ws_probe) {
        const bool still_slower = ws ? c->ws_mean[1] >= c->ws_mean[0] : c->ws_mean[0] > c->ws_mean[1];
        c->ws_every = still_slower ? std::min(2 * c->ws_every, kWsEveryMax) : kWsEvery;
        c->ws_next = c->ws_solves + c->ws_every;
        c->ws_probe = false;
    }
    c->ws_solves++;
}

// u = K^-1 b for b = D m (already in c->b), from the caller's u, with the learned warm
// start when a network is enabled and the policy picks it.
// Secant start (mm_secant_start): for b = D m of a later evaluation of the line search, u
// (the previous evaluation's solution) moves by the Galerkin step along d = u - u_sec, the
// change of the solution since the line search's first evaluation: gamma = d.(b - Ku) /
// d.Kd (device-side, no host sync; d is zero on the Dirichlet nodes).  The PCG then ends
// at the unchanged tolerance.
void secant_start(mm_ctx *c, const double *b, double *u) {
    Launch L = c->L();
    const int64_t n = c->N;
    double *d = c->z, *Ku = c->q, *Kd = c->r;   // PCG workspace, free before the solve
    launch_axpby(L, n, 1.0, u, -1.0, c->u_sec, d);
    op_K(c, u, Ku, nullptr, nullptr);
    op_K(c, d, Kd, nullptr, nullptr);
    launch_reduce(L, RED_YZD, n, Ku, b, d, d, 0, 0, 0, c->sec_red);   // (b - Ku).d
    launch_reduce(L, RED_DOT, n, d, Kd, nullptr, nullptr, 0, 0, 0, c->sec_red + 2);
    launch_mlp_galerkin(L, n, c->sec_red, c->sec_red + 2, d, u, c->sec_red + 3);
    c->sec_calls++;
}

int potential_solve(mm_ctx *c, const double *m, double *u, double tol, int *it) {
    if (c->sec_next) {
        c->sec_next = false;
        c->ws_have_prev = false;
        secant_start(c, c->b, u);
        return pcg(c, c->b, u, tol, 0.0, 10000, it);
    }
    const bool ws = c->ws_net && ws_choose(c);
    const auto t0 = std::chrono::steady_clock::now();
    if (ws) learned_warm_start(c, m, c->b, u);
    else c->ws_have_prev = false;   // the next direction starts from the prediction
    const int rc = pcg(c, c->b, u, tol, 0.0, 10000, it);   // returns after the solve has ended
    if (c->ws_net && rc >= 0)
        ws_record(c, ws, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return rc;
}

// A reduction over the NCG's interleaved 3-vectors queued behind an evaluation, so that
// its totals arrive with the energy in one host round trip (into red[0 .. nout)).
struct PostReduce {
    RedOp op;
    const double *a, *b, *cc, *d;
    int nout;
    double *out;
};
void launch_reduce3(mm_ctx *c, RedOp op, const double *a, const double *b, const double *cc, const double *d);

// E, g_proj, g_full for magnetisation m (energies.py:67-99).  off_zero: on a structured
// grid, m is zero off the magnet block and gproj, gfull already hold zero there.
int total_eg(mm_ctx *c, const double *m, double hext, const double *hdir, double tol_u,
             int hmag_on, double *u, double *E, double *gproj, double *gfull, int *pcg_it,
             bool off_zero = false, const PostReduce *post = nullptr) {
    Launch L = c->L();
    op_C(c, m, c->gC, hmag_on > 0 ? c->b : nullptr, nullptr);
    int it = 0, rc = 0;
    if (hmag_on > 0) {
        rc = potential_solve(c, m, u, tol_u, &it);
        if (rc < 0) return rc;
        op_grad(c, u, c->gd);
    }
    const double zc0 = -hext * hdir[0] / c->V, zc1 = -hext * hdir[1] / c->V,
                 zc2 = -hext * hdir[2] / c->V;
    launch_finish_eg(L, c->N, c->structured ? &c->G : nullptr, m, c->meas, zc0, zc1, zc2, c->gC, c->gd, gfull,
                     gproj, &c->dsc->energy, hmag_on > 0, !(off_zero && c->structured));
    if (post) launch_reduce3(c, post->op, post->a, post->b, post->cc, post->d);
    int rc2 = sync_scalars(c);
    if (rc2) return rc2;
    *E = c->hsc->energy;
    if (post)
        for (int k = 0; k < post->nout; ++k) post->out[k] = c->hsc->red[k];
    if (pcg_it) *pcg_it = it;
    return rc;
}

// M_inv (loop.py:40-55): truncated CG on H with grad_fk = -g.  No host sync.
// zout_off_zero: on a structured grid, g is zero off the magnet block and zout already
// holds zero there.
// The truncated CG's direction ring for at least `slots` directions (max_iter + 1); a
// larger ring replaces the old one (freed once the stream has drained).
constexpr int kTcgSlots = 11;   // precond_iter 10, the reference's default
template <typename T>
int tcg_ring(mm_ctx *c, TcgBlock<T> &B, int slots) {
    if (B.slots >= slots) return 0;
    const size_t n1 = (size_t)c->G.mn + 1, n3 = 3 * n1 * n1 * n1;
    if (B.dring) {
        HIPC(hipStreamSynchronize(c->s));
        for (void *old : {(void *)B.dring, (void *)B.alpha}) {
            c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), old));
            (void)hipFree(old);
        }
        c->bytes -= (size_t)B.slots * n3 * sizeof(T) + (size_t)B.slots * sizeof(double);
        B.dring = nullptr;
        B.alpha = nullptr;
        B.slots = 0;
    }
    int rc;
    if ((rc = c->alloc(&B.dring, (size_t)slots * n3)) || (rc = c->alloc(&B.alpha, (size_t)slots))) return rc;
    B.slots = slots;
    return 0;
}

int precond(mm_ctx *c, const double *m_ref, const double *gF, const double *g, int max_iter,
            double *zout, bool zout_off_zero = false) {
    Launch L = c->L();
    const int64_t n = c->N, n3 = 3 * c->N;
    if (c->structured) {   // on the magnet block only (FP32 vectors in the mixed form)
        const int rc = c->precision == MM_PREC_MIXED ? tcg_ring(c, c->tb32, max_iter + 1)
                                                     : tcg_ring(c, c->tb64, max_iter + 1);
        if (rc) return rc;
        if (c->precision == MM_PREC_MIXED)
            tcgc_run<float>(L, c->G, n, c->tb32, m_ref, gF, g, c->Dex, max_iter, zout, !zout_off_zero);
        else
            tcgc_run<double>(L, c->G, n, c->tb64, m_ref, gF, g, c->Dex, max_iter, zout, !zout_off_zero);
        return 0;
    }
    launch_tcg_init(L, n3, g, c->Dex, c->tp, c->tr, c->tz, c->td);
    for (int i = 0; i < max_iter; ++i) {
        op_C(c, c->td, c->tCd, nullptr, &c->dsc->tcg_done);
        launch_hvec(L, n, m_ref, gF, c->tCd, c->td, c->tHd, &c->dsc->tcg_den, &c->dsc->tcg_done);
        launch_tcg_update(L, n3, c->tp, c->tr, c->tz, c->td, c->tHd, c->Dex);
        launch_tcg_dupdate(L, n3, c->td, c->tz);
    }
    launch_reduce(L, RED_NORMINF, n3, c->tp, nullptr, nullptr, nullptr, 0, 0, 0, &c->dsc->tcg_pmax);
    launch_tcg_final(L, n3, c->tp, g, zout);
    return 0;
}

int reduce(mm_ctx *c, RedOp op, int64_t n, const double *a, const double *b, const double *cc,
           const double *d, double *out, int nout, double h0 = 0, double h1 = 0, double h2 = 0) {
    launch_reduce(c->L(), op, n, a, b, cc, d, h0, h1, h2, c->dsc->red);
    int rc = sync_scalars(c);
    if (rc) return rc;
    for (int k = 0; k < nout; ++k) out[k] = c->hsc->red[k];
    return 0;
}
// reduce over the NCG's interleaved 3-vectors; on the magnet block only when the field
// step's m is zero off it (c->ncg_block: then so are the gradient, direction and iterates)
void launch_reduce3(mm_ctx *c, RedOp op, const double *a, const double *b, const double *cc, const double *d) {
    launch_reduce(c->L(), op, 3 * c->N, a, b, cc, d, 0, 0, 0, c->dsc->red, c->ncg_block ? &c->G : nullptr, 3);
}


struct PairEntry {
    int mat, col;
    double S, M, D[3];
};

// The rows of node i of the assembled C and D (PairCDev) from its slots 4 q + k
// (ascending): per magnet element q and local node k, for each local node l (node j) the
// exchange value 2 A_eff / V vol_q (grad phi_k . grad phi_l), the consistent-mass value
// -2 K / V vol_q (1 + [k == l]) / 20 and the divergence's Js vol_q grad phi_k / 4 of
// material m, summed per (m, j) in slot order.
void pair_rows(const HostOps &H, const std::vector<MatCoef> &mc, const int *slot, int nslot,
               std::vector<PairEntry> &out) {
    out.clear();
    const int64_t n = (int64_t)H.mag_vol.size();
    for (int t = 0; t < nslot; ++t) {
        const int q = slot[t] / 4, k = slot[t] % 4;
        const int m = H.mag_mat[q];
        const double vol = H.mag_vol[q];
        double G[4][3];
        for (int l = 1; l < 4; ++l)
            for (int a = 0; a < 3; ++a) G[l][a] = H.mag_G[(3 * (l - 1) + a) * n + q];
        for (int a = 0; a < 3; ++a) G[0][a] = -(G[1][a] + G[2][a] + G[3][a]);
        const double cex = mc[m].cex * vol, ck = mc[m].cK * vol / 20.0, cd = 0.25 * mc[m].Js * vol;
        for (int l = 0; l < 4; ++l)
            out.push_back({m, H.mag_tets[4 * q + l],
                           cex * (G[k][0] * G[l][0] + G[k][1] * G[l][1] + G[k][2] * G[l][2]),
                           ck * (k == l ? 2.0 : 1.0),
                           {cd * G[k][0], cd * G[k][1], cd * G[k][2]}});
    }
    std::stable_sort(out.begin(), out.end(), [](const PairEntry &a, const PairEntry &b) {
        return a.mat != b.mat ? a.mat < b.mat : a.col < b.col;
    });
    size_t w = 0;
    for (size_t t = 0; t < out.size(); ++t) {
        if (w > 0 && out[w - 1].mat == out[t].mat && out[w - 1].col == out[t].col) {
            out[w - 1].S += out[t].S;
            out[w - 1].M += out[t].M;
            for (int a = 0; a < 3; ++a) out[w - 1].D[a] += out[t].D[a];
        } else {
            out[w++] = out[t];
        }
    }
    out.resize(w);
}

// Assembled C of a general context (two passes over the nodes: sizes, then the rows).
int build_pair_c(mm_ctx *c, const HostOps &H, const std::vector<MatCoef> &mc, const std::vector<int> &iptr,
                 const std::vector<int> &islot) {
    const int64_t N = c->N;
    std::vector<int> nrows(N, 0), nent(N, 0);
#pragma omp parallel
    {
        std::vector<PairEntry> buf;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < N; ++i) {
            pair_rows(H, mc, islot.data() + iptr[i], iptr[i + 1] - iptr[i], buf);
            nent[i] = (int)buf.size();
            for (size_t t = 0; t < buf.size(); ++t) nrows[i] += t == 0 || buf[t].mat != buf[t - 1].mat;
        }
    }
    std::vector<int> node_ptr(N + 1, 0);
    std::vector<int64_t> ent0(N + 1, 0);
    for (int64_t i = 0; i < N; ++i) {
        node_ptr[i + 1] = node_ptr[i] + nrows[i];
        ent0[i + 1] = ent0[i] + nent[i];
    }
    if (ent0[N] >= INT32_MAX) return fail(MM_E_STATE, "assembled C: more than 2^31 entries");
    const int R = node_ptr[N];
    std::vector<int> row_mat(R), row_ptr(R + 1), col(ent0[N]);
    std::vector<double2> sm(ent0[N]);
    std::vector<double> d3(3 * ent0[N]);
    row_ptr[R] = (int)ent0[N];
#pragma omp parallel
    {
        std::vector<PairEntry> buf;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < N; ++i) {
            pair_rows(H, mc, islot.data() + iptr[i], iptr[i + 1] - iptr[i], buf);
            int r = node_ptr[i];
            for (size_t t = 0; t < buf.size(); ++t) {
                const int64_t e = ent0[i] + (int64_t)t;
                if (t == 0 || buf[t].mat != buf[t - 1].mat) {
                    row_mat[r] = buf[t].mat;
                    row_ptr[r++] = (int)e;
                }
                col[e] = buf[t].col;
                sm[e] = make_double2(buf[t].S, buf[t].M);
                for (int a = 0; a < 3; ++a) d3[3 * e + a] = buf[t].D[a];
            }
        }
    }
    int *dnp, *drm, *drp, *dcol;
    double2 *dsm;
    double *dd3;
    int rc;
    if ((rc = c->upload(&dnp, node_ptr)) || (rc = c->upload(&drm, row_mat)) || (rc = c->upload(&drp, row_ptr)) ||
        (rc = c->upload(&dcol, col)) || (rc = c->upload(&dsm, sm)) || (rc = c->upload(&dd3, d3)))
        return rc;
    c->PC.nn = N;
    c->PC.node_ptr = dnp;
    c->PC.row_mat = drm;
    c->PC.row_ptr = drp;
    c->PC.col = dcol;
    c->PC.sm = dsm;
    c->PC.d3 = dd3;
    c->pc_entries = ent0[N];
    return 0;
}

std::vector<MatCoef> material_table(const mm_mesh_desc *d, double V) {
    const double mu0 = 4.0e-7 * M_PI, size = d->length_scale > 0 ? d->length_scale : 1e-9;
    std::vector<MatCoef> mc(d->n_mats + 1);
    for (int m = 0; m < d->n_mats; ++m) {
        const bool mag = d->mat_Js[m] > 0;
        const double Aeff = mag ? mu0 * d->mat_A[m] / (size * size) : 0.0;
        const double Keff = mag ? mu0 * d->mat_K1[m] : 0.0;
        mc[m].cex = 2.0 * Aeff / V;
        mc[m].cK = -2.0 * Keff / V;
        mc[m].Js = mag ? d->mat_Js[m] : 0.0;
        for (int a = 0; a < 3; ++a) mc[m].u[a] = mag ? d->mat_u[3 * m + a] : 0.0;
    }
    mc[d->n_mats] = MatCoef{0.0, 0.0, 0.0, {0.0, 0.0, 0.0}};  // voxels outside the magnet block
    return mc;
}

template <typename T>
std::vector<T> convert(const std::vector<double> &v) {
    return std::vector<T>(v.begin(), v.end());
}

}  // namespace

namespace {
// Coarse node set of an axis and its transfers.  Spacing-aware coarsening: walking the
// intervals from the first node, two adjacent intervals merge (their middle node drops)
// when both are at most kMgMerge * hmin, hmin the smallest spacing of the level over all
// three axes; wider intervals stay.  On a graded air box the fine cells coarsen first and
// the coarse levels approach uniform spacing, instead of stretching the already wide
// cells (which point smoothers cannot handle).  Transfers: per fine index the coarse
// interval (c0, weight t of c0 + 1); per coarse index up to 3 (fine index, weight) pairs
// of the transposed interpolation (at most one dropped node on each side).
struct AxisTransfer {
    std::vector<double> coarse;
    std::vector<int> pc, ri;
    std::vector<float> pt, rw;
};
constexpr double kMgMerge = 3.0;
constexpr int kMgCoarsest = 16;   // stop once every axis has at most this many nodes
AxisTransfer coarsen_axis(const std::vector<double> &xf, double hmin) {
    AxisTransfer t;
    const int nf = (int)xf.size();
    std::vector<int> keep{0};
    for (int i = 0; i < nf - 1;) {
        const bool merge = i + 2 <= nf - 1 && xf[i + 1] - xf[i] <= kMgMerge * hmin &&
                           xf[i + 2] - xf[i + 1] <= kMgMerge * hmin;
        i += merge ? 2 : 1;
        keep.push_back(i);
    }
    const int nc = (int)keep.size();
    for (int k : keep) t.coarse.push_back(xf[k]);
    t.pc.assign(nf, 0);
    t.pt.assign(nf, 0.f);
    t.ri.assign(3 * nc, 0);
    t.rw.assign(3 * nc, 0.f);
    std::vector<int> fill(nc, 0);
    int c0 = 0;
    for (int i = 0; i < nf; ++i) {
        while (c0 + 1 < nc && keep[c0 + 1] <= i) ++c0;
        const double tt = keep[c0] == i ? 0.0 : (xf[i] - t.coarse[c0]) / (t.coarse[c0 + 1] - t.coarse[c0]);
        t.pc[i] = c0;
        t.pt[i] = (float)tt;
        // restriction: fine i feeds coarse c0 (1 - t) and c0 + 1 (t)
        const double w[2] = {1.0 - tt, tt};
        for (int q = 0; q < 2; ++q) {
            if (w[q] == 0.0) continue;
            const int C = c0 + q;
            if (fill[C] == 3) {   // cannot happen with merges of two intervals
                t.coarse.clear();
                return t;
            }
            t.ri[3 * C + fill[C]] = i;
            t.rw[3 * C + fill[C]] = (float)w[q];
            ++fill[C];
        }
    }
    return t;
}
}  // namespace

extern "C" {

const char *mm_last_error(void) { return g_err.c_str(); }

int mm_ctx_create(const mm_mesh_desc *d, const mm_options *opts, mm_ctx **out) {
    if (!d || !out) return fail(MM_E_ARG, "null argument");
    *out = nullptr;
    if (d->n_nodes <= 0 || d->n_elems <= 0 || !d->coords || !d->tets || !d->elem_mat ||
        d->n_mats <= 0 || !d->mat_Js || !d->mat_A || !d->mat_K1 || !d->mat_u)
        return fail(MM_E_ARG, "incomplete mesh description");
    mm_options o{};
    if (opts) o = *opts;
    if (o.precision != MM_PREC_FP64 && o.precision != MM_PREC_MIXED)
        return fail(MM_E_ARG, "unknown precision");
    if (o.eval_precision != MM_PREC_FP64 && o.eval_precision != MM_PREC_MIXED)
        return fail(MM_E_ARG, "unknown evaluation precision");
    for (int64_t i = 0; i < 4 * (int64_t)d->n_elems; ++i)
        if (d->tets[i] < 0 || d->tets[i] >= d->n_nodes) return fail(MM_E_MESH, "element node index out of range");
    for (int64_t e = 0; e < d->n_elems; ++e)
        if (d->elem_mat[e] < 0 || d->elem_mat[e] >= d->n_mats)
            return fail(MM_E_MESH, "element material index out of range");

    const bool structured = o.use_structured != 0;
    HostOps H;
    StructHost S;
    if (structured) {
        std::string err = structured_setup(*d, S);
        if (!err.empty()) return fail(MM_E_MESH, err);
        if (!(S.V > 0.0)) return fail(MM_E_MESH, "no magnetic element (all Js <= 0)");
        if (!ka_grid_fits(d->grid_n[0], d->grid_n[1], d->grid_n[2]))
            return fail(MM_E_MESH, "structured: grid too large for the PCG tiles; use use_structured = 0");
        if (!tcg_tiles_fit(S.mn))
            return fail(MM_E_MESH, "structured: magnet block too large for the truncated-CG tiles; use use_structured = 0");
    } else {
        std::string err = assemble(*d, o.drop_tol, H, true);
        if (!err.empty()) return fail(MM_E_MESH, err);
    }

    HIPC(hipSetDevice(o.device));
    mm_ctx *c = new mm_ctx();
    c->device = o.device;
    c->precision = o.precision;
    c->eval32 = o.precision == MM_PREC_MIXED && o.eval_precision == MM_PREC_MIXED;
    c->structured = structured;
    c->N = d->n_nodes;
    c->NE = d->n_elems;
    c->V = structured ? S.V : H.V;
    c->Nm = structured ? S.n_mag_nodes : H.n_mag_nodes;
    if (o.pcg_batch > 0) c->pcg_batch = o.pcg_batch;
    if (o.inner_tol > 0) {
        c->ru_delta = o.inner_tol;
        c->ru_delta_set = true;
    }
    int rc = 0;
    auto bail = [&](int code) {
        delete c;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(MM_E_HIP, "hipStreamCreate failed"));

    const std::vector<double> &meas = structured ? S.meas : H.meas;
    const std::vector<double> &Dex = structured ? S.Dex : H.Dex;
    const std::vector<double> &diagK = structured ? S.diagK : H.diagK;
    if ((rc = c->upload(&c->meas, meas))) return bail(rc);
    if ((rc = c->upload(&c->Dex, Dex))) return bail(rc);
    {
        std::vector<double> inv(c->N);
        for (int64_t i = 0; i < c->N; ++i) inv[i] = 1.0 / diagK[i];
        if ((rc = c->upload(&c->invDK, inv))) return bail(rc);
        if ((rc = c->upload(&c->invDK32, convert<float>(inv)))) return bail(rc);
    }
    {
        std::vector<MatCoef> mc = material_table(d, c->V);
        if ((rc = c->upload(&c->mats, mc))) return bail(rc);
        c->n_mats = d->n_mats;
    }
    if (structured) {
        // 7-point Poisson stencil: voxel sizes padded with zeros and 1 / (6 h), FP64 and FP32
        auto up = [&](const std::vector<double> &h, const double **ph, const double **pinv,
                      const float **ph32, const float **pinv32) -> int {
            std::vector<double> hp(h.size() + 2, 0.0), inv(h.size() + 2, 0.0);
            for (size_t i = 0; i < h.size(); ++i) {
                hp[i + 1] = h[i];
                inv[i + 1] = 1.0 / (6.0 * h[i]);
            }
            double *a, *b;
            float *a32, *b32;
            int r_;
            if ((r_ = c->upload(&a, hp)) || (r_ = c->upload(&b, inv)) || (r_ = c->upload(&a32, convert<float>(hp))) ||
                (r_ = c->upload(&b32, convert<float>(inv))))
                return r_;
            *ph = a;
            *pinv = b;
            *ph32 = a32;
            *pinv32 = b32;
            return 0;
        };
        c->P.nx = c->P32.nx = S.nx;
        c->P.ny = c->P32.ny = S.ny;
        c->P.nz = c->P32.nz = S.nz;
        if ((rc = up(S.hx, &c->P.hxp, &c->P.ixp, &c->P32.hxp, &c->P32.ixp)) ||
            (rc = up(S.hy, &c->P.hyp, &c->P.iyp, &c->P32.hyp, &c->P32.iyp)) ||
            (rc = up(S.hz, &c->P.hzp, &c->P.izp, &c->P32.hzp, &c->P32.izp)))
            return bail(rc);
        uint16_t *vm;
        if ((rc = c->upload(&vm, S.vmat))) return bail(rc);
        c->G.nx = S.nx;
        c->G.ny = S.ny;
        c->G.m0x = S.m0x;
        c->G.m0y = S.m0y;
        c->G.m0z = S.m0z;
        c->G.mn = S.mn;
        c->G.n_mats = d->n_mats;
        c->G.vmat = vm;
        c->G.fS = S.h / 6.0;
        c->G.fM = S.h * S.h * S.h / 120.0;
        c->G.fB = S.h * S.h / 24.0;
        c->G.fG = S.h * S.h / (24.0 * c->V);
        c->NEm = 6 * (int64_t)S.mn * S.mn * S.mn;
        // the learned warm start's view of the grid
        {
            std::vector<double> ax(d->grid_x, d->grid_x + S.nx), ay(d->grid_y, d->grid_y + S.ny),
                az(d->grid_z, d->grid_z + S.nz);
            c->axis[0] = ax;
            c->axis[1] = ay;
            c->axis[2] = az;
            double *dx, *dy, *dz;
            if ((rc = c->upload(&dx, ax)) || (rc = c->upload(&dy, ay)) || (rc = c->upload(&dz, az))) return bail(rc);
            double half = 0.0, jsmax = 0.0;
            for (double v : ax) half = std::max(half, std::fabs(v));
            for (double v : ay) half = std::max(half, std::fabs(v));
            for (double v : az) half = std::max(half, std::fabs(v));
            for (int q = 0; q < d->n_mats; ++q) jsmax = std::max(jsmax, d->mat_Js[q]);
            FeatGrid &F = c->F;
            F.nx = S.nx;
            F.ny = S.ny;
            F.nz = S.nz;
            F.x = dx;
            F.y = dy;
            F.z = dz;
            F.inv_half = half > 0 ? 1.0 / half : 0.0;
            F.inv_V = 1.0 / c->V;
            F.inv_js = jsmax > 0 ? 1.0 / jsmax : 0.0;
            F.m0x = S.m0x;
            F.m0y = S.m0y;
            F.m0z = S.m0z;
            F.mn = S.mn;
            F.vmat = vm;
            if ((rc = c->alloc(&c->ws_red, 8))) return bail(rc);
        }
    } else {
        {   // K on the host for the AMG set-up, columns ascending within each row
            CsrHost &K = c->hK;
            K.n = K.m = (int)c->N;
            K.ptr = H.k_rowptr;
            K.col = H.k_col;
            K.val = H.k_val;
            std::vector<std::pair<int, double>> row;
            for (int i = 0; i < K.n; ++i) {
                row.clear();
                for (int q = K.ptr[i]; q < K.ptr[i + 1]; ++q) row.emplace_back(K.col[q], K.val[q]);
                std::sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
                for (size_t t = 0; t < row.size(); ++t) {
                    K.col[K.ptr[i] + t] = row[t].first;
                    K.val[K.ptr[i] + t] = row[t].second;
                }
            }
        }
        std::vector<int> sp, scol;
        std::vector<double> sval;
        csr_to_sell((int)c->N, H.k_rowptr, H.k_col, H.k_val, sp, scol, sval);
        int *d_sp, *d_col;
        double *d_val;
        float *d_val32;
        if ((rc = c->upload(&d_sp, sp)) || (rc = c->upload(&d_col, scol)) || (rc = c->upload(&d_val, sval)) ||
            (rc = c->upload(&d_val32, convert<float>(sval))))
            return bail(rc);
        c->K.n_rows = c->K32.n_rows = (int)c->N;
        c->K.n_slices = c->K32.n_slices = (int)sp.size() - 1;
        c->K.slice_ptr = c->K32.slice_ptr = d_sp;
        c->K.col = c->K32.col = d_col;
        c->K.val = d_val;
        c->K32.val = d_val32;
        c->k_nnz = (int64_t)H.k_val.size();
        c->k_entries = (int64_t)sval.size();
        c->NEm = (int64_t)H.mag_elems.size();
        std::vector<int4> t4(c->NEm);
        for (int64_t q = 0; q < c->NEm; ++q)
            t4[q] = make_int4(H.mag_tets[4 * q], H.mag_tets[4 * q + 1], H.mag_tets[4 * q + 2], H.mag_tets[4 * q + 3]);
        int4 *dt;
        double *dG, *dv;
        int *dm;
        if ((rc = c->upload(&dt, t4)) || (rc = c->upload(&dG, H.mag_G)) || (rc = c->upload(&dv, H.mag_vol)) ||
            (rc = c->upload(&dm, H.mag_mat)))
            return bail(rc);
        c->E.n = (int)c->NEm;
        c->E.tets = dt;
        c->E.G = dG;
        c->E.vol = dv;
        c->E.mat = dm;
        {   // node -> slots 4 q + k of its magnet elements, ascending (a fixed summation order)
            std::vector<int> iptr(c->N + 1, 0), islot(4 * (size_t)c->NEm);
            for (size_t s4 = 0; s4 < islot.size(); ++s4) iptr[H.mag_tets[s4] + 1]++;
            for (int64_t i = 0; i < c->N; ++i) iptr[i + 1] += iptr[i];
            std::vector<int> fill(iptr.begin(), iptr.end() - 1);
            for (size_t s4 = 0; s4 < islot.size(); ++s4) islot[fill[H.mag_tets[s4]]++] = (int)s4;
            int *dp, *dsl;
            double *dcy;
            if ((rc = c->upload(&dp, iptr)) || (rc = c->upload(&dsl, islot)) ||
                (rc = c->alloc(&dcy, 3 * (size_t)c->NEm)))
                return bail(rc);
            c->E.nn = c->N;
            c->E.inc_ptr = dp;
            c->E.inc_slot = dsl;
            c->E.cy = dcy;
            if ((rc = build_pair_c(c, H, material_table(d, c->V), iptr, islot))) return bail(rc);
        }
    }
    // algorithmic bytes of one KA launch of the PCG loop: z and p_old read once, p and
    // q = K p written once, plus the stored matrix for SELL (entries: value + column,
    // slice pointers); the stencil stores no matrix.
    const int s_k = c->precision == MM_PREC_MIXED ? 4 : 8;
    c->t_spmv.bytes = structured ? 4 * (int64_t)s_k * c->N
                                 : c->k_nnz * (s_k + 4) + 4 * (c->N + 1) + 4 * (int64_t)s_k * c->N;
    // PCG update: reads x, r, p, q, invD; writes x, r, z
    c->t_upd.bytes = 8 * (int64_t)s_k * c->N;
    c->t_mg.bytes = 3 * (int64_t)sizeof(float) * c->N;   // b, x read; the new x written

    if ((rc = c->alloc(&c->dsc, 1))) return bail(rc);
    {
        DevScalars init{};
        init.pcg_done = 1;
        init.tcg_done = 1;
        init.ir_done = 1;
        HIPC(hipMemcpy(c->dsc, &init, sizeof(DevScalars), hipMemcpyHostToDevice));
    }
    {   // the mirror and, on its own 64-byte line after it, the sequence number; coherent, so
        // that the device's writes bypass its caches
        constexpr size_t kMirror = (sizeof(DevScalars) + 63) / 64 * 64;
        void *hm = nullptr;
        if (hipHostMalloc(&hm, kMirror + 64, hipHostMallocCoherent) != hipSuccess)
            return bail(fail(MM_E_HIP, "hipHostMalloc failed"));
        std::memset(hm, 0, kMirror + 64);
        c->hsc = (DevScalars *)hm;
        c->hseq = (unsigned long long *)((char *)hm + kMirror);
    }
    if ((rc = c->alloc(&c->red.partials, 4 * (size_t)kMaxPartials))) return bail(rc);

    const size_t n = c->N, n3 = 3 * n;
    if ((rc = c->alloc(&c->partsA, 3 * (size_t)kMaxPartials)) || (rc = c->alloc(&c->partsB, 3 * (size_t)kMaxPartials)) ||
        (rc = c->alloc(&c->ncg_parts, NCG_NV * (size_t)kMaxPartials)))
        return bail(rc);
    double **vN[] = {&c->b, &c->r, &c->z, &c->p, &c->p2, &c->q};
    for (double **v : vN)
        if ((rc = c->alloc(v, n))) return bail(rc);
    if (c->precision == MM_PREC_MIXED) {
        float **vF[] = {&c->r32, &c->z32, &c->p32, &c->p32b, &c->q32, &c->dx32};
        for (float **v : vF)
            if ((rc = c->alloc(v, n))) return bail(rc);
    }
    double **v3[] = {&c->gC, &c->gd, &c->X, &c->X0, &c->X1, &c->Gp, &c->G1, &c->Dd, &c->Z1, &c->gtmp};
    for (double **v : v3)
        if ((rc = c->alloc(v, n3))) return bail(rc);
    if (structured) {   // truncated CG on the magnet block
        const size_t n1 = (size_t)c->G.mn + 1, nb = n1 * n1 * n1;
        auto block = [&](auto &B) -> int {
            int r_;
            if ((r_ = c->alloc(&B.mg, nb))) return r_;
            for (auto **v : {&B.m0, &B.D, &B.p, &B.r, &B.z, &B.Hd})
                if ((r_ = c->alloc(v, 3 * nb))) return r_;
            return tcg_ring(c, B, kTcgSlots);
        };
        if ((rc = c->precision == MM_PREC_MIXED ? block(c->tb32) : block(c->tb64))) return bail(rc);
    } else {
        double **vt[] = {&c->tp, &c->tr, &c->tz, &c->td, &c->tCd, &c->tHd};
        for (double **v : vt)
            if ((rc = c->alloc(v, n3))) return bail(rc);
    }
    HIPC(hipDeviceSynchronize());
    *out = c;
    return 0;
}

int mm_ctx_destroy(mm_ctx *c) {
    if (!c) return fail(MM_E_ARG, "null context");
    delete c;
    return 0;
}

int mm_ctx_info(mm_ctx *c, mm_info *o) {
    if (!c || !o) return fail(MM_E_ARG, "null argument");
    o->n_nodes = c->N;
    o->n_elems = c->NE;
    o->n_mag_elems = c->NEm;
    o->n_mag_nodes = c->Nm;
    o->k_nnz = c->k_nnz;
    o->k_sell_entries = c->k_entries;
    o->volume = c->V;
    o->structured = c->structured ? 1 : 0;
    o->precision = c->precision;
    o->device_bytes = (int64_t)c->bytes;
    return 0;
}

int mm_get_meas_host(mm_ctx *c, double *meas_host) {
    if (!c || !meas_host) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemcpy(meas_host, c->meas, c->N * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int mm_get_diag_host(mm_ctx *c, double *dex_host, double *diagK_host) {
    if (!c) return fail(MM_E_ARG, "null argument");
    if (dex_host) HIPC(hipMemcpy(dex_host, c->Dex, 3 * c->N * sizeof(double), hipMemcpyDeviceToHost));
    if (diagK_host) {
        std::vector<double> inv(c->N);
        HIPC(hipMemcpy(inv.data(), c->invDK, c->N * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < c->N; ++i) diagK_host[i] = 1.0 / inv[i];
    }
    return 0;
}

int mm_external_eg(mm_ctx *c, const double *m, double hext, const double hdir[3], double *E_host,
                   double *g) {
    if (!c || !m || !hdir || !E_host) return fail(MM_E_ARG, "null argument");
    launch_external(c->L(), c->N, m, c->meas, -hext * hdir[0] / c->V, -hext * hdir[1] / c->V,
                    -hext * hdir[2] / c->V, g, &c->dsc->energy);
    int rc = sync_scalars(c);
    if (rc) return rc;
    *E_host = c->hsc->energy;
    return 0;
}

int mm_exani_eg(mm_ctx *c, const double *m, double *E_host, double *g) {
    if (!c || !m || !E_host || !g) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemsetAsync(g, 0, 3 * (size_t)c->N * sizeof(double), c->s));
    op_C(c, m, g, nullptr, nullptr);
    launch_reduce(c->L(), RED_DOT, 3 * c->N, m, g, nullptr, nullptr, 0, 0, 0, c->dsc->red);
    int rc = sync_scalars(c);
    if (rc) return rc;
    *E_host = 0.5 * c->hsc->red[0];
    return 0;
}

int mm_hmag_eg(mm_ctx *c, const double *m, double *u, double tol, double *E_host, double *g,
               int32_t *pcg_iters_host) {
    if (!c || !m || !u || !E_host || !g) return fail(MM_E_ARG, "null argument");
    op_C(c, m, c->gC, c->b, nullptr);
    int it = 0;
    int rc = potential_solve(c, m, u, tol, &it);
    if (rc < 0) return rc;
    HIPC(hipMemsetAsync(g, 0, 3 * (size_t)c->N * sizeof(double), c->s));
    op_grad(c, u, g);
    launch_reduce(c->L(), RED_DOT, 3 * c->N, m, g, nullptr, nullptr, 0, 0, 0, c->dsc->red);
    int rc2 = sync_scalars(c);
    if (rc2) return rc2;
    *E_host = 0.5 * c->hsc->red[0];
    if (pcg_iters_host) *pcg_iters_host = it;
    return rc;
}

int mm_total_eg(mm_ctx *c, const double *m, double hext, const double hdir[3], double tol_u,
                int32_t hmag_on, double *u, double *E_host, double *g_proj, double *g_full,
                int32_t *pcg_iters_host) {
    if (!c || !m || !hdir || !u || !E_host || !g_proj || !g_full) return fail(MM_E_ARG, "null argument");
    int it = 0;
    int rc = total_eg(c, m, hext, hdir, tol_u, hmag_on, u, E_host, g_proj, g_full, &it);
    if (pcg_iters_host) *pcg_iters_host = it;
    return rc;
}

int mm_pcg(mm_ctx *c, const double *b, double *x, double tol, double atol, int32_t maxiter,
           int32_t *iters_host) {
    if (!c || !b || !x) return fail(MM_E_ARG, "null argument");
    if (maxiter < 0) return fail(MM_E_ARG, "maxiter < 0");
    int it = 0;
    int rc = pcg(c, b, x, tol, atol, maxiter, &it);
    if (iters_host) *iters_host = it;
    return rc;
}

int mm_precond(mm_ctx *c, const double *m_ref, const double *g_full, const double *g,
               int32_t max_iter, double *z, int32_t *iters_host) {
    if (!c || !m_ref || !g_full || !g || !z) return fail(MM_E_ARG, "null argument");
    if (max_iter < 0) return fail(MM_E_ARG, "max_iter < 0");
    int rc = precond(c, m_ref, g_full, g, max_iter, z);
    if (rc) return rc;
    rc = sync_scalars(c);
    if (rc) return rc;
    if (iters_host) *iters_host = max_iter > 0 ? c->hsc->tcg_iter : 0;
    return 0;
}

int mm_update_m(mm_ctx *c, const double *x, double alpha, const double *d, double *out) {
    if (!c || !x || !d || !out) return fail(MM_E_ARG, "null argument");
    launch_update_m(c->L(), c->N, x, alpha, d, out);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_dot(mm_ctx *c, const double *a, const double *b, int64_t n, double *out) {
    if (!c || !a || !b || !out || n < 0) return fail(MM_E_ARG, "bad argument");
    return reduce(c, RED_DOT, n, a, b, nullptr, nullptr, out, 1);
}

// ---------------------------------------------------------------- mm_dot_noalloc
// mm_dot for callers that cannot allocate (no context): its reduction keeps the per-block
// partials and the result in this file-scope device array instead of a context's
// workspace.  k_reduce runs at most kMaxPartials blocks (one partial each), so every
// n >= 0 fits; the array is 2080 doubles (16.3 KB) per device, held until the process
// ends.  The array's device address is looked up once per device (hipGetSymbolAddress,
// with that device current) and only that address is handed to kernels and copies.
static __device__ __align__(256) double scratch[kMaxPartials + 32];

namespace {
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocResult = kMaxPartials;   // result slot after the partials
std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device (set-up on its first use).
int scratch_ptr(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return 0;
}
}  // namespace

int mm_dot_noalloc(int32_t device, const double *a, const double *b, int64_t n, double *out_host) {
    if (!out_host || n < 0 || (n > 0 && (!a || !b))) return fail(MM_E_ARG, "bad argument");
    int count = 0;
    HIPC(hipGetDeviceCount(&count));
    if (device < 0 || device >= count || device >= kNoAllocMaxDevices)
        return fail(MM_E_ARG, "device ordinal out of range");
    int prev = 0;
    HIPC(hipGetDevice(&prev));
    if (prev != device) HIPC(hipSetDevice(device));
    double *ws = nullptr;
    int rc = scratch_ptr(device, &ws);
    hipError_t e = hipSuccess;
    if (!rc) {
        Launch L{};
        L.s = nullptr;   // the device's null stream: ordered after the caller's default-stream work
        L.red.partials = ws;
        launch_reduce(L, RED_DOT, n, a, b, nullptr, nullptr, 0, 0, 0, ws + kNoAllocResult);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(out_host, ws + kNoAllocResult, sizeof(double), hipMemcpyDeviceToHost);
    }
    if (prev != device) hipSetDevice(prev);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MM_E_HIP, std::string("mm_dot_noalloc: ") + hipGetErrorString(e));
    return 0;
}

int mm_norm_inf(mm_ctx *c, const double *a, int64_t n, double *out) {
    if (!c || !a || !out || n < 0) return fail(MM_E_ARG, "bad argument");
    return reduce(c, RED_NORMINF, n, a, nullptr, nullptr, nullptr, out, 1);
}

int mm_compute_mh(mm_ctx *c, const double *m, const double hdir[3], double *mh) {
    if (!c || !m || !hdir || !mh) return fail(MM_E_ARG, "null argument");
    double s = 0;
    int rc = reduce(c, RED_MH, c->N, m, c->meas, nullptr, nullptr, &s, 1, hdir[0], hdir[1], hdir[2]);
    if (rc) return rc;
    *mh = s / c->V;
    return 0;
}

int mm_apply_C(mm_ctx *c, const double *v, double *y) {
    if (!c || !v || !y) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemsetAsync(y, 0, 3 * (size_t)c->N * sizeof(double), c->s));
    op_C(c, v, y, nullptr, nullptr);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_apply_K(mm_ctx *c, const double *x, double *y) {
    if (!c || !x || !y) return fail(MM_E_ARG, "null argument");
    op_K(c, x, y, nullptr, nullptr);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_divergence(mm_ctx *c, const double *m, double *b) {
    if (!c || !m || !b) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemsetAsync(b, 0, (size_t)c->N * sizeof(double), c->s));
    op_C(c, m, c->gtmp, b, nullptr);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_gradient(mm_ctx *c, const double *u, double *g) {
    if (!c || !u || !g) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemsetAsync(g, 0, 3 * (size_t)c->N * sizeof(double), c->s));
    op_grad(c, u, g);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_minimize(mm_ctx *c, double *m_inout, double hext, const double hdir[3],
                const mm_min_params *P, const double *m_ref, double *u, double *gF, double *E_host,
                mm_stats *st) {
    if (!c || !m_inout || !hdir || !P || !m_ref || !u || !gF || !E_host)
        return fail(MM_E_ARG, "null argument");
    const int64_t n = c->N;
    const int64_t n3 = 3 * n;
    const size_t bytes3 = n3 * sizeof(double);
    Launch L = c->L();
    const double tol = P->tol_fun;
    const int iter_max = P->iter_max > 0 ? P->iter_max : 100000;
    mm_stats loc{};
    int rc = 0;
    double *X = c->X, *X0 = c->X0, *X1 = c->X1, *G = c->Gp, *G1 = c->G1, *D = c->Dd, *Z1 = c->Z1;
    c->sec_next = false;
    HIPC(hipMemcpyAsync(X, m_inout, bytes3, hipMemcpyDeviceToDevice, c->s));
    if (c->structured) launch_outside_max(L, n, c->G, X, &c->dsc->red[0]);
    if ((rc = sync_scalars(c))) return rc;
    c->ncg_block = c->structured && c->hsc->red[0] == 0.0;
    const long long tcg0 = c->hsc->tcg_total;
    // When m is zero off the magnet block (ncg_block), so are every iterate, gradient and
    // direction of this call: each buffer's entries off the block are written (zero) the
    // first time this call writes it, and after that only its block entries.
    const double *zeroed[8] = {};
    int nzeroed = 0;
    auto off_zero = [&](const double *p) {
        if (!c->ncg_block) return false;
        for (int q = 0; q < nzeroed; ++q)
            if (zeroed[q] == p) return true;
        if (nzeroed < 8) zeroed[nzeroed++] = p;   // written in full by the launch that asked
        return false;
    };
    auto blk = [&](const double *out) { return off_zero(out) ? &c->G : nullptr; };

    auto feval = [&](const double *x, double *E, double *g, const PostReduce *post = nullptr) -> int {
        int it = 0;
        const bool oz_g = off_zero(g), oz_f = off_zero(gF);
        int r = total_eg(c, x, hext, hdir, P->tol_u, P->hmag_on, u, E, g, gF, &it, oz_g && oz_f, post);
        loc.fcalls++;
        loc.pcg_iters += it;
        loc.pcg_calls += P->hmag_on > 0;
        loc.c_applies++;
        return r < 0 ? r : 0;
    };
    auto minv = [&](const double *g, double *z) -> int {
        loc.c_applies += P->precond_iter;
        return precond(c, m_ref, gF, g, P->precond_iter, z, off_zero(z));
    };
    const GridDev *gb = c->ncg_block ? &c->G : nullptr;
    // The totals of every test from the end of one iteration to the next line search
    // (the restart after the HS+ update, the stop test, the restart before the line search,
    // the line search's first step) come from one scan (k_ncg_scan) and one host round
    // trip, read at the top of the loop; the HS+ beta stays on the device.
    auto scan = [&](const double *xo, const double *xn, const double *go, const double *gn, bool upd) {
        if (upd) launch_reduce3(c, RED_YZD, go, gn, Z1, D);   // ((g1 - g).z, (g1 - g).d) into red
        launch_ncg_scan(L, n3, gb, xo, xn, go, gn, Z1, D, upd ? c->dsc->red : nullptr, c->ncg_parts, c->dsc->ncg);
    };

    const double FMAX = DBL_MAX;
    double f0 = FMAX, f = 0.0;
    if ((rc = feval(X, &f, G))) return rc;
    if ((rc = minv(G, Z1))) return rc;
    launch_axpby(L, n3, -1.0, Z1, 0.0, nullptr, D, blk(D));  // d = -M_inv(g)
    int64_t k = 1;
    off_zero(X);    // the copy of m, in full
    scan(X, X, G, G, false);   // x0 = x: the first stop test's max|x0 - x| is 0
    bool after_update = false;   // the scan's d is an HS+ update (its restart test pending)
    const double sqrt_tol = std::sqrt(tol), cbrt_tol = std::cbrt(tol);
    const double sqrt_eps = std::sqrt(DBL_EPSILON);
    for (;;) {
        if ((rc = sync_scalars(c))) return rc;
        double v[NCG_NV];
        for (int q = 0; q < NCG_NV; ++q) v[q] = c->hsc->ncg[q];
        // restart if the HS+ direction is not a sufficient descent direction
        // (minimizers.py:182-184; tested against the gradient it was formed with)
        bool rz = false;
        if (after_update && v[NCG_DG_OLD] > -0.001 * std::sqrt(v[NCG_DD]) * std::sqrt(v[NCG_GG_OLD])) {
            launch_axpby(L, n3, -1.0, Z1, 0.0, nullptr, D, blk(D));
            loc.restarts_z++;
            rz = true;
        }
        // stop test (minimizers.py:164-166): max|x0 - x|, max|x|, max|g|
        const bool a = (f0 - f) > tol * (1 + std::fabs(f));
        const bool b = (v[NCG_DX_MAX] > sqrt_tol * (1 + v[NCG_X_MAX])) || k == 0;
        const bool cc = v[NCG_G_MAX] > cbrt_tol * (1 + std::fabs(f));
        if (!((a || b || cc) && k < iter_max)) break;
        // restart if d is not a sufficient descent direction (minimizers.py:171-173):
        // d.g, d.d, g.g for d as it stands (-z after the restart above)
        const double dg3[3] = {rz ? -v[NCG_ZG] : v[NCG_DG], rz ? v[NCG_ZZ] : v[NCG_DD], v[NCG_GG]};
        bool rg = false;
        if (dg3[0] > -0.001 * std::sqrt(dg3[1]) * std::sqrt(dg3[2])) {
            launch_axpby(L, n3, -1.0, G, 0.0, nullptr, D, blk(D));
            loc.restarts_g++;
            rg = true;
        }
        // ---- line search (minimizers.py:92-137)
        double f1 = 0.0;
        {
            const double rho = 0.5, cA = 1e-4, delta = 0.1, eps = 1e-6;
            const int max_ls = 20;
            // x.x, d.d, max|d| for the final d
            const double h3[3] = {v[NCG_XX], rg ? v[NCG_GG] : rz ? v[NCG_ZZ] : v[NCG_DD],
                                  rg ? v[NCG_G_MAX] : rz ? v[NCG_Z_MAX] : v[NCG_D_MAX]};
            const double t1 = 2.0 * sqrt_eps * (1.0 + std::sqrt(h3[0])) / std::sqrt(h3[1]);
            const double t2 = sqrt_eps / h3[2];
            const double h = t1 < t2 ? t1 : t2;
            // the secant start's reference solution: u at x, before the first evaluation
            const bool sec = c->sec_on && P->hmag_on > 0;
            if (sec) HIPC(hipMemcpyAsync(c->u_sec, u, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
            launch_update_m(L, n, X, h, D, X1, blk(X1));
            double gd2[2];
            const PostReduce pgd2{RED_GD2, G, G1, D, nullptr, 2, gd2};
            if ((rc = feval(X1, &f1, G1, &pgd2))) return rc;
            const double gd = gd2[0];
            double alpha;
            {  // alpha_init (minimizers.py:34-57), c = 0.1, o = 0.99
                double term1 = -h * (gd / (gd2[1] - gd));
                if (term1 < 0) term1 = 0.1 * std::fabs(term1);
                double term2 = 2.0 * (f - f0) / gd;
                if (term2 < 0) term2 = 1.0e30;
                alpha = term1 < 0.99 * term2 ? term1 : term2;
            }
            double gd1;
            const PostReduce pdot{RED_DOT, G1, D, nullptr, nullptr, 1, &gd1};
            launch_update_m(L, n, X, alpha, D, X1, blk(X1));
            c->sec_next = sec;
            if ((rc = feval(X1, &f1, G1, &pdot))) return rc;
            int i = 0;
            for (;;) {
                const double eps_k = eps * std::fabs(f);
                const bool wolfe = ((2 * delta - 1) * gd >= gd1) && (f1 <= f + eps_k);
                const bool done = wolfe || (f1 <= f + cA * alpha * gd) || (i > max_ls);
                if (done) break;
                alpha *= rho;
                launch_update_m(L, n, X, alpha, D, X1, blk(X1));
                c->sec_next = sec;
                if ((rc = feval(X1, &f1, G1, &pdot))) return rc;
                ++i;
                loc.backtracks++;
            }
        }
        // ---- HS+ direction (minimizers.py:175-184): d = -z1 + beta d on the device, its
        // restart test at the top of the next iteration
        if ((rc = minv(G1, Z1))) return rc;
        scan(X, X1, G, G1, true);
        after_update = true;
        // shift state: x0 <- x, x <- x1, f0 <- f, f <- f1, g <- g1
        std::swap(X0, X);
        std::swap(X, X1);
        std::swap(G, G1);
        f0 = f;
        f = f1;
        ++k;
    }
    HIPC(hipMemcpyAsync(m_inout, X, bytes3, hipMemcpyDeviceToDevice, c->s));
    if ((rc = sync_scalars(c))) return rc;
    loc.ncg_iters = k;
    loc.tcg_iters = c->hsc->tcg_total - tcg0;
    *E_host = f;
    if (st) {
        st->ncg_iters += loc.ncg_iters;
        st->fcalls += loc.fcalls;
        st->pcg_iters += loc.pcg_iters;
        st->pcg_calls += loc.pcg_calls;
        st->tcg_iters += loc.tcg_iters;
        st->c_applies += loc.c_applies;
        st->backtracks += loc.backtracks;
        st->restarts_g += loc.restarts_g;
        st->restarts_z += loc.restarts_z;
    }
    // keep the rotated buffers for the next call
    c->X = X;
    c->X0 = X0;
    c->X1 = X1;
    c->Gp = G;
    c->G1 = G1;
    return 0;
}

// ---------------------------------------------------------------- multigrid


int mm_multigrid(mm_ctx *c, int32_t max_levels) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (max_levels == 0) {   // back to Jacobi, with Jacobi's reliable-update interval
        c->mg_nl = 0;
        if (!c->ru_delta_set) c->ru_delta = 0.1;
        return 0;
    }
    if (!c->structured || c->precision != MM_PREC_MIXED)
        return fail(MM_E_STATE, "the multigrid preconditioner needs a structured mixed-precision context");
    if (max_levels < 2 || max_levels > kMgMaxLevels) return fail(MM_E_ARG, "levels outside [2, 8]");
    // the levels are built once per context: a later call re-enables them (same bound) or
    // is refused, so that no set-up rewrites levels in use or allocates them again
    if (c->mg_built > 0) {
        if (max_levels != c->mg_built_max)
            return fail(MM_E_STATE, "multigrid already set up with another level bound; use a new context");
        c->mg_nl = c->mg_built;
        if (!c->ru_delta_set) c->ru_delta = 1e-3;
        return 0;
    }
    static_assert(kMgMaxLevels == MM_MG_MAX_LEVELS, "level bound of the header");
    int rc = 0;
    std::vector<double> ax[3] = {c->axis[0], c->axis[1], c->axis[2]};
    int nl = 0;
    c->mg[0].P = c->P32;
    for (int l = 0;; ++l) {
        const int64_t n = (int64_t)ax[0].size() * ax[1].size() * ax[2].size();
        if (l > 0) {
            if ((rc = c->alloc(&c->mg_b[l], (size_t)n)) || (rc = c->alloc(&c->mg_xa[l], (size_t)n)) ||
                (rc = c->alloc(&c->mg_xb[l], (size_t)n)))
                return rc;
        } else if (!c->mg_xb[0]) {
            if ((rc = c->alloc(&c->mg_xb[0], (size_t)n))) return rc;
        }
        nl = l + 1;
        const size_t mx = std::max(ax[0].size(), std::max(ax[1].size(), ax[2].size()));
        if (mx <= (size_t)kMgCoarsest || nl == max_levels) break;
        // transfers to level l + 1 and its operator
        double hmin = HUGE_VAL;
        for (int a = 0; a < 3; ++a)
            for (size_t i = 1; i < ax[a].size(); ++i) hmin = std::min(hmin, ax[a][i] - ax[a][i - 1]);
        AxisTransfer t[3] = {coarsen_axis(ax[0], hmin), coarsen_axis(ax[1], hmin), coarsen_axis(ax[2], hmin)};
        if (t[0].coarse.size() == ax[0].size() && t[1].coarse.size() == ax[1].size() &&
            t[2].coarse.size() == ax[2].size())
            break;   // nothing merges any more
        for (int a = 0; a < 3; ++a) {   // host-side bounds of every index the kernels use
            if (t[a].coarse.size() < 3) return fail(MM_E_STATE, "multigrid: degenerate coarse axis");
            const int nf = (int)ax[a].size(), nc = (int)t[a].coarse.size();
            // the prolongation reads coarse nodes c0 and c0 + 1 of every interior fine node
            for (int i = 0; i < nf; ++i)
                if (t[a].pc[i] < 0 || t[a].pc[i] >= nc ||
                    ((t[a].pt[i] != 0.f || (i > 0 && i < nf - 1)) && t[a].pc[i] + 1 >= nc))
                    return fail(MM_E_STATE, "multigrid: prolongation index out of range");
            for (int q = 0; q < 3 * nc; ++q)
                if (t[a].ri[q] < 0 || t[a].ri[q] >= nf) return fail(MM_E_STATE, "multigrid: restriction index out of range");
        }
        if (!mg_restrict_span_fits(t[0].ri.data(), t[0].rw.data(), (int)t[0].coarse.size()))
            return fail(MM_E_STATE, "multigrid: an x restriction span exceeds the restriction kernel's staged row");
        MgLevelDev &f = c->mg[l];
        int *pci[3], *rii[3];
        float *ptf[3], *rwf[3];
        for (int a = 0; a < 3; ++a)
            if ((rc = c->upload(&pci[a], t[a].pc)) || (rc = c->upload(&ptf[a], t[a].pt)) ||
                (rc = c->upload(&rii[a], t[a].ri)) || (rc = c->upload(&rwf[a], t[a].rw)))
                return rc;
        f.pxc = pci[0]; f.pyc = pci[1]; f.pzc = pci[2];
        f.pxt = ptf[0]; f.pyt = ptf[1]; f.pzt = ptf[2];
        f.rxi = rii[0]; f.ryi = rii[1]; f.rzi = rii[2];
        f.rxw = rwf[0]; f.ryw = rwf[1]; f.rzw = rwf[2];
        if ((rc = c->alloc(&c->mg_r[l], (size_t)n))) return rc;
        for (int a = 0; a < 3; ++a) ax[a] = t[a].coarse;
        PoissonGrid<float> &P = c->mg[l + 1].P;
        P.nx = (int)ax[0].size();
        P.ny = (int)ax[1].size();
        P.nz = (int)ax[2].size();
        const float **hp[3] = {&P.hxp, &P.hyp, &P.hzp}, **ip[3] = {&P.ixp, &P.iyp, &P.izp};
        for (int a = 0; a < 3; ++a) {
            std::vector<float> h(ax[a].size() + 1, 0.f), inv(ax[a].size() + 1, 0.f);
            for (size_t i = 1; i < ax[a].size(); ++i) {
                const double d = ax[a][i] - ax[a][i - 1];
                h[i] = (float)d;
                inv[i] = (float)(1.0 / (6.0 * d));
            }
            float *dh, *di;
            if ((rc = c->upload(&dh, h)) || (rc = c->upload(&di, inv))) return rc;
            *hp[a] = dh;
            *ip[a] = di;
        }
    }
    // Chebyshev-Richardson steps: reciprocals of the Chebyshev nodes of [lo, 2], from the
    // largest node down; with leja set, in greedy Leja order instead (the largest node,
    // then each time the node with the largest product of distances to those taken; ties,
    // which the symmetric nodes make, go to the lower index).  Same polynomial; in the
    // descending order the last steps of the coarsest polynomial (w up to ~100) multiply
    // the FP32 rounding of the earlier ones by ~1e3, in Leja order it stays at ~4e-7.
    auto omegas = [](float *om, int deg, double lo, bool leja) {
        const double hi = 2.0;
        double node[kMgMaxDeg];
        int order[kMgMaxDeg];
        bool used[kMgMaxDeg] = {};
        for (int k = 0; k < deg; ++k) {
            node[k] = (hi + lo) / 2 + (hi - lo) / 2 * std::cos(M_PI * (2 * k + 1) / (2.0 * deg));
            order[k] = k;
        }
        for (int t = 1; leja && t < deg; ++t) {
            used[order[t - 1]] = true;
            double best = -1;
            for (int j = 0; j < deg; ++j) {
                if (used[j]) continue;
                double p = 1;
                for (int i = 0; i < t; ++i) p *= std::fabs(node[j] - node[order[i]]);
                if (p > best * (1 + 1e-9)) best = p, order[t] = j;
            }
        }
        for (int k = 0; k < deg; ++k) om[k] = (float)(1.0 / node[order[k]]);
    };
    // the finest level smooths with degree 4 on [2/8, 2] (a sweep on the 154^3 grid: 20.8
    // PCG iterations per solve against 25 with degree 3 on [2/15, 2], +2.7% field steps/s;
    // coarser levels gain nothing from more smoothing)
    omegas(c->mg_om0, c->mg_deg0, 2.0 / 8.0, false);
    omegas(c->mg_om, c->mg_deg, 2.0 / 15.0, false);
    omegas(c->mg_omc, c->mg_cdeg, 2.0 / 100.0, true);
    // invw of every level: its first step's weight over the stencil's own diagonal (the
    // first-step kernel applied to ones), 0 on the faces
    {
        float *ones = nullptr;
        const std::vector<float> h1((size_t)c->N, 1.f);
        if ((rc = c->upload(&ones, h1))) return rc;
        Launch L = c->L();
        for (int l = 0; l < nl; ++l) {
            const PoissonGrid<float> &P = c->mg[l].P;
            if ((rc = c->alloc(&c->mg[l].invw, (size_t)P.nx * P.ny * P.nz))) return rc;
            launch_mg_first(L, P, ones, c->mg[l].invw,
                            l + 1 == nl ? c->mg_omc[0] : l == 0 ? c->mg_om0[0] : c->mg_om[0], nullptr);
        }
        HIPC(hipStreamSynchronize(c->s));
    }
    // the coarse tail: the most levels (at least one, not level 0) whose vectors fit the
    // tail kernel's LDS
    c->mg_tail = nl;
    auto tail_args = [&](int t) {
        MgTailArgs a;
        a.nt = nl - t;
        a.s = c->mg_deg;
        a.cd = c->mg_cdeg;
        for (int k = 0; k < kMgMaxDeg; ++k) {
            a.om[k] = c->mg_om[k];
            a.omc[k] = c->mg_omc[k];
        }
        for (int q = 0; q < a.nt; ++q) {
            a.lev[q] = c->mg[t + q];
            a.n[q] = c->mg[t + q].P.nx * c->mg[t + q].P.ny * c->mg[t + q].P.nz;
        }
        a.b0 = c->mg_b[t];
        a.x0 = c->mg_xa[t];
        a.count = &c->dsc->mg_vcycles;
        mg_tail_layout(a);
        return a;
    };
    for (int t = nl - 1; t >= 1 && nl - t <= kMgTailMax; --t) {
        const MgTailArgs a = tail_args(t);
        if (!mg_tail_fits(a)) break;
        c->mg_tail = t;
        c->tail = a;
    }
    if (c->mg_tail < nl) {   // the tail's stencil factors and transfer lists, formed once
        float *img = nullptr;
        if ((rc = c->alloc(&img, (size_t)c->tail.img_n))) return rc;
        launch_mg_tail_image(c->L(), c->tail, img);
        HIPC(hipStreamSynchronize(c->s));
        c->tail.img = img;
    }
    // a multigrid-preconditioned FP32 recursion stays accurate over a longer stretch:
    // replace the residual every 1e-3 of reduction unless the caller chose otherwise
    if (!c->ru_delta_set) c->ru_delta = 1e-3;
    c->mg_nl = nl;   // published last: a set-up that failed before leaves the Jacobi form in use
    c->mg_built = nl;
    c->mg_built_max = max_levels;
    return 0;
}

namespace {
// SELL-64 of a CSR matrix (rectangular allowed: padding entries take column 0), FP32
int upload_sell32(mm_ctx *c, const CsrHost &M, SellDev<float> &S) {
    const int C = kWave, ns = (M.n + C - 1) / C;
    std::vector<int> sp(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        int w = 0;
        for (int r = s * C; r < std::min(M.n, (s + 1) * C); ++r) w = std::max(w, M.ptr[r + 1] - M.ptr[r]);
        sp[s + 1] = sp[s] + w * C;
    }
    std::vector<int> col(sp[ns], 0);
    std::vector<float> val(sp[ns], 0.f);
    for (int s = 0; s < ns; ++s) {
        const int w = (sp[s + 1] - sp[s]) / C;
        for (int lane = 0; lane < C; ++lane) {
            const int r = s * C + lane;
            const int len = r < M.n ? M.ptr[r + 1] - M.ptr[r] : 0;
            for (int j = 0; j < len && j < w; ++j) {
                col[sp[s] + (size_t)j * C + lane] = M.col[M.ptr[r] + j];
                val[sp[s] + (size_t)j * C + lane] = (float)M.val[M.ptr[r] + j];
            }
        }
    }
    int *dsp, *dcol;
    float *dval;
    int rc;
    if ((rc = c->upload(&dsp, sp)) || (rc = c->upload(&dcol, col)) || (rc = c->upload(&dval, val))) return rc;
    S.n_rows = M.n;
    S.n_slices = ns;
    S.slice_ptr = dsp;
    S.col = dcol;
    S.val = dval;
    return 0;
}

// CSR (FP32) with lanes per row from the mean row length: about four entries per lane
int upload_csr32(mm_ctx *c, const CsrHost &M, CsrDevF &S) {
    const double mean = M.n > 0 ? (double)M.ptr[M.n] / M.n : 0.0;
    int g = 4;
    while (g < kWave && 4.0 * g < mean) g *= 2;
    std::vector<float> val(M.val.begin(), M.val.end());
    int *dptr, *dcol;
    float *dval;
    int rc;
    if ((rc = c->upload(&dptr, M.ptr)) || (rc = c->upload(&dcol, M.col)) || (rc = c->upload(&dval, val))) return rc;
    S.n_rows = M.n;
    S.group = g;
    S.ptr = dptr;
    S.col = dcol;
    S.val = dval;
    return 0;
}
}  // namespace

int mm_amg(mm_ctx *c, int32_t max_levels) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (max_levels == 0) {   // back to Jacobi, with Jacobi's reliable-update interval
        c->amg_nl = 0;
        if (!c->ru_delta_set) c->ru_delta = 0.1;
        return 0;
    }
    if (c->structured || c->precision != MM_PREC_MIXED)
        return fail(MM_E_STATE, "the AMG preconditioner needs a general (unstructured) mixed-precision context");
    if (max_levels < 2 || max_levels > kAmgMaxLevels) return fail(MM_E_ARG, "levels outside [2, 10]");
    if (c->amg_built > 0) {   // built once per context (a later call re-enables it)
        if (max_levels != c->amg_built_max)
            return fail(MM_E_STATE, "AMG already set up with another level bound; use a new context");
        c->amg_nl = c->amg_built;
        c->t_mg.bytes = c->k_nnz * 8 + 4 * (c->N + 1) + 16 * (int64_t)c->N;
        if (!c->ru_delta_set) c->ru_delta = 1e-3;
        return 0;
    }
    // strength threshold 0.02: aggregates follow the strong direction of anisotropic
    // couplings (graded air boxes, the spherical shell's tensor); 0 (every coupling strong)
    // took 1.5x (graded box) to 3.9x (shell) the PCG iterations, 0.05 and above grew the
    // coarse operators faster than it cut iterations
    // smoothing degree 2 on level 0, 1 on the coarser levels: one PCG iteration more per
    // solve than degree 2 everywhere (19 -> 20 on a graded air-box mesh, 29 -> 30 on a
    // shell mesh) for ~19% less work per V-cycle (level 1 carries ~0.6 of level 0's entries)
    constexpr int kCoarse = 500, kDeg0 = 2, kDeg = 1;
    constexpr double kAlpha = 30.0, kTheta = 0.02;
    AmgHost H;
    const std::string err = amg_setup(c->hK, max_levels, kCoarse, kDeg0, kDeg, kAlpha, kTheta, H);
    if (!err.empty()) return fail(MM_E_STATE, err);
    const int nl = (int)H.levels.size();
    if (nl < 2) return fail(MM_E_STATE, "amg: the matrix is already at the coarsest size");
    int rc;
    for (int l = 0; l < nl; ++l) {
        const AmgLevelHost &h = H.levels[l];
        AmgLevelDev &d = c->amg[l];
        d.n = h.A.n;
        d.deg = (int)h.om.size();
        for (int k = 0; k < d.deg; ++k) d.om[k] = (float)h.om[k];
        if (l == 0) {
            d.A = c->K32;
            d.dinv = c->invDK32;
            if ((rc = c->alloc(&d.xb, (size_t)d.n)) || (rc = c->alloc(&c->amg_r0, (size_t)d.n))) return rc;
            std::vector<float> w0(d.n);
            for (int i = 0; i < d.n; ++i) w0[i] = (float)(h.om[0] * h.dinv[i]);
            if ((rc = c->upload(&c->amg_invw0, w0))) return rc;
        } else {
            float *dv;
            std::vector<float> di(h.dinv.begin(), h.dinv.end());
            if ((rc = c->upload(&dv, di))) return rc;
            d.dinv = dv;
            if ((rc = upload_csr32(c, h.A, d.Ac))) return rc;
            for (float **v : {&d.b, &d.xa, &d.xb, &d.r})
                if ((rc = c->alloc(v, (size_t)d.n))) return rc;
        }
        if (l + 1 < nl && ((rc = upload_sell32(c, h.P, d.P)) || (rc = upload_csr32(c, h.R, d.R)))) return rc;
    }
    c->amg_nc = H.levels.back().A.n;
    std::vector<float> inv(H.inv.begin(), H.inv.end());
    if ((rc = c->upload(&c->amg_inv, inv))) return rc;
    c->amg_built = c->amg_nl = nl;
    c->amg_built_max = max_levels;
    // a level-0 smoothing step reads the SELL matrix (value + column per stored entry, the
    // slice pointers), 1 / diag, b and x, and writes the new x
    c->t_mg.bytes = c->k_nnz * 8 + 4 * (c->N + 1) + 16 * (int64_t)c->N;
    if (!c->ru_delta_set) c->ru_delta = 1e-3;
    HIPC(hipDeviceSynchronize());
    return 0;
}

int mm_amg_levels(mm_ctx *c, int32_t *levels, int64_t *sizes) {
    if (!c || !levels) return fail(MM_E_ARG, "null argument");
    *levels = c->amg_nl;
    if (sizes)
        for (int l = 0; l < c->amg_nl; ++l) sizes[l] = c->amg[l].n;
    return 0;
}

int mm_amg_apply(mm_ctx *c, const float *r, float *z) {
    if (!c || !r || !z) return fail(MM_E_ARG, "null argument");
    if (c->amg_nl <= 0) return fail(MM_E_STATE, "AMG not set up (mm_amg)");
    float *res = amg_vcycle(c, r, z, c->amg[0].xb, nullptr, false);
    if (res != z) HIPC(hipMemcpyAsync(z, res, (size_t)c->N * sizeof(float), hipMemcpyDeviceToDevice, c->s));
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_mg_levels(mm_ctx *c, int32_t *levels, int32_t *dims) {
    if (!c || !levels) return fail(MM_E_ARG, "null argument");
    *levels = c->mg_nl;
    if (dims)
        for (int l = 0; l < c->mg_nl; ++l) {
            dims[3 * l] = c->mg[l].P.nx;
            dims[3 * l + 1] = c->mg[l].P.ny;
            dims[3 * l + 2] = c->mg[l].P.nz;
        }
    return 0;
}

int mm_mg_apply(mm_ctx *c, const float *r, float *z) {
    if (!c || !r || !z) return fail(MM_E_ARG, "null argument");
    if (c->mg_nl == 0) return fail(MM_E_STATE, "no multigrid preconditioner (mm_multigrid)");
    float *res = mg_vcycle(c, r, z, c->mg_xb[0], nullptr, false);
    if (res != z) HIPC(hipMemcpyAsync(z, res, (size_t)c->N * sizeof(float), hipMemcpyDeviceToDevice, c->s));
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

// ---------------------------------------------------------------- learned warm start

int mm_mlp_create(mm_ctx *c, int64_t max_batch, uint64_t seed, mm_mlp **out) {
    if (!c || !out || max_batch <= 0) return fail(MM_E_ARG, "null argument or max_batch <= 0");
    *out = nullptr;
    HIPC(hipSetDevice(c->device));
    mm_mlp *w = new mm_mlp();
    w->ctx = c;
    w->max_batch = max_batch;
    w->ld_max = (max_batch + 31) / 32 * 32;
    int rc = 0;
    auto bail = [&](int code) {
        delete w;
        return code;
    };
    if ((rc = w->alloc(&w->p, kMlpParams)) || (rc = w->alloc(&w->g, kMlpParams)) ||
        (rc = w->alloc(&w->m1, kMlpParams)) || (rc = w->alloc(&w->m2, kMlpParams)) ||
        (rc = w->alloc(&w->w1t, (size_t)kMlpW * kMlpP1)) || (rc = w->alloc(&w->w2t, (size_t)kMlpW * kMlpPW)) ||
        (rc = w->alloc(&w->w3t, (size_t)kMlpW * kMlpPW)) || (rc = w->alloc(&w->w2n, (size_t)kMlpW * kMlpPW)) ||
        (rc = w->alloc(&w->w3n, (size_t)kMlpW * kMlpPW)) || (rc = w->alloc(&w->y, (size_t)w->ld_max)) ||
        (rc = w->alloc(&w->dy, (size_t)w->ld_max)) || (rc = w->alloc(&w->loss, 1)))
        return bail(rc);
    for (int l = 0; l < 3; ++l)
        if ((rc = w->alloc(&w->h[l], (size_t)kMlpW * w->ld_max)) || (rc = w->alloc(&w->dz[l], (size_t)kMlpW * w->ld_max)))
            return bail(rc);
    // He-uniform initialisation (splitmix64 stream), zero biases
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t st = seed;
    auto uni = [&st]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const int offs[4] = {kMlpOffW1, kMlpOffW2, kMlpOffW3, kMlpOffW4};
    const int fan_in[4] = {kMlpIn, kMlpW, kMlpW, kMlpW}, count[4] = {kMlpIn * kMlpW, kMlpW * kMlpW, kMlpW * kMlpW, kMlpW};
    for (int l = 0; l < 4; ++l) {
        const double lim = std::sqrt((l < 3 ? 6.0 : 3.0) / fan_in[l]);
        for (int i = 0; i < count[l]; ++i) p[offs[l] + i] = (float)((2.0 * uni() - 1.0) * lim);
    }
    if (hipMemcpy(w->p, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(MM_E_HIP, "hipMemcpy of the initial parameters failed"));
    launch_mlp_publish(c->L(), w->p, w->w1t, w->w2t, w->w3t, w->w2n, w->w3n);
    if (hipStreamSynchronize(c->s) != hipSuccess) return bail(fail(MM_E_HIP, "stream synchronise failed"));
    *out = w;
    return 0;
}

int mm_mlp_destroy(mm_mlp *w) {
    if (!w) return 0;
    if (w->ctx && w->ctx->ws_net == w) w->ctx->ws_net = nullptr;
    delete w;
    return 0;
}

int mm_mlp_set_params(mm_mlp *w, const float *params_host) {
    if (!w || !params_host) return fail(MM_E_ARG, "null argument");
    mm_ctx *c = w->ctx;
    HIPC(hipMemcpyAsync(w->p, params_host, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->s));
    HIPC(hipMemsetAsync(w->m1, 0, kMlpParams * sizeof(float), c->s));
    HIPC(hipMemsetAsync(w->m2, 0, kMlpParams * sizeof(float), c->s));
    w->step = 0;
    launch_mlp_publish(c->L(), w->p, w->w1t, w->w2t, w->w3t, w->w2n, w->w3n);
    HIPC(hipStreamSynchronize(c->s));
    c->ws_have_prev = false;
    return 0;
}

int mm_mlp_get_params(mm_mlp *w, float *params_host) {
    if (!w || !params_host) return fail(MM_E_ARG, "null argument");
    HIPC(hipMemcpyAsync(params_host, w->p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, w->ctx->s));
    HIPC(hipStreamSynchronize(w->ctx->s));
    return 0;
}

int mm_mlp_features(mm_ctx *c, const double *m, uint16_t *X) {
    if (!c || !m || !X) return fail(MM_E_ARG, "null argument");
    if (!c->structured) return fail(MM_E_STATE, "the learned warm start needs a structured (Kuhn grid) context");
    Launch L = c->L();
    for (int a = 0; a < 3; ++a)
        launch_reduce(L, RED_MH, c->N, m, c->meas, nullptr, nullptr, a == 0, a == 1, a == 2, c->ws_red + a);
    launch_mlp_features(L, c->F, m, c->ws_red, X);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_mlp_forward(mm_mlp *w, int64_t n, const uint16_t *X, float *y) {
    if (!w || !X || !y || n < 0) return fail(MM_E_ARG, "null argument or n < 0");
    launch_mlp_forward(w->ctx->L(), w->dev(), n, X, y);
    HIPC(hipStreamSynchronize(w->ctx->s));
    return 0;
}

namespace {
// forward with stored activations, loss, backward, all parameter gradients into w->g
int mlp_grads(mm_mlp *w, int64_t n, const uint16_t *X, const float *t) {
    mm_ctx *c = w->ctx;
    Launch L = c->L();
    const int64_t ld = (n + 31) / 32 * 32;
    HIPC(hipMemsetAsync(w->loss, 0, sizeof(double), c->s));
    HIPC(hipMemsetAsync(w->g, 0, kMlpParams * sizeof(float), c->s));
    const MlpDev d = w->dev();
    launch_mlp_forward_train(L, d, n, X, w->y, w->h[0], w->h[1], w->h[2], ld);
    launch_mlp_loss(L, n, w->y, t, w->dy, w->loss);
    launch_mlp_backward(L, d, n, w->dy, w->h[0], w->h[1], w->h[2], w->dz[0], w->dz[1], w->dz[2], ld);
    launch_mlp_param_grads(L, n, ld, X, w->dy, w->h[0], w->h[1], w->h[2], w->dz[0], w->dz[1], w->dz[2], w->g);
    return 0;
}
}  // namespace

int mm_mlp_grads(mm_mlp *w, int64_t n, const uint16_t *X, const float *target, float *grad_host,
                 double *loss_host) {
    if (!w || !X || !target || !grad_host) return fail(MM_E_ARG, "null argument");
    if (n <= 0 || n > w->max_batch) return fail(MM_E_ARG, "batch size outside (0, max_batch]");
    int rc = mlp_grads(w, n, X, target);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(grad_host, w->g, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, w->ctx->s));
    double l = 0;
    HIPC(hipMemcpyAsync(&l, w->loss, sizeof(double), hipMemcpyDeviceToHost, w->ctx->s));
    HIPC(hipStreamSynchronize(w->ctx->s));
    if (loss_host) *loss_host = l;
    return 0;
}

int mm_mlp_train_step(mm_mlp *w, int64_t n, const uint16_t *X, const float *target, float lr, double *loss_host) {
    if (!w || !X || !target) return fail(MM_E_ARG, "null argument");
    if (n <= 0 || n > w->max_batch) return fail(MM_E_ARG, "batch size outside (0, max_batch]");
    int rc = mlp_grads(w, n, X, target);
    if (rc) return rc;
    mm_ctx *c = w->ctx;
    launch_mlp_adam(c->L(), w->p, w->g, w->m1, w->m2, lr, ++w->step);
    launch_mlp_publish(c->L(), w->p, w->w1t, w->w2t, w->w3t, w->w2n, w->w3n);
    c->ws_have_prev = false;
    double l = 0.0;
    HIPC(hipMemcpyAsync(&l, w->loss, sizeof(double), hipMemcpyDeviceToHost, c->s));
    // the batch buffers are the caller's: return only when the step has read them
    HIPC(hipStreamSynchronize(c->s));
    if (loss_host) *loss_host = l;
    return 0;
}

int mm_warm_start(mm_ctx *c, mm_mlp *w, double scale) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (w && !c->structured) return fail(MM_E_STATE, "the learned warm start needs a structured (Kuhn grid) context");
    if (w && w->ctx != c) return fail(MM_E_ARG, "the network belongs to another context");
    if (w && !c->ws_X) {
        if (c->alloc(&c->ws_X, (size_t)c->N * kMlpIn) || c->alloc(&c->ws_pred, (size_t)c->N) ||
            c->alloc(&c->ws_prev, (size_t)c->N))
            return MM_E_HIP;
    }
    c->ws_net = w;
    c->ws_scale = scale;
    c->ws_have_prev = false;
    c->ws_calls = 0;
    return 0;
}

// ---------------------------------------------------------------- learned preconditioner

int mm_lp_setup(mm_ctx *c) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (!c->structured) return fail(MM_E_STATE, "the learned preconditioner needs a structured (Kuhn grid) context");
    if (c->lp_r) return 0;
    const size_t n = (size_t)c->N;
    if (c->alloc(&c->lp_r, n) || c->alloc(&c->lp_z, n) || c->alloc(&c->lp_p, n) || c->alloc(&c->lp_q, n) ||
        c->alloc(&c->lp_rold, n) || c->alloc(&c->lp_w, n) || c->alloc(&c->lp_X, n * kMlpIn) ||
        c->alloc(&c->lp_y, n))
        return MM_E_HIP;
    return 0;
}

namespace {
// z = net correction of r (in units of the RMS of r / diag K), or z = r / diag K
void lp_precond(mm_ctx *c, mm_mlp *net, double scale, const double *r, double *z, bool learned) {
    Launch L = c->L();
    if (!learned) {
        launch_lp_scaled(L, c->N, r, c->invDK, z);
        return;
    }
    launch_lp_scaled(L, c->N, r, c->invDK, c->lp_w);
    launch_reduce(L, RED_SS, c->N, c->lp_w, nullptr, nullptr, nullptr, 0, 0, 0, &c->dsc->red[4]);
    launch_lp_features(L, c->F, c->P, c->lp_w, &c->dsc->red[4], c->lp_X);
    launch_mlp_forward(L, net->dev(), c->N, c->lp_X, c->lp_y);
    launch_lp_apply(L, c->F, c->lp_y, &c->dsc->red[4], scale, z);
}
}  // namespace

int mm_lp_features(mm_ctx *c, const double *r, uint16_t *X) {
    if (!c || !r || !X) return fail(MM_E_ARG, "null argument");
    if (!c->lp_r) return fail(MM_E_STATE, "call mm_lp_setup first");
    Launch L = c->L();
    launch_lp_scaled(L, c->N, r, c->invDK, c->lp_w);
    launch_reduce(L, RED_SS, c->N, c->lp_w, nullptr, nullptr, nullptr, 0, 0, 0, &c->dsc->red[4]);
    launch_lp_features(L, c->F, c->P, c->lp_w, &c->dsc->red[4], X);
    HIPC(hipStreamSynchronize(c->s));
    return 0;
}

int mm_lp_pcg(mm_ctx *c, mm_mlp *net, double scale, const double *b, double *x, double tol, int32_t maxiter,
              int32_t every, int32_t *iters_host, int32_t *lp_apps_host) {
    if (!c || !b || !x) return fail(MM_E_ARG, "null argument");
    if (maxiter < 0 || every < 0) return fail(MM_E_ARG, "maxiter and every must be >= 0");
    if (!c->lp_r) return fail(MM_E_STATE, "call mm_lp_setup first");
    if (every > 0 && (!net || net->ctx != c)) return fail(MM_E_ARG, "a network of this context is needed");
    Launch L = c->L();
    const int64_t n = c->N;
    double *r = c->lp_r, *z = c->lp_z, *p = c->lp_p, *q = c->lp_q, *d = c->lp_rold;
    int rc = 0, it = 0, apps = 0;
    auto dot = [&](const double *a, const double *bb, double *out) {
        return reduce(c, RED_DOT, n, a, bb, nullptr, nullptr, out, 1);
    };
    double bb = 0, rr = 0, rz = 0;
    auto residual = [&]() -> int {   // r = b - K x, r.r
        op_K(c, x, q, nullptr, nullptr);
        launch_axpby(L, n, 1.0, b, -1.0, q, r, nullptr);
        return dot(r, r, &rr);
    };
    // Galerkin step along the network's predicted error d = N(r): gamma = d.r / d.K d
    // minimises the K-norm of the error along d, so it can only decrease
    auto galerkin = [&]() -> int {
        lp_precond(c, net, scale, r, d, true);
        ++apps;
        double dr = 0, dkd = 0;
        op_K(c, d, q, nullptr, nullptr);
        int e = dot(d, r, &dr);
        if (!e) e = dot(d, q, &dkd);
        if (e) return e;
        if (dkd > 0.0 && std::isfinite(dr / dkd)) {
            launch_axpby(L, n, dr / dkd, d, 1.0, x, x, nullptr);
            return residual();
        }
        return 0;
    };
    if ((rc = dot(b, b, &bb)) || (rc = residual())) return rc;
    const double bound = tol * tol * bb;
    bool restart = true;
    while (rr > bound && it < maxiter) {
        if (restart) {   // (re)start the Jacobi CG from the current iterate
            if (every > 0 && (rc = galerkin())) return rc;
            if (!(rr > bound)) break;
            lp_precond(c, net, scale, r, z, false);
            if ((rc = dot(r, z, &rz))) return rc;
            HIPC(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
            restart = false;
        }
        double pq = 0;
        op_K(c, p, q, nullptr, nullptr);
        if ((rc = dot(p, q, &pq))) return rc;
        const double alpha = rz / pq;
        launch_axpby(L, n, alpha, p, 1.0, x, x, nullptr);    // x += alpha p
        launch_axpby(L, n, -alpha, q, 1.0, r, r, nullptr);   // r -= alpha q
        ++it;
        if ((rc = dot(r, r, &rr))) return rc;
        if (!(rr > bound)) break;
        if (every > 0 && it % every == 0) {
            restart = true;
            continue;
        }
        double rz_new = 0;
        lp_precond(c, net, scale, r, z, false);
        if ((rc = dot(r, z, &rz_new))) return rc;
        launch_axpby(L, n, 1.0, z, rz_new / rz, p, p, nullptr);   // p = z + beta p
        rz = rz_new;
    }
    HIPC(hipStreamSynchronize(c->s));
    if (iters_host) *iters_host = it;
    if (lp_apps_host) *lp_apps_host = apps;
    return rr > bound ? MM_W_MAXITER : 0;
}

int mm_secant_start(mm_ctx *c, int32_t on) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (on && !c->u_sec) {
        int rc;
        if ((rc = c->alloc(&c->u_sec, (size_t)c->N)) || (rc = c->alloc(&c->sec_red, 8))) return rc;
    }
    c->sec_on = on != 0;
    c->sec_next = false;
    return 0;
}

int mm_warm_start_policy(mm_ctx *c, int32_t policy) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (policy != MM_WS_EVERY && policy != MM_WS_MEASURED) return fail(MM_E_ARG, "unknown warm-start policy");
    c->ws_policy = policy;
    c->ws_solves = c->ws_n[0] = c->ws_n[1] = 0;
    c->ws_every = kWsEvery;
    c->ws_next = kWsEvery - 1;
    c->ws_probe = false;
    c->ws_mean[0] = c->ws_mean[1] = 0.0;
    return 0;
}

int mm_warm_start_report(mm_ctx *c, int64_t *solves, int64_t *warm_solves, double *mean_s_plain,
                         double *mean_s_warm) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (solves) *solves = c->ws_solves;
    if (warm_solves) *warm_solves = c->ws_n[1];
    if (mean_s_plain) *mean_s_plain = c->ws_mean[0];
    if (mean_s_warm) *mean_s_warm = c->ws_mean[1];
    return 0;
}

int mm_warm_start_stats(mm_ctx *c, int64_t *calls, double *last_gamma) {
    if (!c) return fail(MM_E_ARG, "null context");
    if (calls) *calls = c->ws_calls;
    if (last_gamma) {
        *last_gamma = 0.0;
        if (c->ws_red) {
            HIPC(hipMemcpyAsync(last_gamma, c->ws_red + 7, sizeof(double), hipMemcpyDeviceToHost, c->s));
            HIPC(hipStreamSynchronize(c->s));
        }
    }
    return 0;
}

int mm_timing_enable(mm_ctx *c, int32_t on) {
    if (!c) return fail(MM_E_ARG, "null context");
    c->timing = on != 0;
    for (KernelTimer *t : {&c->t_spmv, &c->t_upd, &c->t_mg}) {
        if (c->timing && t->ev.empty()) {
            t->ev.resize(2 * 1024);   // sampled pairs: iterations up to 16 x 1024 per solve
            // no system-scope fence at the events: they would write back the L2 around
            // the timed dispatch and slow the kernel they measure
            for (auto &e : t->ev) HIPC(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        }
        t->launches = 0;
        t->ms = 0.0;
    }
    c->n_ru = c->n_corr = 0;
    int rc = sync_scalars(c);
    if (rc) return rc;
    c->vcyc0 = c->hsc->mg_vcycles;
    return 0;
}

int mm_timing_get(mm_ctx *c, mm_kernel_timing *o) {
    if (!c || !o) return fail(MM_E_ARG, "null argument");
    o->spmv_launches = c->t_spmv.launches;
    o->spmv_ms = c->t_spmv.ms;
    o->spmv_bytes = c->t_spmv.bytes;
    o->update_launches = c->t_upd.launches;
    o->update_ms = c->t_upd.ms;
    o->update_bytes = c->t_upd.bytes;
    o->mg_smooth_launches = c->t_mg.launches;
    o->mg_smooth_ms = c->t_mg.ms;
    o->mg_smooth_bytes = c->t_mg.bytes;
    int rc = sync_scalars(c);
    if (rc) return rc;
    o->vcycles = c->mg_nl > 0 && c->mg_tail < c->mg_nl ? (int64_t)(c->hsc->mg_vcycles - c->vcyc0) : -1;
    o->ru_passes = c->n_ru;
    o->corrections = c->n_corr;
    return 0;
}

}  // extern "C"
