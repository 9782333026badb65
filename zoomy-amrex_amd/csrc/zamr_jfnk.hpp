#pragma once
// solver.implicit_global: the global backward-Euler source solve of
// Source/imex_solver.H:102-160 on the device.  After a stage's flux update
// (Qexp), Newton on F(Q) = Q - Qexp - dt S(Q): at most 12 iterations, stop
// at |F| < 1e-12 or |F| < 1e-9 |F_0|; each correction solves J d = -F with
// restarted GMRES(30) (zero initial guess, no preconditioner, tolerance
// 1e-4 relative) whose operator is the finite-difference product
// J x = x - (dt/eps) (S(Q + eps x) - S(Q)), eps = 1e-7 (1 + |Q|) / |x|.
// S is ComputeSourceField's: aux = update_aux_variables(q, 0), then
// Model::source, cell by cell (every source here is local, so its ghost
// refresh does not enter).  Gram-Schmidt is classical, applied twice.
//
// Vectors hold the valid cells of every box of the level, box by box, row
// by row, component-major over the level: v[n * N + c].  Everything runs on
// the device, the control flow included: an Arnoldi step is three passes
// over the vectors (operator + first Gram-Schmidt dots; first subtraction +
// second dots; second subtraction + norm).  Each pass leaves per-block
// partial sums, summed in a fixed order (deterministic): the Gram-Schmidt
// coefficients by every block of the pass that uses them, the norm by a
// one-block kernel after the third, which applies the Givens rotations,
// updates the residual estimate and decides whether the cycle goes on.  A
// state word (`phase`) in device memory steers which of the queued kernels
// act, so the host queues several Arnoldi steps, the cycle's end (least
// squares, update, restart residual) and reads the state back once per
// Newton iteration.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "zamr_error.hpp"
#include "zamr_kernels.hpp"

namespace zamr {

// Rows of Model::source that can be nonzero: M::zamr_source_rows when the
// model declares it (the shallow-water friction acts on the momentum rows
// only), every row otherwise.  The solve's vectors hold those rows only: in
// the others F = Q - Qexp - dt S is 0, the operator is the identity on a
// zero component, and every Krylov vector, dot product and update is the
// same with them left out (they only add exact zeros).  The residual pass
// checks S there and stops the solve if a declared-zero row is not.
template <class M, class = void>
struct SourceRows {
    static constexpr int n = M::n_dof_q;
    __host__ __device__ static constexpr int row(int a) { return a; }
    __host__ __device__ static constexpr bool listed(int r) { return true; }
};
template <class M>
struct SourceRows<M, std::void_t<decltype(M::zamr_source_rows)>> {
    static constexpr int n = (int)(sizeof(M::zamr_source_rows) / sizeof(int));
    __host__ __device__ static constexpr int row(int a) { return M::zamr_source_rows[a]; }
    __host__ __device__ static constexpr bool listed(int r) {
        for (int a = 0; a < n; ++a)
            if (M::zamr_source_rows[a] == r) return true;
        return false;
    }
};

// Box of compact cell c: bstart[b] <= c < bstart[b + 1].
__device__ __forceinline__ int jf_box(const long long* __restrict__ bstart, int nb, long long c) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct JfCell {
    int b, i, j;
};
__device__ __forceinline__ JfCell jf_cell(const BoxDesc* __restrict__ boxes, const long long* __restrict__ bstart,
                                          int nb, long long c) {
    JfCell r;
    r.b = jf_box(bstart, nb, c);
    const BoxDesc& B = boxes[r.b];
    const int w = B.hi[0] - B.lo[0] + 1;
    const int k = (int)(c - bstart[r.b]);
    r.i = B.lo[0] + k % w;
    r.j = B.lo[1] + k / w;
    return r;
}

// Device state of one solve (GMRES(30) cycle + Newton bookkeeping).
constexpr int kJfRestart = 30;
struct JfState {
    double H[kJfRestart + 1][kJfRestart];
    double cs[kJfRestart], sn[kJfRestart], g[kJfRestart + 1], y[kJfRestart];
    double gp[kJfRestart + 1];  // g[k] before column k's rotation (each written once a cycle)
    double h1[kJfRestart + 1], h2[kJfRestart + 1];
    double vnorm[kJfRestart + 1];
    double r0, eps_num, xnorm, resid, resid0;
    int phase;  // 0 Arnoldi, 2 update, 3 restart residual, 6 Q update, 5 finished
    int k;      // Arnoldi column of the current cycle
    int kd;     // columns of the cycle done
    int done;   // the cycle reached the tolerance
    int src0;   // the cycle's first vector before normalising: 0 rhs, 1 r
    int first;  // the first cycle of this Newton iteration (x starts at 0)
    int nd;     // Newton converged
    int newton_its;
    long long its;          // Arnoldi steps of this Newton iteration's GMRES
    long long gmres_total;  // over the solve
    int bad_rows;           // a row declared source-free had a nonzero source
    int nrun;               // 1 + the last Newton iteration whose residual pass ran
    int umark, ustep;       // sequence number and Arnoldi step of the last step whose third pass ran
    int its0, pad_;         // its at the cycle's start
    int its_hist[16];       // Arnoldi steps of each finished Newton iteration
};

// What the host reads after a round: the state's words it steers by, copied
// into host-visible memory by one small kernel at the end of the round (the
// kernels of the round write device memory only).
struct JfStatus {
    int phase, nd, bad_rows, pad_;
    long long its, newton_its, gmres_total;
    int its_hist[16];  // Arnoldi steps of each finished Newton iteration (the host's next queueing)
};

enum : int { kJfArnoldi = 0, kJfUpdate = 2, kJfRestartRes = 3, kJfFinished = 5, kJfQUpdate = 6 };

// Cells per thread of a cell pass (elements per thread of a vector pass: 4x
// that).  Measured at 1024^2 (ms per step): 1 -> 1.57, 2 -> 1.22, 4 -> 1.26,
// 8 -> 1.42; fewer blocks cut the per-block partial sums every pass reads,
// more cells per thread cut the loads a thread keeps in flight.
constexpr int kJfCellsPerThread = 2;
constexpr long kJfMaxIter = 2000;
constexpr int kJfNewtonMax = 12;

// The status of the state, for the host; also each finished iteration's
// Arnoldi step count.
// (all threads of a block, after the state's last write and a barrier)
__device__ __forceinline__ void jf_publish(const JfState* st, JfStatus* hs) {
    if (threadIdx.x < 16) hs->its_hist[threadIdx.x] = st->its_hist[threadIdx.x];
    if (threadIdx.x != 0) return;
    hs->phase = st->phase;
    hs->nd = st->nd;
    hs->bad_rows = st->bad_rows;
    hs->its = st->its;
    hs->newton_its = st->newton_its;
    hs->gmres_total = st->gmres_total;
}

// Sum of a per-thread value over the block (256 threads), in every thread.
// A fixed order: wave sums by xor butterflies, then the four wave sums in
// order.
__device__ __forceinline__ double jf_block_sum(double v, double* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[wv] = v;
    __syncthreads();
    return ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

// out[j] = sum over the nblk blocks of part[j * nblk + b], j < m <= 31, by a
// fixed tree -- thread t sums blocks t, t + 256, ... in order, a wave's sums
// by xor butterflies, then the four wave sums in order (jf_block_sum's) --
// so every block gets the same values.  A thread's loads for eight of its
// blocks go out together (indices clamped into range, the extra values not
// added), so a row costs a memory latency per 2048 blocks instead of one per
// load; one barrier for all rows.  One row at a time: more loads in flight
// would take registers the vector passes that call this need for their
// occupancy.  `out` is shared memory, valid in every thread on return.
__device__ __forceinline__ void jf_sums(const double* __restrict__ part, int nblk, int m, double* out) {
    constexpr int kR = 1, kU = 8;
    __shared__ double wsum[kJfRestart + 1][4];
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
    for (int j0 = 0; j0 < m; j0 += kR) {
        double s[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) s[r] = 0.0;
        for (int b0 = t; b0 < nblk; b0 += kU * 256) {
            double v[kR][kU];
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                const long long row = (long long)min(j0 + r, m - 1) * nblk;
#pragma unroll
                for (int u = 0; u < kU; ++u) v[r][u] = part[row + min(b0 + u * 256, nblk - 1)];
            }
#pragma unroll
            for (int r = 0; r < kR; ++r)
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (b0 + u * 256 < nblk) s[r] += v[r][u];
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
            double v = s[r];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0 && j0 + r < m) wsum[j0 + r][wv] = v;
        }
    }
    __syncthreads();
    if (t < m) out[t] = ((wsum[t][0] + wsum[t][1]) + wsum[t][2]) + wsum[t][3];
    __syncthreads();
}

// One cell's operator on the source rows: out = (u + (-dt/eps) S(q + eps u))
// + (dt/eps) Sbase, eps = eps_num / |u| (|u| = unorm; 0 gives 0); q has all
// rows, u, Sbase and out the source rows.
template <class M>
__device__ __forceinline__ void jf_apply_cell(const BoxDesc* __restrict__ boxes, const long long* __restrict__ bstart,
                                              int nb, long long c, const double* qc, const double* uc,
                                              const double* sb, double eps_num, double unorm, double dt,
                                              const typename Sys<M>::VP& p, double time, double* out) {
    using S = Sys<M>;
    using R = SourceRows<M>;
    constexpr int nq = S::nq;
    if (unorm == 0.0) {
#pragma unroll
        for (int a = 0; a < R::n; ++a) out[a] = 0.0;
        return;
    }
    const double eps = eps_num / unorm;
    const JfCell e = jf_cell(boxes, bstart, nb, c);
    const BoxView B = view(boxes, e.b);
    typename S::VQ qq;
#pragma unroll
    for (int n = 0; n < nq; ++n) qq(n, 0) = qc[n];
#pragma unroll
    for (int a = 0; a < R::n; ++a) qq(R::row(a), 0) = 1.0 * qc[R::row(a)] + eps * uc[a];
    const typename S::VX X = S::center(B, e.i, e.j);
    typename S::VA av{};
    av = M::update_aux_variables(qq, av, p, time, X);
    const typename S::VQ sv = M::source(qq, av, p, time, X);
    const double r = dt / eps;
#pragma unroll
    for (int a = 0; a < R::n; ++a) out[a] = (1.0 * uc[a] + (-r) * sv(R::row(a), 0)) + r * sb[a];
}

#define JF_CELLS_LOOP                                                                              \
    for (int pc = 0; pc < kJfCellsPerThread; ++pc)                                                 \
        if (const long long c = ((long long)blockIdx.x * kJfCellsPerThread + pc) * blockDim.x + threadIdx.x; \
            c < N)

// Zero the state (all threads of one block, 8-byte words), before the
// block's threads part ways.
__device__ __forceinline__ void jf_init_state(JfState* st) {
    static_assert(sizeof(JfState) % 8 == 0, "the state is a whole number of 8-byte words");
    unsigned long long* w = reinterpret_cast<unsigned long long*>(st);
    for (int i = threadIdx.x; i < (int)(sizeof(JfState) / 8); i += blockDim.x) w[i] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) st->umark = -1;  // no step yet (sequence numbers start at 1)
}

// Valid cells of buffer `buf` -> compact q (all rows) and Qexp (the source
// rows), or compact q -> buffer (scatter).  With `st` (the gather that
// starts a solve) block 0 also zeroes the solver state.
template <class M>
__global__ void __launch_bounds__(256) jf_move_kernel(const BoxDesc* __restrict__ boxes,
                                                      const long long* __restrict__ bstart, int nb, long long N,
                                                      int buf, double* __restrict__ v, int to_box,
                                                      double* __restrict__ qexp, JfState* st) {
    constexpr int nq = M::n_dof_q;
    using R = SourceRows<M>;
    if (st && blockIdx.x == 0) jf_init_state(st);  // the gather starts a solve
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const JfCell e = jf_cell(boxes, bstart, nb, c);
    const BoxDesc& B = boxes[e.b];
    double* q = buf == 0 ? B.q[0] : (buf == 1 ? B.q[1] : B.q[2]);
    const long long idx = (long long)(e.i - B.gl[0]) + (long long)(e.j - B.gl[1]) * B.sy;
    if (to_box) {
#pragma unroll
        for (int a = 0; a < R::n; ++a) q[idx + R::row(a) * B.sc] = v[R::row(a) * N + c];
        return;
    }
#pragma unroll
    for (int n = 0; n < nq; ++n) v[n * N + c] = q[idx + n * B.sc];
#pragma unroll
    for (int a = 0; a < R::n; ++a) qexp[a * N + c] = q[idx + R::row(a) * B.sc];
}


// A new cycle's Hessenberg matrix and residual vector, zeroed by all threads.
__device__ __forceinline__ void jf_clear_cycle(JfState* st) {
    for (int i = threadIdx.x; i < (kJfRestart + 1) * kJfRestart; i += blockDim.x) (&st->H[0][0])[i] = 0.0;
    for (int j = threadIdx.x; j <= kJfRestart; j += blockDim.x) st->g[j] = 0.0;
    __syncthreads();
}

// A finished Newton iteration's bookkeeping (one thread).
__device__ __forceinline__ void jf_newton_done(JfState* st) {
    if (st->newton_its < 16) st->its_hist[st->newton_its] = (int)st->its;
    st->gmres_total += st->its;
    st->newton_its += 1;
    st->phase = kJfFinished;
}

// Newton iteration `it` runs once the previous one is finished, or, fused,
// right after the previous one's GMRES cycle converged with its correction
// not yet applied (phase kJfUpdate, done): the Newton residual pass then
// applies it first.  Read by the residual pass and its tail alike (the pass
// changes none of these words).
__device__ __forceinline__ bool jf_newton_fused(const JfState* st, int it) {
    return it > 0 && st->phase == kJfUpdate && st->done && st->newton_its == it - 1;
}
__device__ __forceinline__ bool jf_newton_runs(const JfState* st, int it) {
    return !st->nd && ((st->newton_its == it && (it == 0 || st->phase == kJfFinished)) || jf_newton_fused(st, it));
}

// Newton residual: (fused) Q += x + y_j V_j on the source rows, the converged
// cycle's correction as jf_update_kernel applies it; then S = S(q),
// F = (q - Qexp) + (-dt) S, rhs = -F on the source rows; partials of |F|^2
// (row 0) and of |q|^2 over every row (row 1).
template <class M>
__global__ void __launch_bounds__(256) jf_newton_kernel(const BoxDesc* __restrict__ boxes,
                                                        const long long* __restrict__ bstart, int nb, long long N,
                                                        double* __restrict__ q, const double* __restrict__ qexp,
                                                        const double* __restrict__ x, double* const* __restrict__ V,
                                                        double* __restrict__ S, double* __restrict__ rhs, double dt,
                                                        Params pp, double time, JfState* st, double* part, int it) {
    using Sy = Sys<M>;
    using R = SourceRows<M>;
    constexpr int nq = Sy::nq;
    __shared__ double lds[4];
    __shared__ double yv[kJfRestart];
    if (!jf_newton_runs(st, it)) return;
    const bool fused = jf_newton_fused(st, it);
    const int kd = fused ? st->kd : 0, first = st->first;
    if ((int)threadIdx.x < kd) yv[threadIdx.x] = st->y[threadIdx.x];
    __syncthreads();
    typename Sy::VP p;
    for (int k = 0; k < Sy::np; ++k) p(k, 0) = pp.v[k];
    double f2 = 0.0, q2 = 0.0;
    bool bad = false;
    JF_CELLS_LOOP {
        const JfCell e = jf_cell(boxes, bstart, nb, c);
        const BoxView B = view(boxes, e.b);
        typename Sy::VQ qq;
#pragma unroll
        for (int n = 0; n < nq; ++n) qq(n, 0) = q[n * N + c];
        if (fused) {
#pragma unroll
            for (int a = 0; a < R::n; ++a) {
                const long long i = a * N + c;
                double xv = first ? 0.0 : x[i];
                for (int j = 0; j < kd; ++j) xv += yv[j] * V[j][i];
                qq(R::row(a), 0) = qq(R::row(a), 0) + 1.0 * xv;
                q[R::row(a) * N + c] = qq(R::row(a), 0);
            }
        }
        const typename Sy::VX X = Sy::center(B, e.i, e.j);
        typename Sy::VA av{};
        av = M::update_aux_variables(qq, av, p, time, X);
        const typename Sy::VQ sv = M::source(qq, av, p, time, X);
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            q2 += qq(n, 0) * qq(n, 0);
            if (!R::listed(n) && sv(n, 0) != 0.0) bad = true;
        }
#pragma unroll
        for (int a = 0; a < R::n; ++a) {
            const double F = (qq(R::row(a), 0) - qexp[a * N + c]) + (-dt) * sv(R::row(a), 0);
            S[a * N + c] = sv(R::row(a), 0);
            rhs[a * N + c] = -F;
            f2 += F * F;
        }
    }
    if (bad) st->bad_rows = 1;
    double v = jf_block_sum(f2, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
    v = jf_block_sum(q2, lds);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = v;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->nrun = it + 1;
}

// The end of Newton iteration `it`'s residual pass, from its |F|^2 and |q|^2
// (one block, all threads; the caller read `fused` before any of it): the
// previous iteration's bookkeeping when the pass applied its correction,
// the convergence test, and the start of the GMRES cycle.
__device__ __forceinline__ void jf_newton_end(JfState* st, int it, double f2, double q2, bool fused) {
    const double resid = sqrt(f2);
    const double r0 = it == 0 ? resid : st->resid0;
    const bool conv = resid < 1e-12 || (r0 > 0.0 && resid < 1e-9 * r0) || st->bad_rows;
    if (!conv) jf_clear_cycle(st);
    if (threadIdx.x != 0) return;
    if (fused) jf_newton_done(st);  // the previous iteration's end
    if (it == 0) st->resid0 = resid;
    st->resid = resid;
    if (conv) {
        st->nd = 1;
        return;
    }
    st->eps_num = 1e-7 * (1.0 + sqrt(q2));
    st->r0 = resid;  // |rhs| = |F|
    st->vnorm[0] = resid;
    st->src0 = 0;
    st->k = 0;
    st->kd = 0;
    st->done = 0;
    st->first = 1;
    st->its = 0;
    st->its0 = 0;
    st->g[0] = resid;
    st->gp[0] = resid;
    st->phase = kJfArnoldi;
}

// Newton's convergence test and the start of the GMRES cycle.
// The end of a residual pass that no Arnoldi step follows in the queue (the
// converging test the host queues last).
__device__ __forceinline__ void jf_newton_tail_body(JfState* st, const double* part, int nblk, int it) {
    __shared__ double tot[2];
    if (st->nd || st->nrun != it + 1) return;
    const bool fused = jf_newton_fused(st, it);
    jf_sums(part, nblk, 2, tot);
    jf_newton_end(st, it, tot[0], tot[1], fused);
}

__global__ void __launch_bounds__(256) jf_newton_tail(JfState* st, const double* part, int nblk, int it,
                                                      JfStatus* hs, int pub) {
    jf_newton_tail_body(st, part, nblk, it);
    if (pub) {  // the last kernel of a queued round: the status for the host
        __syncthreads();
        jf_publish(st, hs);
    }
}

// The end of Arnoldi step k from w2 = |w|^2 (one block, all threads):
// H[j][k] = h1 + h2, H[k+1][k] = |w|, the Givens rotations, the residual
// estimate, the cycle's next move and, at its end, the least-squares
// coefficients y = H^-1 g.  It reads only words written by earlier kernels
// and never rewritten by this one (the step's h1 and h2, the earlier
// rotations and rotated g, the pre-rotation gp[k], r0, its0), so every
// block of a pass can work it out while one of them (`store`) writes the
// results.  Returns, in every thread, whether the cycle ends here.
__device__ __forceinline__ bool jf_column_end(JfState* st, int k, double w2, double rtol, bool store) {
    // the column, the rotations and g in shared memory: loaded by all
    // threads at once, worked on by one
    __shared__ double hc[kJfRestart + 1], cs[kJfRestart], sn[kJfRestart], g[kJfRestart + 1], y[kJfRestart];
    __shared__ double Hs[kJfRestart][kJfRestart];
    __shared__ int fin;
    const int t = threadIdx.x;
    if (t <= k) hc[t] = 0.0 + st->h1[t] + st->h2[t];
    if (t < k) {
        cs[t] = st->cs[t];
        sn[t] = st->sn[t];
        g[t] = st->g[t];
    }
    __syncthreads();
    if (t == 0) {
        const double hn = sqrt(w2);
        hc[k + 1] = hn;
        for (int j = 0; j < k; ++j) {
            const double u = cs[j] * hc[j] + sn[j] * hc[j + 1];
            hc[j + 1] = -sn[j] * hc[j] + cs[j] * hc[j + 1];
            hc[j] = u;
        }
        const double den = hypot(hc[k], hc[k + 1]);
        cs[k] = den > 0.0 ? hc[k] / den : 1.0;
        sn[k] = den > 0.0 ? hc[k + 1] / den : 0.0;
        hc[k] = den;
        hc[k + 1] = 0.0;
        const double gk = st->gp[k];
        const double gnext = -sn[k] * gk;
        g[k] = cs[k] * gk;
        const long long its = st->its0 + k + 1;
        const bool conv = fabs(gnext) < st->r0 * rtol;
        fin = conv || hn == 0.0 || its >= kJfMaxIter || k + 1 == kJfRestart;
        if (store) {
            st->cs[k] = cs[k];
            st->sn[k] = sn[k];
            st->g[k] = g[k];
            st->gp[k + 1] = gnext;
            st->its = its;
            st->kd = k + 1;
            if (fin) {
                st->done = conv ? 1 : 0;
                st->phase = kJfUpdate;
            } else {
                st->vnorm[k + 1] = hn;
                st->k = k + 1;
            }
        }
    }
    __syncthreads();
    const bool f = fin;
    if (!store) return f;
    if (t <= k + 1) st->H[t][k] = hc[t];
    if (!f) return f;
    // the cycle's end: y over its kd = k + 1 columns (H upper triangular
    // after the rotations; column k is the one above)
    const int kd = k + 1;
    for (int e = t; e < kd * kd; e += blockDim.x) {
        const int i = e / kd, j = e % kd;
        if (j >= i) Hs[i][j] = j == k ? hc[i] : st->H[i][j];
    }
    __syncthreads();
    if (t == 0) {
        for (int i = kd - 1; i >= 0; --i) {
            double u = g[i];
            for (int j = i + 1; j < kd; ++j) u -= Hs[i][j] * y[j];
            y[i] = Hs[i][i] != 0.0 ? u / Hs[i][i] : 0.0;
        }
    }
    __syncthreads();
    if (t < kd) st->y[t] = y[t];
    return f;
}

// The end of the Arnoldi step a round of queued steps closes with (sequence
// number seq - 1), when its third pass ran: the next step's first pass ends
// the step before it itself.
__device__ __forceinline__ void jf_column_tail_body(JfState* st, const double* part3, int nblk, double rtol, int seq) {
    __shared__ double nn[1];
    if (st->nd || st->umark != seq - 1 || st->phase != kJfArnoldi) return;
    const int k = st->ustep;
    jf_sums(part3, nblk, 1, nn);
    jf_column_end(st, k, nn[0], rtol, true);
}

__global__ void __launch_bounds__(256) jf_column_tail(JfState* st, const double* part3, int nblk, double rtol,
                                                      int seq, JfStatus* hs, int pub) {
    jf_column_tail_body(st, part3, nblk, rtol, seq);
    if (pub) {  // the last kernel of a queued round: the status for the host
        __syncthreads();
        jf_publish(st, hs);
    }
}

// Arnoldi pass 1: V_k = src / vnorm_k (stored; src = rhs or r at k = 0, the
// orthogonalised w after), w = J V_k, partials of <V_j, w>, j <= k.  The
// operator's |V_k| is taken as 1: V_k is a normalised vector (the reference
// recomputes its norm, 1 to rounding).  newton_it >= 0: queued right after
// Newton iteration newton_it's residual pass; if that pass ran (`nrun`,
// which only it writes), this is its cycle's first step and applies the
// pass's end first: every block works it out from the pass's partials
// `npart` (|rhs| = |F| is vnorm_0, and eps), block 0 also writes it to the
// state, the other blocks read none of what it writes, and a converged test
// stops them all.  Otherwise (the pass did not run: an earlier cycle is
// still going) it is an ordinary step, as the passes after it are.  An
// ordinary step whose sequence number `seq` follows that of the step whose
// third pass ran last (`umark`, written only by third passes) first ends
// that step (jf_column_end from its |w|^2 partials `part3`: every block,
// block 0 storing it) and then, unless the cycle ended there, runs the next.
template <class M>
__global__ void __launch_bounds__(256) jf_arnoldi1_kernel(const BoxDesc* __restrict__ boxes,
                                                          const long long* __restrict__ bstart, int nb, long long N,
                                                          const double* __restrict__ q, const double* __restrict__ S,
                                                          const double* __restrict__ rhs, const double* __restrict__ r,
                                                          double* __restrict__ w, double* const* __restrict__ V,
                                                          double dt, Params pp, double time, JfState* st,
                                                          double* part, const double* __restrict__ npart,
                                                          int newton_it, const double* __restrict__ part3,
                                                          int nblk3, double rtol, int seq) {
    using Sy = Sys<M>;
    using R = SourceRows<M>;
    constexpr int nq = Sy::nq, na = R::n;
    __shared__ double lds[4];
    __shared__ double tot[2];
    int k;
    double eps_num, vk;
    const double* __restrict__ bk;
    if (newton_it >= 0 && !st->nd && st->nrun == newton_it + 1) {
        const bool fused = blockIdx.x == 0 && jf_newton_fused(st, newton_it);
        jf_sums(npart, gridDim.x, 2, tot);
        const double resid = sqrt(tot[0]);
        const double r0 = newton_it == 0 ? resid : st->resid0;
        const bool conv = resid < 1e-12 || (r0 > 0.0 && resid < 1e-9 * r0) || st->bad_rows;
        if (blockIdx.x == 0) jf_newton_end(st, newton_it, tot[0], tot[1], fused);
        if (conv) return;
        k = 0;
        eps_num = 1e-7 * (1.0 + sqrt(tot[1]));
        vk = resid;
        bk = rhs;
    } else if (!st->nd && st->umark == seq - 1 && st->phase == kJfArnoldi) {
        // the step queued just before this one ran its third pass: its end
        // first (every block, block 0 storing it), then the next step
        const int kp = st->ustep;
        jf_sums(part3, nblk3, 1, tot);
        if (jf_column_end(st, kp, tot[0], rtol, blockIdx.x == 0)) return;
        k = kp + 1;
        eps_num = st->eps_num;
        vk = sqrt(tot[0]);
        bk = w;
    } else {
        if (st->nd || st->phase != kJfArnoldi) return;
        k = st->k;
        eps_num = st->eps_num;
        vk = st->vnorm[k];
        bk = k > 0 ? w : (st->src0 == 0 ? rhs : r);
    }
    double* __restrict__ vout = V[k];
    typename Sy::VP p;
    for (int i = 0; i < Sy::np; ++i) p(i, 0) = pp.v[i];
    double uc[kJfCellsPerThread][na], wc[kJfCellsPerThread][na];
    JF_CELLS_LOOP {
        double qc[nq], sb[na];
#pragma unroll
        for (int n = 0; n < nq; ++n) qc[n] = q[n * N + c];
#pragma unroll
        for (int a = 0; a < na; ++a) {
            uc[pc][a] = bk[a * N + c] / vk;
            sb[a] = S[a * N + c];
        }
        jf_apply_cell<M>(boxes, bstart, nb, c, qc, uc[pc], sb, eps_num, 1.0, dt, p, time, wc[pc]);
#pragma unroll
        for (int a = 0; a < na; ++a) {
            vout[a * N + c] = uc[pc][a];
            w[a * N + c] = wc[pc][a];
        }
    }
    for (int j = 0; j <= k; ++j) {
        const double* __restrict__ vj = V[j];
        double sacc = 0.0;
        JF_CELLS_LOOP {
#pragma unroll
            for (int a = 0; a < na; ++a) sacc += (j == k ? uc[pc][a] : vj[a * N + c]) * wc[pc][a];
        }
        const double v = jf_block_sum(sacc, lds);
        if (threadIdx.x == 0) part[(long long)j * gridDim.x + blockIdx.x] = v;
    }
}

// Arnoldi passes 2 (PASS 0) and 3 (PASS 1) over the len vector elements:
// the Gram-Schmidt coefficients h_j of the previous pass (h1 after pass 1,
// h2 after pass 2) summed from its per-block partials `pin` by every block
// in the same fixed order (so all blocks hold the same values; block 0
// keeps them in the state for the column's end), then w -= h_j V_j for
// j = 0..k in order, w written back, then partials of <V_j, w> (pass 2) or
// of |w|^2 (pass 3) into `pout`, a buffer other than `pin`.  The previous
// pass left `nprev` partials per coefficient (its grid); the grid of these
// passes covers the len elements exactly.
template <int PASS>
__global__ void __launch_bounds__(256) jf_arnoldi23_kernel(long long len, double* __restrict__ w,
                                                           double* const* __restrict__ V, JfState* st,
                                                           const double* __restrict__ pin, int nprev,
                                                           double* __restrict__ pout, int seq) {
    __shared__ double lds[4];
    __shared__ double hv[kJfRestart + 1];
    if (st->nd || st->phase != kJfArnoldi) return;
    const int k = st->k;
    jf_sums(pin, nprev, k + 1, hv);
    if (blockIdx.x == 0 && (int)threadIdx.x <= k) (PASS == 0 ? st->h1 : st->h2)[threadIdx.x] = hv[threadIdx.x];
    if (PASS == 1 && blockIdx.x == 0 && threadIdx.x == 0) {  // this step's end is due (read by later kernels)
        st->umark = seq;
        st->ustep = k;
    }
    constexpr int kE = kJfCellsPerThread * 4;  // elements per thread
    double we[kE];
    const long long base = (long long)blockIdx.x * kE * blockDim.x + threadIdx.x;
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const long long i = base + (long long)e * blockDim.x;
        we[e] = i < len ? w[i] : 0.0;
    }
    for (int j = 0; j <= k; ++j) {
        const double* __restrict__ vj = V[j];
        const double h = hv[j];
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            const long long i = base + (long long)e * blockDim.x;
            if (i < len) we[e] -= h * vj[i];
        }
    }
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const long long i = base + (long long)e * blockDim.x;
        if (i < len) w[i] = we[e];
    }
    const int m = PASS == 0 ? k + 1 : 1;
    for (int j = 0; j < m; ++j) {
        double sacc = 0.0;
        if (PASS == 0) {
            const double* __restrict__ vj = V[j];
#pragma unroll
            for (int e = 0; e < kE; ++e) {
                const long long i = base + (long long)e * blockDim.x;
                if (i < len) sacc += vj[i] * we[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < kE; ++e) sacc += we[e] * we[e];
        }
        const double v = jf_block_sum(sacc, lds);
        if (threadIdx.x == 0) pout[(long long)j * gridDim.x + blockIdx.x] = v;
    }
}

// x += y_j V_j, j in order (x = 0 in the first cycle), per cell; converged:
// Q += x on the source rows; else partials of |x|^2 for the restart residual.
template <class M>
__global__ void __launch_bounds__(256) jf_update_kernel(long long N, double* __restrict__ x, double* __restrict__ q,
                                                        double* const* __restrict__ V, const JfState* st,
                                                        double* part) {
    using R = SourceRows<M>;
    constexpr int na = R::n;
    __shared__ double lds[4];
    __shared__ double yv[kJfRestart];
    if (st->nd || st->phase != kJfUpdate) return;
    const int kd = st->kd, first = st->first, done = st->done;
    if ((int)threadIdx.x < kd) yv[threadIdx.x] = st->y[threadIdx.x];
    __syncthreads();
    double x2 = 0.0;
    JF_CELLS_LOOP {
#pragma unroll
        for (int a = 0; a < na; ++a) {
            const long long i = a * N + c;
            double xv = first ? 0.0 : x[i];
            for (int j = 0; j < kd; ++j) xv += yv[j] * V[j][i];
            if (done) {
                q[R::row(a) * N + c] += 1.0 * xv;
            } else {
                x[i] = xv;
                x2 += xv * xv;
            }
        }
    }
    if (done) return;
    const double v = jf_block_sum(x2, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__device__ __forceinline__ void jf_update_tail_body(JfState* st, const double* part, int nblk) {
    __shared__ double tot[1];
    if (st->nd || st->phase != kJfUpdate) return;
    if (st->done) {
        if (threadIdx.x == 0) jf_newton_done(st);
        return;
    }
    jf_sums(part, nblk, 1, tot);
    if (threadIdx.x == 0) {
        st->xnorm = sqrt(tot[0]);
        st->first = 0;
        st->phase = kJfRestartRes;
    }
}

__global__ void __launch_bounds__(256) jf_update_tail(JfState* st, const double* part, int nblk, JfStatus* hs,
                                                      int pub) {
    jf_update_tail_body(st, part, nblk);
    if (pub) {  // the last kernel of a queued round: the status for the host
        __syncthreads();
        jf_publish(st, hs);
    }
}

// Restart residual: r = rhs - J x, partials of |r|^2.
template <class M>
__global__ void __launch_bounds__(256) jf_restart_kernel(const BoxDesc* __restrict__ boxes,
                                                         const long long* __restrict__ bstart, int nb, long long N,
                                                         const double* __restrict__ q, const double* __restrict__ S,
                                                         const double* __restrict__ rhs, const double* __restrict__ x,
                                                         double* __restrict__ r, double dt, Params pp, double time,
                                                         const JfState* st, double* part) {
    using Sy = Sys<M>;
    using R = SourceRows<M>;
    constexpr int nq = Sy::nq, na = R::n;
    __shared__ double lds[4];
    if (st->nd || st->phase != kJfRestartRes) return;
    typename Sy::VP p;
    for (int i = 0; i < Sy::np; ++i) p(i, 0) = pp.v[i];
    const double eps_num = st->eps_num, xn = st->xnorm;
    double r2 = 0.0;
    JF_CELLS_LOOP {
        double qc[nq], xc[na], sb[na], wc[na];
#pragma unroll
        for (int n = 0; n < nq; ++n) qc[n] = q[n * N + c];
#pragma unroll
        for (int a = 0; a < na; ++a) {
            xc[a] = x[a * N + c];
            sb[a] = S[a * N + c];
        }
        jf_apply_cell<M>(boxes, bstart, nb, c, qc, xc, sb, eps_num, xn, dt, p, time, wc);
#pragma unroll
        for (int a = 0; a < na; ++a) {
            const double rr = 1.0 * rhs[a * N + c] + (-1.0) * wc[a];
            r[a * N + c] = rr;
            r2 += rr * rr;
        }
    }
    const double v = jf_block_sum(r2, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Converged: Q update; else a new cycle from r.
__global__ void __launch_bounds__(256) jf_restart_tail(JfState* st, const double* part, int nblk, double rtol) {
    __shared__ double tot[1];
    if (st->nd || st->phase != kJfRestartRes) return;
    jf_sums(part, nblk, 1, tot);
    const double rn = sqrt(tot[0]);
    const bool fin = rn < st->r0 * rtol || rn == 0.0 || st->its >= kJfMaxIter;
    if (!fin) jf_clear_cycle(st);
    if (threadIdx.x != 0) return;
    if (fin) {
        st->phase = kJfQUpdate;
        return;
    }
    st->vnorm[0] = rn;
    st->src0 = 1;
    st->k = 0;
    st->kd = 0;
    st->done = 0;
    st->its0 = st->its;
    st->g[0] = rn;
    st->gp[0] = rn;
    st->phase = kJfArnoldi;
}

// Q += x on the source rows after a restart residual met the tolerance.
template <class M>
__global__ void __launch_bounds__(256) jf_qupdate_kernel(long long N, double* __restrict__ q,
                                                         const double* __restrict__ x, const JfState* st) {
    using R = SourceRows<M>;
    if (st->nd || st->phase != kJfQUpdate) return;
    JF_CELLS_LOOP {
#pragma unroll
        for (int a = 0; a < R::n; ++a) q[R::row(a) * N + c] += 1.0 * x[a * N + c];
    }
}

__global__ void jf_qupdate_tail(JfState* st) {
    if (st->nd || st->phase != kJfQUpdate || threadIdx.x != 0) return;
    jf_newton_done(st);
}

// End of a step whose stages solved the source globally: Heun with buffer
// `old` (order 2; old < 0 at order 1 or when already applied), end-of-step
// UpdateState and the CFL max, in place on buffer `dst` — the arithmetic of
// the fused stage-2 epilogue.  heun_only: Heun alone, counting the MOOD
// troubled cells of the candidate (h < 0, NaN or +-inf) into the group.
template <class M, class N>
__global__ void __launch_bounds__(NT) finish_step_kernel(const BoxDesc* __restrict__ boxes,
                                                         const WorkItem* __restrict__ items, int nitems, int dst,
                                                         int old, Params pp, int clamp,
                                                         GroupState* __restrict__ gs, int txw, int heun_only) {
    using S = Sys<M>;
    constexpr int nq = S::nq, dim = S::dim;
    typename S::VP p;
    for (int k = 0; k < S::np; ++k) p(k, 0) = pp.v[k];
    ZAMR_ITEM_LOOP(nitems) {
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        GroupState* G = gs + B.group;
        const int c = w.i0 + t;
        const double time = G->time;
        const double rdx = 1.0 / B.dx0, rdy = 1.0 / B.dx1;
        double lam = 0.0;
        int nbad = 0;
        if (t < txw && c <= B.hi0 && G->active) {
            double* __restrict__ qd = qsel(B, dst);
            const double* __restrict__ qo = old >= 0 ? qsel(B, old) : nullptr;
            for (int j = w.j0; j < w.j0 + w.nj; ++j) {
                const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
                typename S::VQ qn;
#pragma unroll
                for (int n = 0; n < nq; ++n) qn(n, 0) = qd[idx + n * B.sc];
                if (qo) {
#pragma unroll
                    for (int n = 0; n < nq; ++n) qn(n, 0) = 0.5 * qo[idx + n * B.sc] + 0.5 * qn(n, 0);
                }
                if (heun_only) {
                    nbad += (!(qn(1, 0) >= 0.0) || !(qn(1, 0) < 1.0e300)) ? 1 : 0;
#pragma unroll
                    for (int n = 0; n < nq; ++n) qd[idx + n * B.sc] = qn(n, 0);
                    continue;
                }
                qn = S::update_state(qn, p, clamp);
                const typename S::VX X = S::center(B, c, j);
                const typename S::VA af = S::aux_at(B, c, j, qn, p, time, X);
#pragma unroll
                for (int d = 0; d < dim; ++d) {
                    typename S::VN nh{};
                    nh(d, 0) = 1.0;
                    const double ev = N::local_max_abs_eigenvalue(qn, af, p, nh)(0, 0);
                    const double inv = kBitwise ? ev / (d == 0 ? B.dx0 : B.dx1) : ev * (d == 0 ? rdx : rdy);
                    if (inv > lam) lam = inv;
                }
#pragma unroll
                for (int n = 0; n < nq; ++n) qd[idx + n * B.sc] = qn(n, 0);
            }
        }
        if (heun_only) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) nbad += __shfl_xor(nbad, off, 64);
            if (t == 0 && nbad > 0) atomicAdd(&G->troubled, nbad);
            continue;
        }
        lam = wave_max(lam);
        if (t == 0) group_max(&G->maxinv_bits, lam);
    }
}

// Host side: the vectors and the queueing.  The host queues Newton
// iterations ahead -- each the residual pass (which first applies the
// previous iteration's converged correction) and a round of Arnoldi steps
// (as many as that iteration took in the last solve) -- as far as the last
// solve went, its converging residual test included, then waits once.
// Kernels act only in their phase (a Newton iteration only after the
// previous one finished or converged), so over-queued work returns at once
// and under-queued work is completed by further rounds (more Arnoldi steps
// and the cycle's update, a restart, the next iteration) after reading the
// state back, which one small kernel copies to host-visible memory at the
// end of each round.
template <class M>
class Jfnk {
   public:
    static constexpr int nq = M::n_dof_q;
    static constexpr int na = SourceRows<M>::n;
    static constexpr int kRestart = kJfRestart;

    hipStream_t stream = nullptr;
    const BoxDesc* boxes = nullptr;
    int nb = 0;
    long long N = 0, len = 0;  // cells, doubles per Krylov vector (source rows)
    int nblk = 0;              // blocks of a cell pass
    int nblk_e = 0;            // blocks of an element pass (Gram-Schmidt passes 2 and 3)
    long long* d_bstart = nullptr;
    double *q = nullptr, *qexp = nullptr, *S = nullptr, *rhs = nullptr, *x = nullptr, *w = nullptr, *r = nullptr;
    std::vector<double*> V;   // basis vectors, kRestart + 1
    double** d_V = nullptr;   // their device pointer list
    double* d_part = nullptr; // (kRestart + 1) x nblk partial sums
    double* d_part2 = nullptr; // the same, for the pass that reads d_part's
    double* d_npart = nullptr; // 2 x nblk partial sums of a Newton residual pass
    double* d_part3 = nullptr; // nblk_e partial sums of |w|^2, the third Arnoldi pass
    JfState* d_st = nullptr;
    JfStatus* h_status = nullptr;  // host-visible, written by the kernels
    long newton_its = 0, gmres_its = 0;  // last solve
    long long total_solves = 0, total_newton = 0, total_gmres = 0;
    int chunk = 2;        // Arnoldi steps queued in the first round
    int newton_pred = 0;  // Newton iterations of the last solve
    std::vector<int> steps_pred;  // Arnoldi steps of each of those iterations

    ~Jfnk() { release(); }
    void release() {
        for (double* p : {q, qexp, S, rhs, x, w, r, d_part, d_part2, d_npart, d_part3})
            if (p) hipFree(p);
        for (double* p : V)
            if (p) hipFree(p);
        V.clear();
        if (d_V) hipFree(d_V);
        if (d_st) hipFree(d_st);
        if (d_bstart) hipFree(d_bstart);
        if (h_status) hipHostFree(h_status);
        q = qexp = S = rhs = x = w = r = d_part = d_part2 = d_npart = d_part3 = nullptr;
        d_V = nullptr;
        d_st = nullptr;
        h_status = nullptr;
        d_bstart = nullptr;
    }

    // Vectors for a level with the given boxes (lo, hi per box, in order).
    void setup(hipStream_t s, const BoxDesc* d_boxes, const std::vector<std::array<int, 4>>& bx) {
        release();
        stream = s;
        boxes = d_boxes;
        nb = (int)bx.size();
        std::vector<long long> st(nb + 1, 0);
        for (int b = 0; b < nb; ++b) st[b + 1] = st[b] + (long long)(bx[b][2] - bx[b][0] + 1) * (bx[b][3] - bx[b][1] + 1);
        N = st[nb];
        len = N * na;
        // a cell pass covers kJfCellsPerThread cells per thread, an element
        // pass 4 kJfCellsPerThread elements per thread, each over a grid of
        // its own (the partial sums a pass leaves are one per block of it)
        const long long per_blk_cells = (long long)kJfCellsPerThread * 256;
        const long long per_blk_elems = (long long)kJfCellsPerThread * 4 * 256;
        nblk = (int)std::max<long long>((N + per_blk_cells - 1) / per_blk_cells, 1);
        nblk_e = (int)std::max<long long>((len + per_blk_elems - 1) / per_blk_elems, 1);
        const int nmax = std::max(nblk, nblk_e);
        ZAMR_HIP(hipMalloc(&d_bstart, st.size() * sizeof(long long)));
        ZAMR_HIP(hipMemcpy(d_bstart, st.data(), st.size() * sizeof(long long), hipMemcpyHostToDevice));
        ZAMR_HIP(hipMalloc(&q, (size_t)N * nq * sizeof(double)));
        for (double** p : {&qexp, &S, &rhs, &x, &w, &r}) ZAMR_HIP(hipMalloc(p, len * sizeof(double)));
        V.assign(kRestart + 1, nullptr);
        for (auto& v : V) ZAMR_HIP(hipMalloc(&v, len * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_V, V.size() * sizeof(double*)));
        ZAMR_HIP(hipMemcpy(d_V, V.data(), V.size() * sizeof(double*), hipMemcpyHostToDevice));
        ZAMR_HIP(hipMalloc(&d_part, (size_t)(kRestart + 1) * nmax * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_part2, (size_t)(kRestart + 1) * nmax * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_npart, (size_t)2 * nblk * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_part3, (size_t)nblk_e * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_st, sizeof(JfState)));
        ZAMR_HIP(hipHostMalloc(&h_status, sizeof(JfStatus), hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(h_status, 0, sizeof(JfStatus));
    }

    static unsigned g1(long long n) { return (unsigned)((n + 255) / 256); }

    // The solve on the valid cells of buffer `buf` (Qexp in, Q out).
    void solve(int buf, double dt, double time, const Params& p) {
        constexpr double rtol = 1e-4;
        const dim3 G(nblk), Ge(nblk_e), T(256), one(1);
        JfStatus* hs = nullptr;
        ZAMR_HIP(hipHostGetDevicePointer((void**)&hs, h_status, 0));
        double* const* dV = (double* const*)d_V;
        const double* cpart = d_part;
        hipLaunchKernelGGL((jf_move_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, buf, q, 0,
                           qexp, d_st);
        int seq = 1;  // sequence numbers of the queued Arnoldi steps and their round-closing kernels
        ZAMR_HIP(hipGetLastError());
        // `steps` Arnoldi steps, the first of them right after Newton
        // iteration newton_it's residual pass (-1: continuing a cycle);
        // with_update: then the cycle's update (a converged cycle's is
        // otherwise applied by the next Newton pass)
        // Every step gets a sequence number; a step ends the one before it
        // (when that one ran), and a one-block kernel ends the round's last.
        // last: the round's final kernel also copies the status for the host.
        auto arnoldi = [&](int steps, bool with_update, int newton_it, bool last) {
            for (int k = 0; k < steps; ++k) {
                const int u = seq++;
                hipLaunchKernelGGL((jf_arnoldi1_kernel<M>), G, T, 0, stream, boxes, d_bstart, nb, N, (const double*)q,
                                   (const double*)S, (const double*)rhs, (const double*)r, w, dV, dt, p, time, d_st,
                                   d_part, (const double*)d_npart, k == 0 ? newton_it : -1, (const double*)d_part3,
                                   nblk_e, rtol, u);
                // pass 1 -> d_part -> pass 2 -> d_part2 -> pass 3 -> d_part3
                hipLaunchKernelGGL(jf_arnoldi23_kernel<0>, Ge, T, 0, stream, len, w, dV, d_st, cpart, nblk, d_part2,
                                   u);
                hipLaunchKernelGGL(jf_arnoldi23_kernel<1>, Ge, T, 0, stream, len, w, dV, d_st,
                                   (const double*)d_part2, nblk_e, d_part3, u);
            }
            hipLaunchKernelGGL(jf_column_tail, one, T, 0, stream, d_st, (const double*)d_part3, nblk_e, rtol, seq++, hs,
                               (int)(last && !with_update));
            if (with_update) {
                hipLaunchKernelGGL((jf_update_kernel<M>), G, T, 0, stream, N, x, q, dV, (const JfState*)d_st, d_part);
                hipLaunchKernelGGL(jf_update_tail, one, T, 0, stream, d_st, cpart, nblk, hs, (int)last);
            }
            ZAMR_HIP(hipGetLastError());
        };
        // Newton iteration `it`: the residual pass, then a round of Arnoldi
        // steps (the first applies the pass's end); queued ahead as far as
        // the last solve went (its converging residual test included, with
        // the pass's end on its own), then read back once
        auto iteration = [&](int it, int steps, bool last) {
            hipLaunchKernelGGL((jf_newton_kernel<M>), G, T, 0, stream, boxes, d_bstart, nb, N, q, (const double*)qexp,
                               (const double*)x, dV, S, rhs, dt, p, time, d_st, d_npart, it);
            arnoldi(steps, false, it, last);
        };
        if ((int)steps_pred.size() == newton_pred && newton_pred > 0 && newton_pred < kJfNewtonMax) {
            // as the last solve went: its iterations with their own Arnoldi
            // step counts, then the residual pass that found it converged
            // (a solve that needs more is finished by the rounds below)
            for (int it = 0; it < newton_pred; ++it) iteration(it, std::max(steps_pred[it], 1), false);
            hipLaunchKernelGGL((jf_newton_kernel<M>), G, T, 0, stream, boxes, d_bstart, nb, N, q, (const double*)qexp,
                               (const double*)x, dV, S, rhs, dt, p, time, d_st, d_npart, newton_pred);
            hipLaunchKernelGGL(jf_newton_tail, one, T, 0, stream, d_st, (const double*)d_npart, nblk, newton_pred, hs, 1);
            ZAMR_HIP(hipGetLastError());
        } else {
            const int ahead = std::min(kJfNewtonMax, std::max(newton_pred, 1) + 1);
            for (int it = 0; it < ahead; ++it) iteration(it, std::max(chunk, 2), it == ahead - 1);
        }
        // the one wait of a round (its final kernel has copied the status)
        auto wait = [&]() {
            ZAMR_HIP(hipGetLastError());
            ZAMR_HIP(hipStreamSynchronize(stream));
        };
        wait();
        long long rounds = 0;
        for (;;) {
            if (h_status->nd || h_status->newton_its >= kJfNewtonMax) break;
            if (++rounds > kJfMaxIter || h_status->its >= kJfMaxIter)
                throw Error(4, "internal: the global source solve's GMRES did not settle");
            if (h_status->phase == kJfFinished) {  // the next iteration was not queued, or ran early
                iteration((int)h_status->newton_its, std::max(chunk, 2), true);
            } else {  // the running cycle needs more steps, a restart, or its update
                hipLaunchKernelGGL((jf_restart_kernel<M>), G, T, 0, stream, boxes, d_bstart, nb, N, (const double*)q,
                                   (const double*)S, (const double*)rhs, (const double*)x, r, dt, p, time,
                                   (const JfState*)d_st, d_part);
                hipLaunchKernelGGL(jf_restart_tail, one, T, 0, stream, d_st, cpart, nblk, rtol);
                hipLaunchKernelGGL((jf_qupdate_kernel<M>), G, T, 0, stream, N, q, (const double*)x,
                                   (const JfState*)d_st);
                hipLaunchKernelGGL(jf_qupdate_tail, one, dim3(64), 0, stream, d_st);
                arnoldi(4, true, -1, true);
            }
            wait();
        }
        const int it = (int)h_status->newton_its;
        if (it > 0) chunk = (int)std::min<long long>(std::max<long long>(h_status->gmres_total / it, 2), kRestart);
        if (h_status->bad_rows)
            throw Error(4, "the global source solve: a row the model declares source-free has a nonzero source");
        newton_pred = it;
        steps_pred.assign(h_status->its_hist, h_status->its_hist + std::min(it, 16));
        hipLaunchKernelGGL((jf_move_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, buf, q, 1,
                           (double*)nullptr, (JfState*)nullptr);
        ZAMR_HIP(hipGetLastError());
        newton_its = h_status->newton_its;
        gmres_its = (long)h_status->gmres_total;
        ++total_solves;
        total_newton += newton_its;
        total_gmres += gmres_its;
    }
};

}  // namespace zamr
