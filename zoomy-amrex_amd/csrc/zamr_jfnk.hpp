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
// refresh does not enter).  Gram-Schmidt is classical, applied twice (one
// fused pass over the basis per sweep).
//
// Vectors hold the valid cells of every box of the level, box by box, row
// by row, component-major over the level: v[n * N + c].  Vector work is
// HBM-streaming (a dot or axpy over k basis vectors reads k + 1 vectors);
// the control flow (Newton, Arnoldi, Givens) runs on the host from the dot
// products, which are reduced deterministically (fixed partials, summed in
// order).
#include <array>
#include <cmath>
#include <vector>

#include "zamr_error.hpp"
#include "zamr_kernels.hpp"

namespace zamr {

constexpr int kJfnkBlocks = 1024;  // partial sums per reduction

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

// Valid cells of buffer `buf` -> compact (gather), or compact -> buffer (scatter).
template <class M>
__global__ void __launch_bounds__(256) jf_move_kernel(const BoxDesc* __restrict__ boxes,
                                                      const long long* __restrict__ bstart, int nb, long long N,
                                                      int buf, double* __restrict__ v, int to_box) {
    constexpr int nq = M::n_dof_q;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const JfCell e = jf_cell(boxes, bstart, nb, c);
    const BoxDesc& B = boxes[e.b];
    double* q = buf == 0 ? B.q[0] : (buf == 1 ? B.q[1] : B.q[2]);
    const long long idx = (long long)(e.i - B.gl[0]) + (long long)(e.j - B.gl[1]) * B.sy;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        if (to_box) q[idx + n * B.sc] = v[n * N + c];
        else v[n * N + c] = q[idx + n * B.sc];
    }
}

// S(q) of every cell (mode x == nullptr: out = S(q)), or the Jacobian
// product out = (x + (-dt/eps) S(q + eps x)) + (dt/eps) Sbase with
// eps = eps_num / |x|, |x|^2 read from the device (*x2); |x| = 0 gives 0.
template <class M>
__global__ void __launch_bounds__(256) jf_source_kernel(const BoxDesc* __restrict__ boxes,
                                                        const long long* __restrict__ bstart, int nb, long long N,
                                                        const double* __restrict__ q, const double* __restrict__ x,
                                                        double eps_num, const double* __restrict__ x2, double dt,
                                                        const double* __restrict__ Sbase, Params pp, double time,
                                                        double* __restrict__ out) {
    using S = Sys<M>;
    constexpr int nq = S::nq;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    double eps = 0.0;
    if (x) {
        const double nx = sqrt(*x2);
        if (nx == 0.0) {
#pragma unroll
            for (int n = 0; n < nq; ++n) out[n * N + c] = 0.0;
            return;
        }
        eps = eps_num / nx;
    }
    typename S::VP p;
    for (int k = 0; k < S::np; ++k) p(k, 0) = pp.v[k];
    const JfCell e = jf_cell(boxes, bstart, nb, c);
    const BoxView B = view(boxes, e.b);
    typename S::VQ qq;
#pragma unroll
    for (int n = 0; n < nq; ++n) qq(n, 0) = x ? 1.0 * q[n * N + c] + eps * x[n * N + c] : q[n * N + c];
    const typename S::VX X = S::center(B, e.i, e.j);
    typename S::VA a{};
    a = M::update_aux_variables(qq, a, p, time, X);
    const typename S::VQ s = M::source(qq, a, p, time, X);
    if (x) {
        const double r = dt / eps;
#pragma unroll
        for (int n = 0; n < nq; ++n) out[n * N + c] = (1.0 * x[n * N + c] + (-r) * s(n, 0)) + r * Sbase[n * N + c];
    } else {
#pragma unroll
        for (int n = 0; n < nq; ++n) out[n * N + c] = s(n, 0);
    }
}

// out = a x + b y (elementwise over len).
__global__ void __launch_bounds__(256) jf_lincomb_kernel(long long len, double a, const double* __restrict__ x,
                                                         double b, const double* __restrict__ y,
                                                         double* __restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < len) out[k] = a * x[k] + b * y[k];
}

// w += sign * sum_j coef[j] V[j]  (coef on the device, m <= 31).
__global__ void __launch_bounds__(256) jf_multi_axpy_kernel(long long len, int m, double* const* __restrict__ V,
                                                            const double* __restrict__ coef, double sign,
                                                            double* __restrict__ w) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= len) return;
    double acc = w[k];
    for (int j = 0; j < m; ++j) acc += (sign * coef[j]) * V[j][k];
    w[k] = acc;
}

// out = x / sqrt(*x2) when *x2 > 0.
__global__ void __launch_bounds__(256) jf_normalize_kernel(long long len, const double* __restrict__ x,
                                                           const double* __restrict__ x2, double* __restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const double n2 = *x2;
    if (k < len && n2 > 0.0) out[k] = x[k] / sqrt(n2);
}

// partial[j * kJfnkBlocks + blk] = this block's share of <V[j], w>.
__global__ void __launch_bounds__(256) jf_dots_kernel(long long len, int m, double* const* __restrict__ V,
                                                      const double* __restrict__ w, double* __restrict__ partial) {
    __shared__ double red[256];
    for (int j = 0; j < m; ++j) {
        const double* __restrict__ v = V[j];
        double s = 0.0;
        for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < len;
             k += (long long)gridDim.x * blockDim.x)
            s += v[k] * w[k];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[j * kJfnkBlocks + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// partial[blk] = this block's share of <a, b> (pointers passed directly).
__global__ void __launch_bounds__(256) jf_dot_kernel(long long len, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < len;
         k += (long long)gridDim.x * blockDim.x)
        s += a[k] * b[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[j] = the sum of partial[j * kJfnkBlocks + 0 .. kJfnkBlocks), one block
// per j: each thread adds its run of kJfnkBlocks / 256 partials in order, then
// a fixed tree over the threads (deterministic, and ~10x faster than one
// thread walking all the partials).
__global__ void __launch_bounds__(256) jf_sum_partials_kernel(int m, const double* __restrict__ partial,
                                                              double* __restrict__ out) {
    static_assert(kJfnkBlocks % 256 == 0, "partials per thread");
    __shared__ double red[256];
    const int j = blockIdx.x;
    if (j >= m) return;
    constexpr int per = kJfnkBlocks / 256;
    const double* __restrict__ p = partial + (long long)j * kJfnkBlocks + threadIdx.x * per;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < per; ++b) s += p[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// End of a step whose stages solved the source globally: Heun with buffer
// `old` (order 2; old < 0 at order 1), end-of-step UpdateState and the CFL
// max, in place on buffer `dst` — the arithmetic of the fused stage-2
// epilogue.
template <class M, class N>
__global__ void __launch_bounds__(NT) finish_step_kernel(const BoxDesc* __restrict__ boxes,
                                                         const WorkItem* __restrict__ items, int nitems, int dst,
                                                         int old, Params pp, int clamp,
                                                         GroupState* __restrict__ gs, int txw) {
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
        lam = wave_max(lam);
        if (t == 0 && lam > 0.0) atomicMax(&G->maxinv_bits, (unsigned long long)__double_as_longlong(lam));
    }
}

// Host side: the vectors, Newton and GMRES.  Scalars that only steer device
// work (Gram-Schmidt coefficients, norms for eps and normalisation) stay on
// the device; one transfer per Arnoldi step brings down what the host needs
// for the Hessenberg matrix and the convergence test.
template <class M>
class Jfnk {
   public:
    static constexpr int nq = M::n_dof_q;
    static constexpr int kRestart = 30;
    static constexpr long kMaxIter = 2000;
    // device scalar slots
    static constexpr int kH1 = 0, kH2 = 32, kHN = 64, kNX = 65, kF2 = 66, kQ2 = 67, kR2 = 68, kSlots = 72;

    hipStream_t stream = nullptr;
    const BoxDesc* boxes = nullptr;
    int nb = 0;
    long long N = 0, len = 0;  // cells, doubles per vector
    long long* d_bstart = nullptr;
    double *q = nullptr, *qexp = nullptr, *S = nullptr, *F = nullptr, *rhs = nullptr, *x = nullptr, *w = nullptr,
           *r = nullptr;
    std::vector<double*> V;        // basis vectors, kRestart + 1
    double* d_partial = nullptr;   // (kRestart + 2) x kJfnkBlocks
    double* d_sc = nullptr;        // kSlots device scalars
    double* d_y = nullptr;         // kRestart coefficients (uploaded)
    double** d_Vsel = nullptr;     // pointer lists, one row per use
    double* h_pin = nullptr;       // pinned: scalars down, coefficients up
    long newton_its = 0, gmres_its = 0;          // last solve
    long long total_solves = 0, total_newton = 0, total_gmres = 0;

    ~Jfnk() { release(); }
    void release() {
        for (double* p : {q, qexp, S, F, rhs, x, w, r, d_partial, d_sc, d_y})
            if (p) hipFree(p);
        for (double* p : V)
            if (p) hipFree(p);
        V.clear();
        if (d_Vsel) hipFree(d_Vsel);
        if (d_bstart) hipFree(d_bstart);
        if (h_pin) hipHostFree(h_pin);
        q = qexp = S = F = rhs = x = w = r = d_partial = d_sc = d_y = nullptr;
        d_Vsel = nullptr;
        d_bstart = nullptr;
        h_pin = nullptr;
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
        len = N * nq;
        ZAMR_HIP(hipMalloc(&d_bstart, st.size() * sizeof(long long)));
        ZAMR_HIP(hipMemcpy(d_bstart, st.data(), st.size() * sizeof(long long), hipMemcpyHostToDevice));
        for (double** p : {&q, &qexp, &S, &F, &rhs, &x, &w, &r}) ZAMR_HIP(hipMalloc(p, len * sizeof(double)));
        V.assign(kRestart + 1, nullptr);
        for (auto& v : V) ZAMR_HIP(hipMalloc(&v, len * sizeof(double)));
        // pointer-list rows: row k (k <= kRestart) = V[0..k]
        ZAMR_HIP(hipMalloc(&d_Vsel, (size_t)(kRestart + 1) * (kRestart + 1) * sizeof(double*)));
        std::vector<double*> rows((size_t)(kRestart + 1) * (kRestart + 1), nullptr);
        for (int k = 0; k <= kRestart; ++k)
            for (int j = 0; j <= k; ++j) rows[(size_t)k * (kRestart + 1) + j] = V[j];
        ZAMR_HIP(hipMemcpy(d_Vsel, rows.data(), rows.size() * sizeof(double*), hipMemcpyHostToDevice));
        ZAMR_HIP(hipMalloc(&d_partial, (kRestart + 2) * (size_t)kJfnkBlocks * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_sc, kSlots * sizeof(double)));
        ZAMR_HIP(hipMalloc(&d_y, kRestart * sizeof(double)));
        ZAMR_HIP(hipHostMalloc(&h_pin, (kSlots + kRestart) * sizeof(double)));
    }

    static unsigned g1(long long n) { return (unsigned)((n + 255) / 256); }
    double* const* basis(int k) const { return d_Vsel + (size_t)k * (kRestart + 1); }  // V[0..k]

    // out[j] = <vs[j], y> for the m vectors of a pointer row, on the device.
    void dots_dev(double* const* vs, int m, const double* y, double* out) {
        if (m <= 0) return;
        hipLaunchKernelGGL(jf_dots_kernel, dim3(kJfnkBlocks), dim3(256), 0, stream, len, m, vs, y, d_partial);
        hipLaunchKernelGGL(jf_sum_partials_kernel, dim3((unsigned)m), dim3(256), 0, stream, m, d_partial, out);
        ZAMR_HIP(hipGetLastError());
    }
    // <y, y> into a device slot.
    void norm2_dev(const double* y, double* out) {
        hipLaunchKernelGGL(jf_dot_kernel, dim3(kJfnkBlocks), dim3(256), 0, stream, len, y, y, d_partial);
        hipLaunchKernelGGL(jf_sum_partials_kernel, dim3(1), dim3(256), 0, stream, 1, d_partial, out);
        ZAMR_HIP(hipGetLastError());
    }
    // Device scalars [first, first + n) to the host.
    void fetch(int first, int n) {
        ZAMR_HIP(hipMemcpyAsync(h_pin + first, d_sc + first, n * sizeof(double), hipMemcpyDeviceToHost, stream));
        ZAMR_HIP(hipStreamSynchronize(stream));
    }
    void lincomb(double* out, double a, const double* xx, double b, const double* yy) {
        hipLaunchKernelGGL(jf_lincomb_kernel, dim3(g1(len)), dim3(256), 0, stream, len, a, xx, b, yy, out);
        ZAMR_HIP(hipGetLastError());
    }

    // out = J v at the frozen iterate (q, S); |v|^2 is computed into slot kNX.
    void apply(const double* v, double* out, double dt, double nrmQ, const Params& p, double time) {
        norm2_dev(v, d_sc + kNX);
        hipLaunchKernelGGL((jf_source_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, q, v,
                           1e-7 * (1.0 + nrmQ), (const double*)(d_sc + kNX), dt, S, p, time, out);
        ZAMR_HIP(hipGetLastError());
    }

    // x = GMRES(J, rhs) to rtol (zero initial guess), |rhs| = r0 known.
    long gmres(double r0, double dt, double nrmQ, const Params& p, double time, double rtol) {
        ZAMR_HIP(hipMemsetAsync(x, 0, len * sizeof(double), stream));
        if (r0 == 0.0) return 0;
        auto conv = [&](double rr) { return rr < r0 * rtol; };
        const double* rr = rhs;  // the residual of x = 0
        double rnorm = r0;
        long its = 0;
        while (its < kMaxIter) {
            lincomb(V[0], 1.0 / rnorm, rr, 0.0, rr);
            std::vector<double> g(kRestart + 1, 0.0), cs(kRestart, 0.0), sn(kRestart, 0.0);
            std::vector<std::vector<double>> H(kRestart + 1, std::vector<double>(kRestart, 0.0));
            g[0] = rnorm;
            int kd = 0;
            bool done = false;
            for (int k = 0; k < kRestart; ++k) {
                apply(V[k], w, dt, nrmQ, p, time);
                // classical Gram-Schmidt, twice, coefficients kept on the device
                dots_dev(basis(k), k + 1, w, d_sc + kH1);
                hipLaunchKernelGGL(jf_multi_axpy_kernel, dim3(g1(len)), dim3(256), 0, stream, len, k + 1, basis(k),
                                   (const double*)(d_sc + kH1), -1.0, w);
                dots_dev(basis(k), k + 1, w, d_sc + kH2);
                hipLaunchKernelGGL(jf_multi_axpy_kernel, dim3(g1(len)), dim3(256), 0, stream, len, k + 1, basis(k),
                                   (const double*)(d_sc + kH2), -1.0, w);
                norm2_dev(w, d_sc + kHN);
                hipLaunchKernelGGL(jf_normalize_kernel, dim3(g1(len)), dim3(256), 0, stream, len, w,
                                   (const double*)(d_sc + kHN), V[k + 1]);
                ZAMR_HIP(hipGetLastError());
                fetch(0, kHN + 1);
                for (int j = 0; j <= k; ++j) H[j][k] = h_pin[kH1 + j] + h_pin[kH2 + j];
                const double hn = std::sqrt(h_pin[kHN]);
                H[k + 1][k] = hn;
                for (int j = 0; j < k; ++j) {
                    const double t = cs[j] * H[j][k] + sn[j] * H[j + 1][k];
                    H[j + 1][k] = -sn[j] * H[j][k] + cs[j] * H[j + 1][k];
                    H[j][k] = t;
                }
                const double den = std::hypot(H[k][k], H[k + 1][k]);
                cs[k] = den > 0.0 ? H[k][k] / den : 1.0;
                sn[k] = den > 0.0 ? H[k + 1][k] / den : 0.0;
                H[k][k] = den;
                H[k + 1][k] = 0.0;
                g[k + 1] = -sn[k] * g[k];
                g[k] = cs[k] * g[k];
                ++its;
                kd = k + 1;
                if (conv(std::fabs(g[k + 1])) || hn == 0.0 || its >= kMaxIter) {
                    done = conv(std::fabs(g[k + 1]));
                    break;
                }
            }
            std::vector<double> y(kd, 0.0);
            for (int i = kd - 1; i >= 0; --i) {
                double t = g[i];
                for (int j = i + 1; j < kd; ++j) t -= H[i][j] * y[j];
                y[i] = H[i][i] != 0.0 ? t / H[i][i] : 0.0;
            }
            for (int j = 0; j < kd; ++j) h_pin[kSlots + j] = y[j];
            ZAMR_HIP(hipMemcpyAsync(d_y, h_pin + kSlots, kd * sizeof(double), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(jf_multi_axpy_kernel, dim3(g1(len)), dim3(256), 0, stream, len, kd, basis(kd - 1),
                               (const double*)d_y, 1.0, x);
            ZAMR_HIP(hipGetLastError());
            ZAMR_HIP(hipStreamSynchronize(stream));  // the pinned coefficients are reused
            if (done) break;
            apply(x, w, dt, nrmQ, p, time);
            lincomb(r, 1.0, rhs, -1.0, w);
            rr = r;
            norm2_dev(r, d_sc + kR2);
            fetch(kR2, 1);
            rnorm = std::sqrt(h_pin[kR2]);
            if (conv(rnorm) || rnorm == 0.0) break;
        }
        return its;
    }

    // The solve on the valid cells of buffer `buf` (Qexp in, Q out).
    void solve(int buf, double dt, double time, const Params& p) {
        hipLaunchKernelGGL((jf_move_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, buf, q, 0);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipMemcpyAsync(qexp, q, len * sizeof(double), hipMemcpyDeviceToDevice, stream));
        double resid0 = -1.0;
        newton_its = 0;
        gmres_its = 0;
        for (int it = 0; it < 12; ++it) {
            hipLaunchKernelGGL((jf_source_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, q,
                               (const double*)nullptr, 0.0, (const double*)nullptr, 0.0, (const double*)nullptr, p,
                               time, S);
            ZAMR_HIP(hipGetLastError());
            lincomb(F, 1.0, q, -1.0, qexp);
            lincomb(F, 1.0, F, -dt, S);
            norm2_dev(F, d_sc + kF2);
            norm2_dev(q, d_sc + kQ2);
            fetch(kF2, 2);
            const double resid = std::sqrt(h_pin[kF2]);
            if (it == 0) resid0 = resid;
            if (resid < 1e-12 || (resid0 > 0 && resid < 1e-9 * resid0)) break;
            lincomb(rhs, -1.0, F, 0.0, F);
            const double nrmQ = std::sqrt(h_pin[kQ2]);
            gmres_its += gmres(resid, dt, nrmQ, p, time, 1e-4);
            lincomb(q, 1.0, q, 1.0, x);
            ++newton_its;
        }
        hipLaunchKernelGGL((jf_move_kernel<M>), dim3(g1(N)), dim3(256), 0, stream, boxes, d_bstart, nb, N, buf, q, 1);
        ZAMR_HIP(hipGetLastError());
        ++total_solves;
        total_newton += newton_its;
        total_gmres += gmres_its;
    }
};

}  // namespace zamr
