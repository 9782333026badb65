#pragma once
// Reverse mode (adjoint) of the frozen-dt SSP-RK2 advance on one box.
//
// Forward, per step with a fixed dt (stage map Phi):
//   V  = U + dt RHS(U)                 flux divergence of make_rhs.H
//   W  = implicit_source(V, aux(U))    point-implicit source, stage-start aux
//   U1 = Phi(U0),  U2 = U0/2 + Phi(U1)/2
// Backward, per step: Ub1 = Phi'(U1)^T (Ub2/2), then
// Ub0 = Phi'(U0)^T Ub1 + Ub2/2, where
//   Phi'(U)^T Wb = Vb + [d aux/dU]^T ab + RHS'(U)^T (dt Vb)
// and (Vb, ab, nb) is the reverse of the point-implicit source at each cell.
//
// One fused kernel per reverse stage (adjoint_stage_kernel): it recomputes
// the face states of the forward stage from the stored stage input, applies
// the model's hand-written face reverse (Adjoint::face_reverse) and the
// minmod-reconstruction reverse, and gathers the transposed stencil onto the
// cells.  The work decomposition mirrors the forward stage: one wave64 per
// strip of 64 - 2*HX columns (HX = order + 1 halo columns each side, since
// the transposed stencil reaches two faces further) and a segment of rows;
// x-faces are exchanged between lanes with DPP shifts, y-faces stream up the
// column with a four-row accumulator ring.  Ghost-cell adjoints are folded
// back onto their mirror cells through the transposed boundary condition.
//
// Driver (AdjointDriver): the forward stores a checkpoint every K steps;
// the backward walks the segments last to first, recomputes each segment's
// stage inputs and pre-implicit states from its checkpoint (the last
// segment stays stored from the forward) and runs the two reverse stages
// per step.  The dt sequence is the caller's (frozen), so the CFL max is not
// differentiated.  Losses enter through per-step gauge adjoints.

#include "zamr_driver.hpp"

namespace zamr {

struct AdjArgs {
    Params p;
    const double* U;    // stage input (box layout)
    const double* V;    // pre-implicit state U + dt RHS(U)
    const double* Wb;   // adjoint of the stage output
    double wscale;      // the stage output adjoint is wscale * Wb
    double* out;        // adjoint of the stage input
    const double* heun; // optional: out += heun / 2 (read at the same cell first)
    double* nbar;       // adjoint of the friction field (one row, box layout), accumulated
    double dt, time;
    int clamp;  // the forward's positivity clamp of the stage input (UpdateState)
    int wb;     // well-balanced (hydrostatic) reconstruction
};

template <class M, class AD>
struct AdjSys {
    using S = Sys<M>;
    using VQ = typename S::VQ;
    using VP = typename S::VP;
    static constexpr int nq = S::nq;

    // Reverse of one face of the forward divergence.  The face sits between
    // L and R (Lm, Rp the outer stencil cells); vtL, vtR are dt * Vb of L and
    // R; rd = 1/dx along DIR.  RHS_L holds (-F - D-) rd and RHS_R (F - D+) rd;
    // with well-balancing (wb) the face states are hydrostatically
    // reconstructed first and RHS_L / RHS_R of the normal momentum also hold
    // +g/2 h*_L^2 rd / -g/2 h*_R^2 rd (the balancing bed source).  Returns the
    // adjoint contributions to the four stencil cells.
    template <int ORDER, int DIR>
    __device__ static __forceinline__ void face(const VQ& Lm, const VQ& L, const VQ& R, const VQ& Rp, const VQ& vtL,
                                                const VQ& vtR, double rd, const VP& p, int wb, VQ& cLm, VQ& cL, VQ& cR,
                                                VQ& cRp) {
        VQ qL, qR;
        unsigned sl[nq], sr[nq];
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            if constexpr (ORDER == 2) {
                const double a1 = L(n, 0) - Lm(n, 0), b1 = R(n, 0) - L(n, 0);
                const double b2 = Rp(n, 0) - R(n, 0);
                sl[n] = minmod_code(a1, b1);
                sr[n] = minmod_code(b1, b2);
                const double gL = (sl[n] == 1u) ? a1 : ((sl[n] == 2u) ? b1 : 0.0);
                const double gR = (sr[n] == 1u) ? b1 : ((sr[n] == 2u) ? b2 : 0.0);
                qL(n, 0) = L(n, 0) + 0.5 * gL;
                qR(n, 0) = R(n, 0) - 0.5 * gR;
            } else {
                sl[n] = sr[n] = 0u;
                qL(n, 0) = L(n, 0);
                qR(n, 0) = R(n, 0);
            }
        }
        VQ Fb, Dpb, Dmb, qLb, qRb;
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            Fb(n, 0) = (vtR(n, 0) - vtL(n, 0)) * rd;
            Dmb(n, 0) = -vtL(n, 0) * rd;
            Dpb(n, 0) = -vtR(n, 0) * rd;
        }
        if (wb) {
            const double kg = p(0, 0) * rd;
            AD::template wb_face_reverse<DIR>(qL, qR, p, Fb, Dpb, Dmb, kg * vtL(2 + DIR, 0), -kg * vtR(2 + DIR, 0),
                                              qLb, qRb);
        } else {
            AD::template face_reverse<DIR>(qL, qR, p, Fb, Dpb, Dmb, qLb, qRb);
        }
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            double lm = 0.0, l = qLb(n, 0), r = qRb(n, 0), rp = 0.0;
            if constexpr (ORDER == 2) {
                const double gl = 0.5 * qLb(n, 0);   // qL = L + gL/2
                const double gr = -0.5 * qRb(n, 0);  // qR = R - gR/2
                if (sl[n] == 1u) { l += gl; lm -= gl; }
                else if (sl[n] == 2u) { r += gl; l -= gl; }
                if (sr[n] == 1u) { r += gr; l -= gr; }
                else if (sr[n] == 2u) { rp += gr; r -= gr; }
            }
            cLm(n, 0) = lm;
            cL(n, 0) = l;
            cR(n, 0) = r;
            cRp(n, 0) = rp;
        }
    }
};

template <class V>
__device__ __forceinline__ void vzero(V& v) {
#pragma unroll
    for (int n = 0; n < V::row_size * V::column_size; ++n) v(n, 0) = 0.0;
}
template <class V>
__device__ __forceinline__ void vadd(V& a, const V& b) {
#pragma unroll
    for (int n = 0; n < V::row_size * V::column_size; ++n) a(n, 0) += b(n, 0);
}

template <class M, class AD, int ORDER>
// Occupancy request for the reverse stage kernel (waves per SIMD); 0 leaves
// it to the compiler.  With the row prefetch the kernel holds ~210 VGPRs, two
// waves per SIMD; forcing three spills 160 bytes per lane and is slower
// (1024^2 backward: 47.5 ms at two waves, 72 ms at three; without the
// prefetch three waves were best, 49.6 ms).
#ifndef ZAMR_ADJ_WAVES
#define ZAMR_ADJ_WAVES 0
#endif
#if ZAMR_ADJ_WAVES > 0
#define ZAMR_ADJ_ATTR __attribute__((amdgpu_waves_per_eu(ZAMR_ADJ_WAVES, ZAMR_ADJ_WAVES)))
#else
#define ZAMR_ADJ_ATTR
#endif
__global__ void __launch_bounds__(NT) ZAMR_ADJ_ATTR adjoint_stage_kernel(const BoxDesc* __restrict__ boxes,
                                                           const WorkItem* __restrict__ items, int nitems, AdjArgs a) {
    using S = Sys<M>;
    using A = AdjSys<M, AD>;
    using VQ = typename S::VQ;
    using VP = typename S::VP;
    constexpr int nq = S::nq;
    constexpr int HX = ORDER + 1;
    constexpr int TXW = WV - 2 * HX;
    static_assert(S::dim == 2, "the adjoint kernel is written for two-dimensional boxes");

    VP p;
#pragma unroll
    for (int k = 0; k < S::np; ++k) p(k, 0) = a.p.v[k];

    const int t = threadIdx.x & (WV - 1);
    const int it = blockIdx.x * (NT / WV) + wave_in_block();
    if (it >= nitems) return;
    const WorkItem w = items[it];
    const BoxView B = view(boxes, w.box);
    const int c = w.i0 - HX + t;
    const bool own = (t >= HX) && (t < HX + TXW) && (c <= B.hi0);
    const bool vx = (c >= B.dlo0) && (c <= B.dhi0);
    const bool vxl = (c - 1 >= B.dlo0) && (c - 1 <= B.dhi0);
    // lanes whose face c-1/2 has all four stencil lanes inside the wave
    const bool face_lane = (t >= 2) && (t <= WV - 2);
    const bool edge_lo = w.i0 - HX < B.dlo0;
    const bool edge_hi = w.i0 - HX + WV - 1 > B.dhi0;
    const double rdx = 1.0 / B.dx0, rdy = 1.0 / B.dx1;
    const int j0 = w.j0, j1 = w.j0 + w.nj;

    // The loads of the pointwise reverse at (c, r), fetched one row ahead of
    // their use: Wb, V, the friction coefficient and the friction adjoint so
    // far.  Lanes or rows outside the domain read an in-domain cell (their
    // values are not used).
    struct PointIn {
        VQ wb, v;
        double nn, nb;
    };
    const int cp = min(max(c, B.dlo0), B.dhi0);
    auto point_fetch = [&](int r, PointIn& q) {
        const int rr = min(max(r, B.dlo1), B.dhi1);
        const long long idx = (long long)(cp - B.gl0) + (long long)(rr - B.gl1) * B.sy;
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            q.wb(n, 0) = a.Wb[idx + n * B.sc];
            q.v(n, 0) = a.V[idx + n * B.sc];
        }
        q.nn = B.aux[idx + AD::friction_row * B.sc];
        q.nb = a.nbar[idx];
    };
    // the prefetched inputs are in (before a row's stores; see land())
    auto land_point = [&](PointIn& q) {
        land(q.wb);
        land(q.v);
        asm volatile("" : "+v"(q.nn), "+v"(q.nb));
    };

    // pointwise reverse of the implicit source at (c, r): dt * Vb and the
    // direct part Vb + [d aux/dU]^T ab; accumulates the friction adjoint
    // (each cell's friction adjoint is updated by one lane once per launch,
    // so the value fetched a row earlier is current)
    auto point = [&](int r, const VQ& u, const PointIn& q, VQ& vt, VQ& d) {
        vzero(vt);
        vzero(d);
        const bool vy = (r >= B.dlo1) && (r <= B.dhi1);
        if (vx && vy) {
            const long long idx = (long long)(c - B.gl0) + (long long)(r - B.gl1) * B.sy;
            VQ wb, vb, ub;
#pragma unroll
            for (int n = 0; n < nq; ++n) wb(n, 0) = a.wscale * q.wb(n, 0);
            double nb;
            AD::source_reverse(u, q.v, q.nn, a.dt, p, wb, vb, ub, nb);
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                vt(n, 0) = a.dt * vb(n, 0);
                d(n, 0) = vb(n, 0) + ub(n, 0);
            }
            if (own && r >= j0 && r < j1) a.nbar[idx] = q.nb + nb;
        }
    };

    // x-faces of one row: every lane evaluates its left face c-1/2; the four
    // stencil contributions are gathered onto their lanes with DPP shifts and
    // ghost lanes are folded onto their mirror cells.
    auto xrow = [&](const VQ& u, const VQ& vt) -> VQ {
        VQ Lm, L, Rp, vl, cLm, cL, cR, cRp;
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            L(n, 0) = from_left(u(n, 0));
            Lm(n, 0) = from_left(L(n, 0));
            Rp(n, 0) = from_right(u(n, 0));
            vl(n, 0) = from_left(vt(n, 0));
        }
        vzero(cLm);
        vzero(cL);
        vzero(cR);
        vzero(cRp);
        if (face_lane && (vx || vxl)) A::template face<ORDER, 0>(Lm, L, u, Rp, vl, vt, rdx, p, a.wb, cLm, cL, cR, cRp);
        VQ G;
#pragma unroll
        for (int n = 0; n < nq; ++n)
            G(n, 0) = cR(n, 0) + from_right(cL(n, 0)) + from_right(from_right(cLm(n, 0))) + from_left(cRp(n, 0));
        if (edge_lo || edge_hi) {
            VQ A1, A3, B1, B3;
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                A1(n, 0) = from_left(G(n, 0));
                A3(n, 0) = from_left(from_left(A1(n, 0)));
                B1(n, 0) = from_right(G(n, 0));
                B3(n, 0) = from_right(from_right(B1(n, 0)));
            }
            // ghost dlo-1-k mirrors dlo+k, ghost dhi+1+k mirrors dhi-k (k = 0, 1)
            if (c == B.dlo0) vadd(G, AD::bc_reverse(B.bc0, 0, A1));
            if (c == B.dlo0 + 1) vadd(G, AD::bc_reverse(B.bc0, 0, A3));
            if (c == B.dhi0) vadd(G, AD::bc_reverse(B.bc1, 0, B1));
            if (c == B.dhi0 - 1) vadd(G, AD::bc_reverse(B.bc1, 0, B3));
        }
        return G;
    };

    // Columns past hi + HX feed no live face: those lanes load column
    // hi + HX instead (in bounds, in the edge lane's cache lines).  The box is
    // the whole domain, so every other column and row mirrors into it.
    const int cl = min(c, B.hi0 + HX);
    auto load = [&](int r) -> VQ { return S::load(B, a.U, cl, r, p, a.clamp, a.time); };

    VQ U0 = load(j0 - 3), U1 = load(j0 - 2), U2 = load(j0 - 1), U3;
    VQ P0, P1, R0, R1, R2, R3;
    PointIn pin;
    {
        VQ d;
        point_fetch(j0 - 2, pin);
        point(j0 - 2, U1, pin, P0, d);
    }
    vzero(R0);
    vzero(R1);
    vzero(R2);
    vzero(R3);
    // the first row's prefetches: U of row j0 and the point inputs of row j0-1
    typename S::Raw unext = S::fetch(B, a.U, cl, j0);
    point_fetch(j0 - 1, pin);
    land(U0);
    land(U1);
    land(U2);
    land(unext.v);
    land_point(pin);
    // iteration j: face j+1/2; rows j-1 .. j+2 in U0..U3 and in the
    // accumulators R0..R3; dt*Vb of rows j, j+1 in P0, P1
    for (int j = j0 - 2; j <= j1; ++j) {
        U3 = S::finish(B, unext, cl, j + 2, p, a.clamp, a.time);
        const PointIn pcur = pin;
        // the next row's loads (the last row's U prefetch re-reads row hi + HX)
        unext = S::fetch(B, a.U, cl, min(j + 3, B.hi1 + HX));
        point_fetch(j + 2, pin);
        // the Heun part of this row's output, used at the row's end
        const bool store_row = j - 1 >= j0 && own && vx;
        VQ hn;
        vzero(hn);
        if (a.heun && store_row) {
            const long long idx = (long long)(c - B.gl0) + (long long)(j - 1 - B.gl1) * B.sy;
#pragma unroll
            for (int n = 0; n < nq; ++n) hn(n, 0) = a.heun[idx + n * B.sc];
        }
        {
            VQ d;
            point(j + 1, U2, pcur, P1, d);
            vadd(R2, d);
        }
        if (j + 1 >= j0 && j + 1 < j1) vadd(R2, xrow(U2, P1));
        VQ cLm, cL, cR, cRp;
        vzero(cLm);
        vzero(cL);
        vzero(cR);
        vzero(cRp);
        const bool vy0 = (j >= B.dlo1) && (j <= B.dhi1);
        const bool vy1 = (j + 1 >= B.dlo1) && (j + 1 <= B.dhi1);
        if (own && vx && (vy0 || vy1)) A::template face<ORDER, 1>(U0, U1, U2, U3, P0, P1, rdy, p, a.wb, cLm, cL, cR, cRp);
        if (j - 1 >= B.dlo1 && j + 2 <= B.dhi1) {
            vadd(R0, cLm);
            vadd(R1, cL);
            vadd(R2, cR);
            vadd(R3, cRp);
        } else {
            // rows outside the domain: onto the mirror row through the
            // transposed boundary condition (all targets stay in the ring)
            auto redirect = [&](int r, VQ& v) -> int {
                int m = r;
                if (r < B.dlo1) {
                    m = 2 * B.dlo1 - 1 - r;
                    v = AD::bc_reverse(B.bc2, 1, v);
                } else if (r > B.dhi1) {
                    m = 2 * B.dhi1 + 1 - r;
                    v = AD::bc_reverse(B.bc3, 1, v);
                }
                return m - (j - 1);
            };
            const int sLm = redirect(j - 1, cLm), sL = redirect(j, cL);
            const int sR = redirect(j + 1, cR), sRp = redirect(j + 2, cRp);
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                R0(n, 0) += (sLm == 0 ? cLm(n, 0) : 0.0) + (sL == 0 ? cL(n, 0) : 0.0) + (sR == 0 ? cR(n, 0) : 0.0) +
                            (sRp == 0 ? cRp(n, 0) : 0.0);
                R1(n, 0) += (sLm == 1 ? cLm(n, 0) : 0.0) + (sL == 1 ? cL(n, 0) : 0.0) + (sR == 1 ? cR(n, 0) : 0.0) +
                            (sRp == 1 ? cRp(n, 0) : 0.0);
                R2(n, 0) += (sLm == 2 ? cLm(n, 0) : 0.0) + (sL == 2 ? cL(n, 0) : 0.0) + (sR == 2 ? cR(n, 0) : 0.0) +
                            (sRp == 2 ? cRp(n, 0) : 0.0);
                R3(n, 0) += (sLm == 3 ? cLm(n, 0) : 0.0) + (sL == 3 ? cL(n, 0) : 0.0) + (sR == 3 ? cR(n, 0) : 0.0) +
                            (sRp == 3 ? cRp(n, 0) : 0.0);
            }
        }
        land(unext.v);
        land_point(pin);
        if (store_row) {
            const long long idx = (long long)(c - B.gl0) + (long long)(j - 1 - B.gl1) * B.sy;
            // the stage read its input through UpdateState's clamp: no
            // adjoint reaches a depth the clamp replaced by 0
            const bool cut = a.clamp && a.U[idx + B.sc] < 0.0;
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                double v = (n == 1 && cut) ? 0.0 : R0(n, 0);
                if (a.heun) v += 0.5 * hn(n, 0);
                a.out[idx + n * B.sc] = v;
            }
        }
        U0 = U1;
        U1 = U2;
        U2 = U3;
        P0 = P1;
        R0 = R1;
        R1 = R2;
        R2 = R3;
        vzero(R3);
    }
}

// h at the gauge cells (component 1) into out[0 .. ng).
__global__ void gauge_sample_kernel(const double* __restrict__ q, long long sc, const long long* __restrict__ gidx,
                                    int ng, double* __restrict__ out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) out[g] = q[gidx[g] + sc];
}

// Loss adjoint of the gauge depths added into the state adjoint (distinct cells).
__global__ void gauge_inject_kernel(double* __restrict__ qb, long long sc, const long long* __restrict__ gidx, int ng,
                                    const double* __restrict__ gbar) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) qb[gidx[g] + sc] += gbar[g];
}

// Reverse of UpdateState's positivity clamp on a step's output, X_{k+1} =
// clamp(Z), Z = X_k / 2 + W / 2 (W's depth is the pre-implicit V's: the
// friction step leaves h alone): the depth adjoint of the cells the clamp
// replaced (stored X_{k+1} depth 0 from a negative Z) is dropped, in place.
__global__ void clamp_reverse_kernel(double* __restrict__ qb, const double* __restrict__ xn,
                                     const double* __restrict__ xo, const double* __restrict__ v, long long sc,
                                     long long cells) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    const long long i = k + sc;  // depth row
    if (xn[i] == 0.0 && 0.5 * xo[i] + 0.5 * v[i] < 0.0) qb[i] = 0.0;
}

// Valid cells of every box of a level (grid.y = box) to or from one box
// covering the domain (grown lower corner (gl0, gl1), row stride sy, component
// stride sc): the state buffer buf of the boxes (buf < 0: their stored Qaux),
// nc components.
__global__ void box_exchange_kernel(const BoxDesc* __restrict__ boxes, int buf, double* __restrict__ one, int gl0,
                                    int gl1, int sy, long long sc, int nc, int to_one) {
    const BoxView B = view(boxes, blockIdx.y);
    const int w = B.hi0 - B.lo0 + 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= w * (B.hi1 - B.lo1 + 1)) return;
    const int i = B.lo0 + k % w, j = B.lo1 + k / w;
    double* q = buf < 0 ? B.aux : qsel(B, buf);
    const long long ib = (long long)(i - B.gl0) + (long long)(j - B.gl1) * B.sy;
    const long long io = (long long)(i - gl0) + (long long)(j - gl1) * sy;
    for (int n = 0; n < nc; ++n) {
        if (to_one) one[io + n * sc] = q[ib + n * B.sc];
        else q[ib + n * B.sc] = one[io + n * sc];
    }
}

// Valid cells of a grown-box array (ncomp components, component stride sc,
// row stride gx, ngrow ghost cells a side) into a dense (ncomp, ny, nx) array.
__global__ void valid_gather_kernel(const double* __restrict__ src, long long sc, int gx, int ngrow, int nx, int ny,
                                    int ncomp, double* __restrict__ dst) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)nx * ny;
    if (k >= per * ncomp) return;
    const int c = (int)(k / per);
    const long long r = k - c * per;
    const int j = (int)(r / nx), i = (int)(r - (long long)j * nx);
    dst[k] = src[(long long)(i + ngrow) + (long long)(j + ngrow) * gx + c * sc];
}

template <class M, class N, class AD>
struct AdjointDriver {
    using Sol = Solver<M, N>;
    static constexpr int nq = M::n_dof_q;
    static constexpr int ORDER = 2;
    static constexpr int HX = ORDER + 1;

    // Per-step storage of one segment: stage inputs X[i] (X[len] = the end),
    // Y[i], pre-implicit states V1[i], V2[i], and a descriptor per step
    // (q0 = X[i], q1 = Y[i]).  Two slots: the backward recomputes segment
    // m-1 into one (second stream) while it reverses segment m from the other.
    struct Slot {
        std::vector<double*> X, Y, V1, V2;
        BoxDesc* d_desc = nullptr;
        int held = -1;  // segment whose states the slot holds
    };

    Sol& s;
    int nsteps = 0, seg = 0, nseg = 0, ng = 0;
    std::vector<double> dts, times;
    std::vector<long long> gidx;
    size_t cells = 0, bufn = 0;  // grown cells per box, doubles per state buffer
    double* pool = nullptr;
    GroupState* d_groups = nullptr;
    long long* d_gidx = nullptr;
    double *d_gh = nullptr, *d_gbar = nullptr;
    WorkItem* d_items = nullptr;
    int nitems = 0;
    std::vector<double*> ck;
    Slot slot[2];
    double *Ab = nullptr, *Bb = nullptr, *nbar = nullptr;
    hipStream_t side = nullptr;  // recompute stream (default priority)
    hipEvent_t ev_go = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    bool have_forward = false;
    bool have_backward = false;
    // device I/O: the gauge, gauge-adjoint and gradient arrays the calls are
    // given are device pointers, and the calls do not wait for the stream
    bool device_io = false;
    // frozen limiter: the first forward after freeze_limiter(1) records every
    // face's minmod choices, later forwards replay them
    bool freeze = false, sw_recorded = false, replayed = false;
    int fwd_sw = 0;  // stage-kernel build of the last forward (the recompute uses the same)
    unsigned* d_sw = nullptr;

    // In device I/O mode every array the caller hands over must be device
    // (or managed) memory; a host address there would be copied as if it
    // were device memory.
    static void require_device(const void* p, const char* what) {
        if (!p) return;
        hipPointerAttribute_t at{};
        const hipError_t e = hipPointerGetAttributes(&at, p);
        if (e != hipSuccess) (void)hipGetLastError();
        if (e != hipSuccess || (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged))
            throw Error(ZAMR_EINVAL, std::string("adjoint: device I/O mode, but ") + what + " is not device memory");
    }

    // The working layout: the context's solver when its one box is the
    // domain, else a single-box solver over the same domain with the same
    // settings (`one`), into which the forward gathers the boxes' state and
    // from which it scatters the end state back.  The forward's values do not
    // depend on the box decomposition, so the gradient is that of the
    // context's own multi-box forward.
    std::unique_ptr<Sol> one;
    Sol& w() { return one ? *one : s; }
    const Sol& w() const { return one ? *one : s; }
    // the working solver runs on the context's current stream
    void bind() {
        if (one) one->stream = s.stream;
    }

    explicit AdjointDriver(Sol& sol) : s(sol) {}
    ~AdjointDriver() { release(); }
    void release() {
        bind();
        if (side) hipStreamSynchronize(side);
        if (w().stream) hipStreamSynchronize(w().stream);
        if (pool) hipFree(pool);
        if (d_groups) hipFree(d_groups);
        if (d_gidx) hipFree(d_gidx);
        for (auto& sl : slot) {
            if (sl.d_desc) hipFree(sl.d_desc);
            sl = Slot{};
        }
        if (d_items) hipFree(d_items);
        if (d_sw) hipFree(d_sw);
        for (hipEvent_t* e : {&ev_go, &ev_rec[0], &ev_rec[1], &ev_rev[0], &ev_rev[1]})
            if (*e) hipEventDestroy(*e), *e = nullptr;
        if (side) hipStreamDestroy(side);
        side = nullptr;
        d_sw = nullptr;
        freeze = sw_recorded = replayed = false;
        pool = nullptr;
        d_groups = nullptr;
        d_gidx = nullptr;
        d_items = nullptr;
        have_forward = have_backward = false;
    }

    void check_config() const {
        if (s.batched || s.hb.empty())
            throw Error(ZAMR_EUNSUP, "the adjoint runs on a single-level domain (not on box lists)");
        if (s.spatial_order != ORDER) throw Error(ZAMR_EUNSUP, "the adjoint is built for solver.spatial_order = 2");
        if (!s.implicit_source) throw Error(ZAMR_EUNSUP, "the adjoint is built for solver.implicit_source = true");
        if (s.implicit_global)
            throw Error(ZAMR_EUNSUP, "the adjoint reverses the point-implicit source (solver.implicit_global = false)");
        if (s.positivity != "none")
            throw Error(ZAMR_EUNSUP, "the adjoint covers solver.positivity = none (with or without the clamp)");
        const HostBox& b = s.hb[0];
        if (b.dhi[0] - b.dlo[0] < 7 || b.dhi[1] - b.dlo[1] < 7)
            throw Error(ZAMR_EINVAL, "the adjoint needs >= 8 cells a side");
    }

    // Valid cells of the context's boxes (state buffer k, or the stored Qaux)
    // to or from the working box.
    void exchange(bool to_one, bool aux) {
        if (!one) return;
        const HostBox& o = one->hb[0];
        int maxc = 0;
        for (const HostBox& b : s.hb) maxc = std::max(maxc, (b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1));
        double* dst = aux ? one->d_pool + o.off_aux : one->d_pool + o.off_q[one->cur];
        const int nc = aux ? (int)std::max(M::n_dof_qaux, 1) : nq;
        hipLaunchKernelGGL(box_exchange_kernel, dim3((unsigned)((maxc + 255) / 256), (unsigned)s.hb.size()), dim3(256),
                           0, w().stream, s.d_boxes, aux ? -1 : s.cur, dst, o.lo[0] - o.ng, o.lo[1] - o.ng, o.gx,
                           (long long)o.gx * o.gy, nc, to_one ? 1 : 0);
        ZAMR_HIP(hipGetLastError());
    }

    void setup(int nsteps_, int seg_, const double* dt, int ng_, const int* gi, const int* gj) {
        check_config();
        if (nsteps_ < 1 || seg_ < 1 || !dt) throw Error(ZAMR_EINVAL, "adjoint: nsteps and segment length must be >= 1");
        if (ng_ < 0 || (ng_ > 0 && (!gi || !gj))) throw Error(ZAMR_EINVAL, "adjoint: bad gauge arrays");
        release();
        if (s.hb.size() > 1 && !one) {
            const HostBox& d = s.hb[0];
            one = std::make_unique<Sol>();
            one->configure_fine(s, 1, {{d.dlo[0], d.dlo[1], d.dhi[0], d.dhi[1]}});
        }
        bind();
        const HostBox& b = w().hb[0];
        nsteps = nsteps_;
        seg = std::min(seg_, nsteps);
        nseg = (nsteps + seg - 1) / seg;
        ng = ng_;
        dts.assign(dt, dt + nsteps);
        times.assign(nsteps, 0.0);
        double tt = 0.0;
        for (int k = 0; k < nsteps; ++k) {
            if (!(dts[k] > 0.0) || !std::isfinite(dts[k])) throw Error(ZAMR_EINVAL, "adjoint: dt must be > 0");
            times[k] = tt;
            tt += dts[k];
        }
        gidx.resize(ng);
        std::vector<long long> seen;
        for (int g = 0; g < ng; ++g) {
            if (gi[g] < b.lo[0] || gi[g] > b.hi[0] || gj[g] < b.lo[1] || gj[g] > b.hi[1])
                throw Error(ZAMR_EINVAL, "adjoint: gauge outside the domain");
            gidx[g] = (long long)(gi[g] - b.lo[0] + b.ng) + (long long)(gj[g] - b.lo[1] + b.ng) * b.gx;
            seen.push_back(gidx[g]);
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            throw Error(ZAMR_EINVAL, "adjoint: gauge cells must be distinct");

        cells = (size_t)b.gx * b.gy;
        auto al = [](size_t x) { return (x + 31) & ~size_t(31); };
        bufn = al(cells * nq);
        const size_t per_slot = (size_t)(seg + 1) + 3 * (size_t)seg;
        const size_t nbuf = (size_t)nseg + 2 * per_slot + 2;
        const size_t total = nbuf * bufn + al(cells) + 2 * al((size_t)nsteps * std::max(ng, 1));
        ZAMR_HIP(hipMalloc(&pool, total * sizeof(double)));
        double* p = pool;
        auto take = [&](size_t n) {
            double* r = p;
            p += n;
            return r;
        };
        ck.clear();
        for (int m = 0; m < nseg; ++m) ck.push_back(take(bufn));
        for (auto& sl : slot) {
            for (int i = 0; i <= seg; ++i) sl.X.push_back(take(bufn));
            for (int i = 0; i < seg; ++i) sl.Y.push_back(take(bufn));
            for (int i = 0; i < seg; ++i) sl.V1.push_back(take(bufn));
            for (int i = 0; i < seg; ++i) sl.V2.push_back(take(bufn));
        }
        Ab = take(bufn);
        Bb = take(bufn);
        nbar = take(al(cells));
        d_gh = take(al((size_t)nsteps * std::max(ng, 1)));
        d_gbar = take(al((size_t)nsteps * std::max(ng, 1)));

        std::vector<GroupState> gs(nsteps);
        for (int k = 0; k < nsteps; ++k) {
            GroupState g{};
            g.dt = dts[k];
            g.time = times[k];
            g.time_end = 1e300;
            g.cfl = w().cfl;
            g.dtmin = w().dtmin;
            g.dtmax = w().dtmax;
            g.divisor = 1.0;
            g.maxinv_bits = 0ull;
            g.step = k;
            g.max_step = -1;
            g.active = 1;
            gs[k] = g;
        }
        ZAMR_HIP(hipMalloc(&d_groups, gs.size() * sizeof(GroupState)));
        ZAMR_HIP(hipMemcpyAsync(d_groups, gs.data(), gs.size() * sizeof(GroupState), hipMemcpyHostToDevice, w().stream));
        if (ng > 0) {
            ZAMR_HIP(hipMalloc(&d_gidx, ng * sizeof(long long)));
            ZAMR_HIP(hipMemcpyAsync(d_gidx, gidx.data(), ng * sizeof(long long), hipMemcpyHostToDevice, w().stream));
        }
        for (auto& sl : slot) {
            std::vector<BoxDesc> dd(seg);
            for (int i = 0; i < seg; ++i) {
                dd[i] = w().desc_of(b, sl.X[i], sl.Y[i]);
                dd[i].group = 0;
            }
            ZAMR_HIP(hipMalloc(&sl.d_desc, dd.size() * sizeof(BoxDesc)));
            ZAMR_HIP(hipMemcpyAsync(sl.d_desc, dd.data(), dd.size() * sizeof(BoxDesc), hipMemcpyHostToDevice,
                                    w().stream));
            ZAMR_HIP(hipStreamSynchronize(w().stream));  // dd leaves scope
        }
        build_items();
        ZAMR_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        for (hipEvent_t* e : {&ev_go, &ev_rec[0], &ev_rec[1], &ev_rev[0], &ev_rev[1]})
            ZAMR_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
        w().items(HX - 1);  // the forward stage's work items, built before any second-stream use
        ZAMR_HIP(hipStreamSynchronize(w().stream));
    }

    void build_items() {
        const HostBox& b = w().hb[0];
        const int rows = w().rows_for(2);
        const int txc = WV - 2 * HX;
        std::vector<WorkItem> v;
        for (int j0 = b.lo[1]; j0 <= b.hi[1]; j0 += rows)
            for (int i0 = b.lo[0]; i0 <= b.hi[0]; i0 += txc) v.push_back(WorkItem{0, i0, j0, std::min(rows, b.hi[1] - j0 + 1)});
        nitems = (int)v.size();
        ZAMR_HIP(hipMalloc(&d_items, v.size() * sizeof(WorkItem)));
        ZAMR_HIP(hipMemcpyAsync(d_items, v.data(), v.size() * sizeof(WorkItem), hipMemcpyHostToDevice, w().stream));
        ZAMR_HIP(hipStreamSynchronize(w().stream));
    }

    int seg_len(int m) const { return std::min(seg, nsteps - m * seg); }
    double* state_buf(int k) const { return w().d_pool + w().hb[0].off_q[k]; }

    // Stage of step `step` on stream st; sw = 0 plain, 1 record, 2 replay.
    template <int EPI>
    void stage(const BoxDesc* boxes, StageArgs a, int step, int sw, hipStream_t st) {
        auto& it = w().items(HX - 1);
        const dim3 grid(slots_for(it.n)), block(NT);
        if (sw == 0) {
            hipLaunchKernelGGL((stage_kernel<M, N, ORDER, EPI, 0>), grid, block, 0, st, boxes, it.d, it.n, a,
                               d_groups + step, w().xcd_remap);
        } else {
            a.sw = d_sw + ((size_t)2 * step + (EPI == 0 ? 0 : 1)) * cells;
            if (sw == 1)
                hipLaunchKernelGGL((stage_kernel<M, N, ORDER, EPI, 1>), grid, block, 0, st, boxes, it.d, it.n, a,
                                   d_groups + step, w().xcd_remap);
            else
                hipLaunchKernelGGL((stage_kernel<M, N, ORDER, EPI, 2>), grid, block, 0, st, boxes, it.d, it.n, a,
                                   d_groups + step, w().xcd_remap);
        }
        ZAMR_HIP(hipGetLastError());
    }

    void freeze_limiter(bool on) {
        if (!pool) throw Error(ZAMR_EINVAL, "adjoint: call zamr_adj_setup first");
        bind();
        if (on && !d_sw) ZAMR_HIP(hipMalloc(&d_sw, (size_t)2 * nsteps * cells * sizeof(unsigned)));
        if (!on && d_sw) {
            ZAMR_HIP(hipStreamSynchronize(w().stream));
            hipFree(d_sw);
            d_sw = nullptr;
        }
        freeze = on;
        sw_recorded = false;
    }

    void sample(const double* q, int step) {
        if (ng == 0) return;
        hipLaunchKernelGGL(gauge_sample_kernel, dim3((ng + 255) / 256), dim3(256), 0, w().stream, q, (long long)cells,
                           d_gidx, ng, d_gh + (size_t)step * ng);
        ZAMR_HIP(hipGetLastError());
    }

    void copy(double* dst, const double* src, hipStream_t st) {
        ZAMR_HIP(hipMemcpyAsync(dst, src, cells * nq * sizeof(double), hipMemcpyDeviceToDevice, st));
    }

    // Steps of segment m from checkpoint ck[m] into slot k, on stream st.
    void run_stored(int m, int k, bool with_gauges, int sw, hipStream_t st) {
        Slot& sl = slot[k];
        copy(sl.X[0], ck[m], st);
        const int len = seg_len(m);
        for (int i = 0; i < len; ++i) {
            const int step = m * seg + i;
            StageArgs a1 = w().stage_args(0, 1, 0);
            a1.pre = sl.V1[i];
            stage<0>(sl.d_desc + i, a1, step, sw, st);
            StageArgs a2 = w().stage_args(1, 0, 0);
            a2.pre = sl.V2[i];
            a2.dstp = sl.X[i + 1];
            stage<1>(sl.d_desc + i, a2, step, sw, st);
            if (with_gauges) sample(sl.X[i + 1], step);
        }
        sl.held = m;
    }

    // Forward from the solver's current state with the frozen dt sequence;
    // gauge depths after every step into gh (nsteps * ng, may be null).
    void forward(double* gh) {
        if (!pool) throw Error(ZAMR_EINVAL, "adjoint: call zamr_adj_setup first");
        check_config();
        if (device_io) require_device(gh, "gauge_h");
        bind();
        // (the stage kernels evaluate Qaux per cell from the state and the
        // stored input rows, so a stale stored Qaux is not read here)
        exchange(true, false);
        exchange(true, true);
        const int sw = freeze ? (sw_recorded ? 2 : 1) : 0;
        slot[0].held = slot[1].held = -1;
        for (int m = 0; m < nseg; ++m) {
            copy(ck[m], state_buf(w().cur), w().stream);
            if (m + 1 < nseg) {
                for (int i = 0; i < seg_len(m); ++i) {
                    const int k = m * seg + i;
                    stage<0>(w().d_boxes, w().stage_args(w().cur, 1 - w().cur, w().cur), k, sw, w().stream);
                    stage<1>(w().d_boxes, w().stage_args(1 - w().cur, w().cur, w().cur), k, sw, w().stream);
                    sample(state_buf(w().cur), k);
                }
            } else {
                run_stored(m, 0, true, sw, w().stream);
                copy(state_buf(w().cur), slot[0].X[seg_len(m)], w().stream);
            }
        }
        exchange(false, false);
        if (sw == 1) sw_recorded = true;
        replayed = sw == 2;
        fwd_sw = sw;
        s.aux_stale = true;
        s.maxinv_fresh = false;
        if (gh && ng > 0)
            ZAMR_HIP(hipMemcpyAsync(gh, d_gh, (size_t)nsteps * ng * sizeof(double),
                                    device_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, w().stream));
        if (!device_io) ZAMR_HIP(hipStreamSynchronize(w().stream));
        have_forward = true;
        have_backward = false;
    }

    void reverse_stage(const Slot& sl, const double* U, const double* V, const double* Wb, double wscale, double* out,
                       const double* heun, int k) {
        AdjArgs a;
        a.p = w().prm;
        a.U = U;
        a.V = V;
        a.Wb = Wb;
        a.wscale = wscale;
        a.out = out;
        a.heun = heun;
        a.nbar = nbar;
        a.dt = dts[k];
        a.time = times[k];
        a.clamp = w().clamp_flag();
        a.wb = w().well_balanced ? 1 : 0;
        hipLaunchKernelGGL((adjoint_stage_kernel<M, AD, ORDER>), dim3((nitems + NT / WV - 1) / (NT / WV)), dim3(NT), 0,
                           w().stream, sl.d_desc, d_items, nitems, a);
        ZAMR_HIP(hipGetLastError());
    }

    // Reverse of segment m from slot k (main stream).
    void reverse_segment(int m, int k) {
        const Slot& sl = slot[k];
        for (int i = seg_len(m) - 1; i >= 0; --i) {
            const int step = m * seg + i;
            if (ng > 0) {
                hipLaunchKernelGGL(gauge_inject_kernel, dim3((ng + 255) / 256), dim3(256), 0, w().stream, Ab,
                                   (long long)cells, d_gidx, ng, d_gbar + (size_t)step * ng);
                ZAMR_HIP(hipGetLastError());
            }
            if (w().clamp_flag()) {
                hipLaunchKernelGGL(clamp_reverse_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, w().stream,
                                   Ab, sl.X[i + 1], sl.X[i], sl.V2[i], (long long)cells, (long long)cells);
                ZAMR_HIP(hipGetLastError());
            }
            reverse_stage(sl, sl.Y[i], sl.V2[i], Ab, 0.5, Bb, nullptr, step);
            reverse_stage(sl, sl.X[i], sl.V1[i], Bb, 1.0, Ab, Ab, step);
        }
    }

    // Backward over all steps; gbar = adjoint of the gauge depths (nsteps * ng).
    // Segment m is reversed on the main stream while the second stream
    // recomputes segment m-1 from its checkpoint into the other slot; events
    // order slot reuse both ways.
    void backward(const double* gbar) {
        if (!have_forward) throw Error(ZAMR_EINVAL, "adjoint: run zamr_adj_forward first");
        bind();
        if (replayed)
            throw Error(ZAMR_EINVAL, "adjoint: the last forward replayed frozen limiter choices; the backward "
                                     "linearises a forward that computed them (freeze off, or the recording forward)");
        if (ng > 0) {
            if (!gbar) throw Error(ZAMR_EINVAL, "adjoint: gauge adjoint array is null");
            if (device_io) require_device(gbar, "gauge_hbar");
            ZAMR_HIP(hipMemcpyAsync(d_gbar, gbar, (size_t)nsteps * ng * sizeof(double),
                                    device_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, w().stream));
        }
        ZAMR_HIP(hipMemsetAsync(Ab, 0, bufn * sizeof(double), w().stream));
        ZAMR_HIP(hipMemsetAsync(Bb, 0, bufn * sizeof(double), w().stream));
        ZAMR_HIP(hipMemsetAsync(nbar, 0, cells * sizeof(double), w().stream));
        w().items(HX - 1);  // (re)built on the main stream, before the recompute stream reads it
        // the recompute stream starts after everything queued so far (the forward)
        ZAMR_HIP(hipEventRecord(ev_go, w().stream));
        ZAMR_HIP(hipStreamWaitEvent(side, ev_go, 0));
        int k = (slot[0].held == nseg - 1) ? 0 : ((slot[1].held == nseg - 1) ? 1 : -1);
        if (k < 0) {
            k = 0;
            run_stored(nseg - 1, 0, false, fwd_sw, w().stream);
        }
        ZAMR_HIP(hipEventRecord(ev_rev[1 - k], w().stream));  // the other slot is free
        for (int m = nseg - 1; m >= 0; --m) {
            if (m > 0) {
                // the recompute runs the build of the stage kernel the forward
                // ran, so that it reproduces the forward bit for bit (FMA
                // contraction may differ between builds)
                ZAMR_HIP(hipStreamWaitEvent(side, ev_rev[1 - k], 0));
                run_stored(m - 1, 1 - k, false, fwd_sw, side);
                ZAMR_HIP(hipEventRecord(ev_rec[1 - k], side));
            }
            reverse_segment(m, k);
            ZAMR_HIP(hipEventRecord(ev_rev[k], w().stream));
            if (m > 0) ZAMR_HIP(hipStreamWaitEvent(w().stream, ev_rec[1 - k], 0));
            k = 1 - k;
        }
        if (!device_io) ZAMR_HIP(hipStreamSynchronize(w().stream));
        have_backward = true;
    }

    // The adjoint of the initial state (nq * ny * nx) and of the friction
    // field (ny * nx), valid cells, raster order.
    void gradient(double* qbar, size_t nqb, double* fbar, size_t nfb) {
        if (!have_backward) throw Error(ZAMR_EINVAL, "adjoint: run zamr_adj_backward first");
        bind();
        const HostBox& b = w().hb[0];
        const int NX = b.hi[0] - b.lo[0] + 1, NY = b.hi[1] - b.lo[1] + 1;
        if ((qbar && nqb != (size_t)nq * NX * NY) || (fbar && nfb != (size_t)NX * NY))
            throw Error(ZAMR_EINVAL, "adjoint: gradient array size mismatch");
        if (device_io) {
            require_device(qbar, "qbar0");
            require_device(fbar, "auxbar");
            auto gather = [&](const double* src, int ncomp, double* dst) {
                const long long n = (long long)ncomp * NX * NY;
                hipLaunchKernelGGL(valid_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, w().stream, src,
                                   (long long)cells, b.gx, b.ng, NX, NY, ncomp, dst);
                ZAMR_HIP(hipGetLastError());
            };
            if (qbar) gather(Ab, nq, qbar);
            if (fbar) gather(nbar, 1, fbar);
            return;
        }
        std::vector<double> tmp(cells * nq), tn(cells);
        ZAMR_HIP(hipMemcpyAsync(tmp.data(), Ab, cells * nq * sizeof(double), hipMemcpyDeviceToHost, w().stream));
        ZAMR_HIP(hipMemcpyAsync(tn.data(), nbar, cells * sizeof(double), hipMemcpyDeviceToHost, w().stream));
        ZAMR_HIP(hipStreamSynchronize(w().stream));
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const size_t li = (size_t)(i + b.ng) + (size_t)(j + b.ng) * b.gx;
                if (qbar)
                    for (int n = 0; n < nq; ++n) qbar[((size_t)n * NY + j) * NX + i] = tmp[li + (size_t)n * cells];
                if (fbar) fbar[(size_t)j * NX + i] = tn[li];
            }
    }
};

}  // namespace zamr
