#pragma once
// gfx950 device kernels for the per-box RHS and time advance.
//
// One fused kernel per RK stage does what the reference spreads over seven
// per-box launches (Source/ZoomyAmr.cpp:561-675): UpdateState, physical-BC
// ghost fill, MUSCL/minmod reconstruction, Rusanov / path-conservative face
// fluxes, flux divergence + well-balanced bed source, ConsUpdate, the
// point-implicit source solve and (stage 2) the Heun combine, the end-of-step
// UpdateState and the CFL max-reduction for the next step.
//
// Work decomposition: a work item is a strip of 64-2*HX columns of one box and
// a segment of up to `rows` rows, and one wave64 owns one item: its 64 lanes
// hold 64 consecutive columns (the strip plus an HX-column halo on each side)
// and march up the segment row by row.  Every lane keeps a (2*HX+1)-row
// register window of its own column, so each y-face is evaluated once, by the
// lane that owns the column, and carried to the next row; each x-face is
// evaluated once, by the lane right of it, and its left-cell share is handed
// over with a cross-lane DPP shift.  No LDS and no barriers: the four waves
// of a workgroup are independent.  State is read row by row with consecutive
// lanes on consecutive cells (coalesced SoA); the next row's load is issued
// one row ahead.
//
// The Model / Numerics headers are the reference's generated device-function
// API (to_amrex.py:326-516), used unchanged.  Models that declare
// `zamr_aux_is_local` have an aux vector that is a per-cell function of the
// state, so aux is evaluated in registers rather than streamed from HBM.

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace zamr {

constexpr int kMaxParams = 32;
constexpr int NT = 256;  // threads per workgroup
constexpr int WV = 64;   // lanes per wave = columns per work item

// Global-memory views of the state pointers.  The box descriptors hold
// generic pointers; without this the compiler emits FLAT accesses, which
// count in both the vector-memory and the LDS/scalar counters, so a wait for
// a prefetched row also waits for every store issued after it.  GLOBAL
// accesses count in the in-order vector-memory counter only.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const gdouble* to_global(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* to_global(double* p) { return (gdouble*)p; }

// This wave's index in its workgroup.  The same on every lane; readfirstlane
// tells the compiler so, which keeps the item, its box descriptor and the row
// loop in scalar registers and scalar loads instead of one copy per lane.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WV); }

struct Params {
    double v[kMaxParams];
};

// One box (patch) on the device: AMReX Array4 layout, i fastest, then j, then
// component; base pointers address the grown-box corner (lo - ng).
struct BoxDesc {
    double* q[3];    // state buffers (the third only with a-posteriori MOOD)
    double* aux;     // stored Qaux (na comps, same layout)
    int lo[2], hi[2];
    int gl[2];       // grown lower corner
    int sy;          // row stride (elements)
    long long sc;    // component stride (elements)
    int dlo[2], dhi[2];  // physical domain of the box's level / member
    double dx[2], plo[2];
    int group;       // time-step group (level or ensemble member)
    int bc[4];       // tag index x_lo, x_hi, y_lo, y_hi
};

struct WorkItem {
    int box, i0, j0, nj;
};

// The descriptor as kernels hold it: scalars only, so no member is ever
// indexed at run time (which would demote the whole struct to scratch).
struct BoxView {
    double *q0, *q1, *q2, *aux;
    int lo0, lo1, hi0, hi1, gl0, gl1, sy;
    long long sc;
    int dlo0, dlo1, dhi0, dhi1;
    double dx0, dx1, plo0, plo1;
    int group, bc0, bc1, bc2, bc3;
};

__device__ __forceinline__ BoxView view(const BoxDesc* __restrict__ boxes, int b) {
    const BoxDesc& d = boxes[b];
    BoxView v;
    v.q0 = d.q[0];
    v.q1 = d.q[1];
    v.q2 = d.q[2];
    v.aux = d.aux;
    v.lo0 = d.lo[0];
    v.lo1 = d.lo[1];
    v.hi0 = d.hi[0];
    v.hi1 = d.hi[1];
    v.gl0 = d.gl[0];
    v.gl1 = d.gl[1];
    v.sy = d.sy;
    v.sc = d.sc;
    v.dlo0 = d.dlo[0];
    v.dlo1 = d.dlo[1];
    v.dhi0 = d.dhi[0];
    v.dhi1 = d.dhi[1];
    v.dx0 = d.dx[0];
    v.dx1 = d.dx[1];
    v.plo0 = d.plo[0];
    v.plo1 = d.plo[1];
    v.group = d.group;
    v.bc0 = d.bc[0];
    v.bc1 = d.bc[1];
    v.bc2 = d.bc[2];
    v.bc3 = d.bc[3];
    return v;
}

// Per-group stepping state, all on the device so that steps never round-trip
// dt through the host (ZoomyAmr.cpp:448-515 and :733-816 moved on-device).
struct GroupState {
    double dt, time, time_end, cfl, dtmin, dtmax, divisor;
    unsigned long long maxinv_bits;  // max over cells of |lambda_d| / dx_d (bit pattern, >= 0)
    int step, max_step, active;
    int troubled;                    // MOOD: troubled cells of this step's order-2 candidate
    long long redos;                 // MOOD: steps redone at order 1
};

struct StageArgs {
    Params p;
    int src, dst, old;           // buffer indices: read, write, Heun's Q^n
    int wb, impl, clamp;         // well_balanced, implicit_source (1 point-implicit solve here;
                                 // 2 solved globally after the stage: no source here), clamp
    int mood;                    // 0; 1 = count the troubled cells of the candidate
                                 // (stage 2); 2 = order-1 redo, troubled groups only
    double* pre;                 // optional: the pre-implicit state Q + dt RHS of a
                                 // single box (box layout), kept for the adjoint
    double* dstp;                // optional: output buffer of a single box in place
                                 // of buffer `dst` (the adjoint's per-step storage)
    unsigned* sw;                // limiter switch table of a single box (record /
                                 // replay builds of the stage kernel only)
};

// The bitwise build (-DZAMR_BITWISE, FMA contraction off) performs every
// division the reference performs, in the reference's order, and agrees with
// the CPU oracle bit for bit.  The throughput build multiplies by reciprocals
// where the reference divides by a cell size or a pivot (results within a few
// ulps; the parity bar is 1e-12 relative).
#ifdef ZAMR_BITWISE
constexpr bool kBitwise = true;
#else
constexpr bool kBitwise = false;
#endif

// 1/x: the hardware estimate refined by two Newton steps (normal range).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    return r;
}

template <class T>
__device__ __forceinline__ T dev_max(T a, T b) { return (a < b) ? b : a; }

// True only for a value the compiler knows to be the constant 0.0 (a literal
// zero entry of a model function's result, seen through inlining); false for
// every value computed at run time.
__device__ __forceinline__ bool zero_literal(double v) { return __builtin_constant_p(v) && v == 0.0; }

template <class V>
__device__ __forceinline__ void vq_zero(V& v) {
#pragma unroll
    for (int n = 0; n < V::row_size * V::column_size; ++n) v(n, 0) = 0.0;
}

// Redefine a loaded vector at this point, through an empty asm statement, so
// that the wait for its loads is placed here.  A loaded value that crosses a
// loop edge is copied there, and the copy waits for the load and so, the
// vector-memory counter being in order, for every store issued after it.
template <class V>
__device__ __forceinline__ void land(V& v) {
#pragma unroll
    for (int n = 0; n < V::row_size * V::column_size; ++n) asm volatile("" : "+v"(v(n, 0)));
}

// make_rhs.H:9-13
__device__ __forceinline__ double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}

// minmod for the throughput build: max(min(a, b), 0) + min(max(a, b), 0)
// picks the argument of smaller magnitude when both have one sign (either,
// when they are equal) and zero otherwise: minmod's value for finite
// arguments (up to the sign of a zero result), in five FP64 operations and
// no selects.
__device__ __forceinline__ double minmod_minmax(double a, double b) {
    return fmax(fmin(a, b), 0.0) + fmin(fmax(a, b), 0.0);
}

// Which argument minmod returns: 0 none (zero), 1 the first, 2 the second.
__device__ __forceinline__ unsigned minmod_code(double a, double b) {
    if (a * b <= 0.0) return 0u;
    return (fabs(a) < fabs(b)) ? 1u : 2u;
}

// Riemann solvers whose numerical_fluctuations are proportional to the bed
// jump qR(0) - qL(0) (N::zamr_fluct_from_bed_jump): after the hydrostatic
// reconstruction both sides carry the face bed, so a well-balanced face has
// none (exactly zero for finite states).
template <class N, class = void>
struct FluctFromBedJump : std::false_type {};
template <class N>
struct FluctFromBedJump<N, std::void_t<decltype(N::zamr_fluct_from_bed_jump)>>
    : std::bool_constant<N::zamr_fluct_from_bed_jump> {};

// Models whose update_aux_variables(Q, 0) is zero but for hinv of the depth
// (M::zamr_aux_only_hinv, with the hinv row and the wet-dry epsilon's index).
template <class M, class = void>
struct AuxOnlyHinv : std::false_type {};
template <class M>
struct AuxOnlyHinv<M, std::void_t<decltype(M::zamr_aux_only_hinv)>> : std::bool_constant<M::zamr_aux_only_hinv> {};

template <class M>
struct Sys {
    static constexpr int nq = M::n_dof_q;
    static constexpr int na = M::n_dof_qaux;
    static constexpr int np = M::n_parameters;
    static constexpr int dim = M::dimension;
    using VQ = amrex::SmallMatrix<double, nq, 1>;
    using VA = amrex::SmallMatrix<double, na, 1>;
    using VP = amrex::SmallMatrix<double, np, 1>;
    using VN = amrex::SmallMatrix<double, dim, 1>;
    using VX = amrex::SmallMatrix<double, 3, 1>;
    using VF = amrex::SmallMatrix<double, 2 * nq, 1>;
    static_assert(M::zamr_aux_is_local,
                  "this build evaluates Qaux in registers; the model must declare zamr_aux_is_local");
    static_assert(np <= kMaxParams, "too many model parameters");
    static_assert(!M::has_diffusion || M::zamr_n_aux_in == 0,
                  "a diffusive model averages both cells' aux at a face; stored aux input rows are not "
                  "read there, so such a model is not supported");
    static_assert(!M::has_diffusion || M::n_dof_gradQ == nq * dim, "gradQ holds n_dof_q x dimension entries");

    __device__ static __forceinline__ VX center(const BoxView& B, int i, int j) {
        VX X{};
        X(0, 0) = B.plo0 + (i + 0.5) * B.dx0;
        X(1, 0) = (dim == 2) ? B.plo1 + (j + 0.5) * B.dx1 : 0.0;
        X(2, 0) = 0.0;
        return X;
    }

    __device__ static __forceinline__ VA aux_of(const VQ& q, const VP& p, double time, const VX& X) {
        VA a0{};
        return M::update_aux_variables(q, a0, p, time, X);
    }

    // Aux of cell (c, j): the model's stored per-cell input rows
    // (M::zamr_aux_in, e.g. a Manning field) read from HBM, the live rows
    // recomputed from q.  Only the source and its Jacobians read input rows.
    __device__ static __forceinline__ VA aux_at(const BoxView& B, int c, int j, const VQ& q, const VP& p,
                                                double time, const VX& X) {
        VA a0{};
        if constexpr (M::zamr_n_aux_in > 0) {
            const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
#pragma unroll
            for (int k = 0; k < M::zamr_n_aux_in; ++k)
                a0(M::zamr_aux_in[k], 0) = B.aux[idx + M::zamr_aux_in[k] * B.sc];
        }
        return M::update_aux_variables(q, a0, p, time, X);
    }

    // UpdateState (ZoomyAmr.cpp:326-327) on one cell.
    __device__ static __forceinline__ VQ update_state(VQ q, const VP& p, int clamp) {
        VA a0{};
        q = M::update_variables(q, a0, p);
        if (clamp && q(1, 0) < 0.0) q(1, 0) = 0.0;
        return q;
    }

    // Mirror index of (c, j) about the physical domain (ZoomyAmr.cpp:399-406):
    // only the outside coordinate is reflected.
    __device__ static __forceinline__ void mirror(const BoxView& B, int c, int j, int& ci, int& cj) {
        ci = c;
        cj = j;
        if (c < B.dlo0) ci = 2 * B.dlo0 - 1 - c;
        else if (c > B.dhi0) ci = 2 * B.dhi0 + 1 - c;
        if (dim == 2) {
            if (j < B.dlo1) cj = 2 * B.dlo1 - 1 - j;
            else if (j > B.dhi1) cj = 2 * B.dhi1 + 1 - j;
        }
    }

    struct Raw {
        VQ v;
    };

    // The stored value behind cell (c, j): itself, or its mirror-index cell.
    template <class P>
    __device__ static __forceinline__ Raw fetch(const BoxView& B, P q, int c, int j) {
        int ci, cj;
        mirror(B, c, j, ci, cj);
        const long long idx = (long long)(ci - B.gl0) + (long long)(cj - B.gl1) * B.sy;
        Raw r;
#pragma unroll
        for (int n = 0; n < nq; ++n) r.v(n, 0) = q[idx + n * B.sc];
        return r;
    }

    // Value of cell (c, j) as the face pass sees it (do_stage: UpdateState,
    // then FillPatch, ZoomyAmr.cpp:563-572): the box's own cells after
    // UpdateState; ghost cells inside the domain as the stage's FillPatch
    // stored them (a neighbour box's updated value, or on a refined level the
    // coarse interpolant, which UpdateState never touches); outside the
    // physical domain the mirror-index cell's value, so obtained, pushed
    // through Model::boundary_conditions (ZoomyAmr.cpp:385-441; corner
    // ghosts take the x-side tag).
    __device__ static __forceinline__ VQ finish(const BoxView& B, const Raw& r, int c, int j, const VP& p, int clamp,
                                                double time) {
        const bool ox = (c < B.dlo0) || (c > B.dhi0);
        const bool oy = (dim == 2) && ((j < B.dlo1) || (j > B.dhi1));
        int ci, cj;
        mirror(B, c, j, ci, cj);
        const bool own = (ci >= B.lo0) && (ci <= B.hi0) && (cj >= B.lo1) && (cj <= B.hi1);
        VQ v = update_state(r.v, p, clamp);
        if (!own) v = r.v;
        if (!ox && !oy) return v;
        // Side index -> tag by arithmetic on all four tags: a select between
        // two field loads would be folded into a load through a selected
        // address, which pins the descriptor in scratch memory.
        const int side = ox ? ((c < B.dlo0) ? 0 : 1) : ((j < B.dlo1) ? 2 : 3);
        const int bc = (side == 0) * B.bc0 + (side == 1) * B.bc1 + (side == 2) * B.bc2 + (side == 3) * B.bc3;
        const double sgn = (side == 0 || side == 2) ? -1.0 : 1.0;
        VN nh{};
        if constexpr (dim == 2) {
            nh(0, 0) = ox ? sgn : 0.0;
            nh(1, 0) = ox ? 0.0 : sgn;
        } else {
            nh(0, 0) = sgn;
        }
        const VA ai = aux_of(v, p, time, center(B, ci, cj));
        return M::boundary_conditions(bc, v, ai, nh, center(B, c, j), time, B.dx0);
    }

    template <class P>
    __device__ static __forceinline__ VQ load(const BoxView& B, P q, int c, int j, const VP& p, int clamp,
                                              double time) {
        return finish(B, fetch(B, q, c, j), c, j, p, clamp, time);
    }

    // Limited slope of a cell from its two neighbours along one direction,
    // minmod(q - qm, qp - q) (make_rhs.H:57-75): the cell's face values are
    // q +- slope/2.  Every slope is shared by the two faces of the cell.
    // SW = 1 also reports which argument minmod returned (2 bits at bit 2n of
    // *code); SW = 2 takes that choice from *code instead of comparing (a
    // forward with the limiter frozen).
    template <int SW>
    __device__ static __forceinline__ double slope(double qm, double q, double qp, unsigned* code, int n) {
        return slope_ab<SW>(q - qm, qp - q, code, n);
    }

    // The same from the two differences a = q - qm and b = qp - q.
    template <int SW>
    __device__ static __forceinline__ double slope_ab(double a, double b, unsigned* code, int n) {
        if constexpr (SW == 2) {
            const unsigned k = (*code >> (2 * n)) & 3u;
            return (k == 1u) ? a : ((k == 2u) ? b : 0.0);
        } else {
            if constexpr (SW == 1) *code |= minmod_code(a, b) << (2 * n);
            if constexpr (SW == 0 && !kBitwise) return minmod_minmax(a, b);
            return minmod(a, b);
        }
    }

    // The rest of eval_face (make_rhs.H:76-131) on the reconstructed face
    // states qL, qR: Audusse hydrostatic reconstruction, face aux refresh,
    // numerical flux and fluctuations.
    template <int DIR, class N>
    __device__ static __forceinline__ void face_states(VQ qL, VQ qR, const VX& XL, const VX& XR, const VP& p,
                                                       double time, int wb, VQ& flux, VF& fluct, double& hLs,
                                                       double& hRs) {
        hLs = qL(1, 0);
        hRs = qR(1, 0);
        VA aL, aR;
        bool aux_set = false;
        if (wb) {  // Audusse hydrostatic reconstruction, make_rhs.H:84-98
            const double bL = qL(0, 0), bR = qR(0, 0);
            const double bf = dev_max(bL, bR);
            const double hL = qL(1, 0), hR = qR(1, 0);
            hLs = dev_max(0.0, hL + bL - bf);
            hRs = dev_max(0.0, hR + bR - bf);
            double sL, sR;
            if constexpr (kBitwise) {
                sL = (hL > 0.0) ? hLs / hL : 0.0;
                sR = (hR > 0.0) ? hRs / hR : 0.0;
            } else if constexpr (AuxOnlyHinv<M>::value) {
                // A face whose depths before and after the reconstruction all
                // reach eps: hinv of the reconstructed depth is 1/h* there, and
                // r = 1/(h h*) per side gives both the scaling h*/h = h* h* r
                // and 1/h* = h r; both sides' r come from one reciprocal of the
                // product of the four depths (a normal number on this path).
                // Decided per face (a face's value never depends on the faces
                // sharing its wave); only the two factors and the two aux rows
                // differ between the paths.
                const double eps = p(M::zamr_idx_eps, 0);
                const double pL = hL * hLs, pR = hR * hRs;
                const double P = pL * pR;
                if (hL >= eps && hR >= eps && hLs >= eps && hRs >= eps && __builtin_isnormal(P)) {
                    const double R = recip(P);
                    const double rL = pR * R, rR = pL * R;
                    sL = hLs * hLs * rL;
                    sR = hRs * hRs * rR;
                    aL = VA{};
                    aR = VA{};
                    aL(M::zamr_idx_hinv, 0) = hL * rL;
                    aR(M::zamr_idx_hinv, 0) = hR * rR;
                    aux_set = true;
                } else {
                    sL = (hL > 0.0) ? hLs * recip(hL) : 0.0;
                    sR = (hR > 0.0) ? hRs * recip(hR) : 0.0;
                }
            } else {
                sL = (hL > 0.0) ? hLs * recip(hL) : 0.0;
                sR = (hR > 0.0) ? hRs * recip(hR) : 0.0;
            }
#pragma unroll
            for (int d = 0; d < dim; ++d) {
                qL(2 + d, 0) *= sL;
                qR(2 + d, 0) *= sR;
            }
            qL(1, 0) = hLs;
            qR(1, 0) = hRs;
            qL(0, 0) = bf;
            qR(0, 0) = bf;
        }
        if (!aux_set) {
            aL = aux_of(qL, p, time, XL);
            aR = aux_of(qR, p, time, XR);
        }
        VN nh{};
        nh(DIR, 0) = 1.0;
        flux = N::numerical_flux(qL, qR, aL, aR, p, nh);
        if (FluctFromBedJump<N>::value && wb) vq_zero(fluct);
        else fluct = N::numerical_fluctuations(qL, qR, aL, aR, p, nh);
    }

    // The face-normal diffusive flux of add_diffusive_flux (make_rhs.H:184-208)
    // through the face between cells qL and qR along DIR: Model::diffusive_flux
    // at the face average of the two cells' states and aux, with the central
    // difference (qR - qL) / dx as the gradient's only entry, projected on the
    // face normal.  Cell values, not the reconstructed face states.
    template <int DIR>
    __device__ static __forceinline__ VQ diffusive_face(const VQ& qL, const VQ& qR, const VX& XL, const VX& XR,
                                                       double dxf, double rdxf, const VP& p, double time) {
        VQ out;
        if constexpr (M::has_diffusion) {
            const VA aL = aux_of(qL, p, time, XL), aR = aux_of(qR, p, time, XR);
            VQ qf;
            VA af;
#pragma unroll
            for (int v = 0; v < nq; ++v) qf(v, 0) = 0.5 * (qL(v, 0) + qR(v, 0));
#pragma unroll
            for (int v = 0; v < na; ++v) af(v, 0) = 0.5 * (aL(v, 0) + aR(v, 0));
            amrex::SmallMatrix<double, M::n_dof_gradQ, 1> g{};
#pragma unroll
            for (int v = 0; v < nq; ++v)
                g(v * dim + DIR, 0) = kBitwise ? (qR(v, 0) - qL(v, 0)) / dxf : (qR(v, 0) - qL(v, 0)) * rdxf;
            const auto Fd = M::diffusive_flux(qf, af, g, p);
#pragma unroll
            for (int v = 0; v < nq; ++v) out(v, 0) = Fd(v * dim + DIR, 0);
        } else {
            vq_zero(out);
        }
        return out;
    }

    // The throughput build's solve of apply_implicit_source's system
    // (I - dt J) dq = dt S.  A row of I - dt J whose entries are all known at
    // compile time to be those of the identity, with a right-hand side known
    // to be zero, has dq = 0 (the printed models' source-free rows: bed and
    // depth for friction); the remaining rows solve D dq_D = rhs_D over the
    // remaining columns, D being the matrix without the identity rows and
    // their columns (the identity columns multiply dq = 0).  D is eliminated
    // as make_rhs.H:467-499 eliminates the whole matrix (partial pivoting,
    // strict '>' search, failure when the pivot is not > 0 or not finite);
    // each pivot's reciprocal is kept for the back substitution.  Entries
    // known at compile time to be zero stay out of the pivot search (|0| is
    // never > the pivot candidate), the row swaps and the row operations
    // (a zero multiplier or a zero pivot-row entry changes nothing), so a
    // block-diagonal D is eliminated block by block.  The whole
    // matrix is singular exactly when D is, and a non-finite entry anywhere
    // fails the full elimination, so the outcome is the reference's up to
    // rounding: on failure the explicit update q + dt S, or q_old where that
    // is not finite (make_rhs.H:501-508).
    __device__ static __forceinline__ VQ solve_blocks(const VQ& q_old, const VQ& S, double dt, double (&A)[nq][nq],
                                                      double (&rhs)[nq], const bool (&lit)[nq][nq]) {
        bool triv[nq];
#pragma unroll
        for (int r = 0; r < nq; ++r) {
            bool t = zero_literal(rhs[r]);
#pragma unroll
            for (int c = 0; c < nq; ++c) t = t && lit[r][c];
            triv[r] = t;
        }
        bool ok = true;
#pragma unroll
        for (int r = 0; r < nq; ++r)
#pragma unroll
            for (int c = 0; c < nq; ++c)
                if (!triv[r] && triv[c] && !lit[r][c]) ok = ok & isfinite(A[r][c]);
        // Two rows left (momentum friction in two dimensions): D = [a b; c d]
        // by its determinant.  The elimination fails exactly when an entry
        // is not finite, both entries of the pivot column are zero or the
        // second pivot (det / first pivot) is zero, i.e. det == 0 for finite
        // entries; the solution is Cramer's, one reciprocal for both rows.
        int nact = 0, ia = 0, ib = 0;
#pragma unroll
        for (int r = 0; r < nq; ++r) {
            if (triv[r]) continue;
            if (nact == 0) ia = r;
            if (nact == 1) ib = r;
            ++nact;
        }
        if (nact == 2) {
            double a = 0.0, b = 0.0, c = 0.0, d = 0.0, r0 = 0.0, r1 = 0.0;
            bool lb = false, lc = false;
#pragma unroll
            for (int r = 0; r < nq; ++r) {
                if (r == ia) {
                    r0 = rhs[r];
#pragma unroll
                    for (int k = 0; k < nq; ++k) {
                        if (k == ia) a = A[r][k];
                        if (k == ib) { b = A[r][k]; lb = lit[r][k]; }
                    }
                }
                if (r == ib) {
                    r1 = rhs[r];
#pragma unroll
                    for (int k = 0; k < nq; ++k) {
                        if (k == ia) { c = A[r][k]; lc = lit[r][k]; }
                        if (k == ib) d = A[r][k];
                    }
                }
            }
            // Literal-zero couplings drop out (a product with one is not
            // folded by the compiler: it could be NaN for a non-finite
            // factor).  A non-finite entry makes det non-finite (or, for an
            // entry whose partner coupling is a literal zero, the solution),
            // so the checks on det and on the solution cover every entry.
            const double bc = (lb || lc) ? 0.0 : b * c;
            const double det = a * d - bc;
            const double rdet = recip(det);
            const double x0 = (lb ? d * r0 : d * r0 - b * r1) * rdet;
            const double x1 = (lc ? a * r1 : a * r1 - c * r0) * rdet;
            ok = ok & (det != 0.0) & isfinite(det) & isfinite(x0) & isfinite(x1);
            double dq[nq];
#pragma unroll
            for (int n = 0; n < nq; ++n) dq[n] = (n == ia) ? x0 : ((n == ib) ? x1 : 0.0);
            return finish_solve(q_old, S, dt, dq, triv, ok);
        }
        double inv[nq];
#pragma unroll
        for (int k = 0; k < nq; ++k) {
            if (triv[k]) continue;
            int piv = k;
            double best = fabs(A[k][k]);
#pragma unroll
            for (int r = k + 1; r < nq; ++r)
                if (!triv[r] && !zero_literal(A[r][k]) && fabs(A[r][k]) > best) {
                    best = fabs(A[r][k]);
                    piv = r;
                }
            ok = ok & (best > 0.0) & isfinite(best);
            // row swap by predication (registers cannot be indexed), skipped as
            // a whole when no lane of the wave pivots
            if (__any(piv != k)) {
#pragma unroll
                for (int r = k + 1; r < nq; ++r) {
                    if (!triv[r] && !zero_literal(A[r][k]) && r == piv) {
#pragma unroll
                        for (int c = k; c < nq; ++c) {
                            if (triv[c]) continue;
                            const double t = A[k][c];
                            A[k][c] = A[r][c];
                            A[r][c] = t;
                        }
                        const double t = rhs[k];
                        rhs[k] = rhs[r];
                        rhs[r] = t;
                    }
                }
            }
            inv[k] = recip(A[k][k]);
#pragma unroll
            for (int r = k + 1; r < nq; ++r) {
                if (triv[r] || zero_literal(A[r][k])) continue;
                const double f = A[r][k] * inv[k];
#pragma unroll
                for (int c = k + 1; c < nq; ++c)
                    if (!triv[c] && !zero_literal(A[k][c])) A[r][c] -= f * A[k][c];
                if (!zero_literal(rhs[k])) rhs[r] -= f * rhs[k];
            }
        }
        double dq[nq];
#pragma unroll
        for (int r = nq - 1; r >= 0; --r) {
            if (triv[r]) {
                dq[r] = 0.0;
                continue;
            }
            double sum = 0.0;
#pragma unroll
            for (int c = r + 1; c < nq; ++c)
                if (!triv[c] && !zero_literal(A[r][c])) sum += A[r][c] * dq[c];
            dq[r] = (rhs[r] - sum) * inv[r];
            ok = ok & isfinite(dq[r]);
        }
        return finish_solve(q_old, S, dt, dq, triv, ok);
    }

    // q_old + dq where the solve succeeded, else the explicit fallback
    // (make_rhs.H:501-508): q_old + dt S where that is finite, else q_old.
    // A row whose source is a literal zero falls back to q_old (q_old +
    // dt * 0 is q_old up to the sign of a zero, or not finite exactly when
    // q_old is not), and a trivial row (dq = 0) keeps q_old on success:
    // such rows take no arithmetic, and a row that is both keeps q_old on
    // either path.
    __device__ static __forceinline__ VQ finish_solve(const VQ& q_old, const VQ& S, double dt, const double (&dq)[nq],
                                                      const bool (&triv)[nq], bool ok) {
        VQ out;
#pragma unroll
        for (int n = 0; n < nq; ++n) out(n, 0) = triv[n] ? q_old(n, 0) : q_old(n, 0) + dq[n];
        // the fallback is evaluated only by waves with a failed solve
        if (__any(!ok)) {
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                double fb = q_old(n, 0);
                if (!zero_literal(S(n, 0))) {
                    const double v = q_old(n, 0) + dt * S(n, 0);
                    fb = isfinite(v) ? v : q_old(n, 0);
                }
                if (!ok) out(n, 0) = fb;
            }
        }
        return out;
    }

    // apply_implicit_source (make_rhs.H:429-509): one linearised backward-Euler
    // step with Gaussian elimination and partial pivoting; explicit fallback.
    __device__ static __forceinline__ VQ implicit_source(const VQ& q_old, const VA& a_in, double dt, const VP& p,
                                                         double time, const VX& X) {
        const VQ S = M::source(q_old, a_in, p, time, X);
        const auto dS_dQ = M::source_jacobian_wrt_variables(q_old, a_in, p);
        const auto dS_dA = M::source_jacobian_wrt_aux_variables(q_old, a_in, p);
        const auto dA_dQ = M::update_aux_variables_jacobian_wrt_variables(q_old, a_in, p, time, X);
        double A[nq][nq];
        double rhs[nq];
        bool lit[nq][nq];  // entry of I - dt J known at compile time (a literal-zero Jacobian entry)
#pragma unroll
        for (int r = 0; r < nq; ++r) {
            rhs[r] = zero_literal(S(r, 0)) ? 0.0 : dt * S(r, 0);
#pragma unroll
            for (int c = 0; c < nq; ++c) {
                // The chain-rule term is summed first and added once.  Entries
                // the model returns as literal zeros (the printer emits dA/dQ
                // == 0, to_amrex.py:478-480, and sparse Jacobians) are known
                // to the compiler after inlining and skipped: for finite
                // values every entry and the elimination below equal
                // make_rhs.H's arithmetic bit for bit, and the identity rows
                // of I - dt J fold away.
                double chain = 0.0;
                bool has_chain = false;
#pragma unroll
                for (int k = 0; k < na; ++k)
                    if (!zero_literal(dA_dQ(k * nq + c, 0))) {
                        chain += dS_dA(r * na + k, 0) * dA_dQ(k * nq + c, 0);
                        has_chain = true;
                    }
                // (the throughput build adds no all-literal-zero chain term:
                // x + 0 differs from x only for x = -0)
                const double jc = (kBitwise || has_chain) ? dS_dQ(r * nq + c, 0) + chain : dS_dQ(r * nq + c, 0);
                lit[r][c] = zero_literal(jc);
                A[r][c] = lit[r][c] ? (r == c ? 1.0 : 0.0) : (r == c ? 1.0 : 0.0) - dt * jc;
            }
        }
        if constexpr (!kBitwise) return solve_blocks(q_old, S, dt, A, rhs, lit);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < nq; ++k) {
            if (ok) {
                int piv = k;
                double best = fabs(A[k][k]);
#pragma unroll
                for (int r = k + 1; r < nq; ++r)
                    if (fabs(A[r][k]) > best) { best = fabs(A[r][k]); piv = r; }
                if (!(best > 0.0) || !isfinite(best)) {
                    ok = false;
                } else {
                    // Row swap by predication (registers cannot be indexed),
                    // skipped as a whole when no lane of the wave pivots.
                    if (__any(piv != k)) {
#pragma unroll
                        for (int r = k + 1; r < nq; ++r) {
                            if (r == piv) {
#pragma unroll
                                for (int c = k; c < nq; ++c) {
                                    const double t = A[k][c];
                                    A[k][c] = A[r][c];
                                    A[r][c] = t;
                                }
                                const double t = rhs[k];
                                rhs[k] = rhs[r];
                                rhs[r] = t;
                            }
                        }
                    }
                    if constexpr (kBitwise) {
#pragma unroll
                        for (int r = k + 1; r < nq; ++r) {
                            const double f = A[r][k] / A[k][k];
#pragma unroll
                            for (int c = k; c < nq; ++c) A[r][c] -= f * A[k][c];
                            rhs[r] -= f * rhs[k];
                        }
                    } else {
                        const double inv = recip(A[k][k]);
#pragma unroll
                        for (int r = k + 1; r < nq; ++r) {
                            const double f = A[r][k] * inv;
#pragma unroll
                            for (int c = k; c < nq; ++c) A[r][c] -= f * A[k][c];
                            rhs[r] -= f * rhs[k];
                        }
                    }
                }
            }
        }
        VQ out;
        if (ok) {
            double dq[nq];
#pragma unroll
            for (int r = nq - 1; r >= 0; --r) {
                double sum = 0.0;
#pragma unroll
                for (int c = r + 1; c < nq; ++c) sum += A[r][c] * dq[c];
                if constexpr (kBitwise) dq[r] = (rhs[r] - sum) / A[r][r];
                else dq[r] = (rhs[r] - sum) * recip(A[r][r]);
                if (!isfinite(dq[r])) ok = false;
            }
            if (ok) {
#pragma unroll
                for (int n = 0; n < nq; ++n) out(n, 0) = q_old(n, 0) + dq[n];
                return out;
            }
        }
#pragma unroll
        for (int n = 0; n < nq; ++n) {
            const double v = q_old(n, 0) + dt * S(n, 0);
            out(n, 0) = isfinite(v) ? v : q_old(n, 0);
        }
        return out;
    }
};

__device__ __forceinline__ double* qsel(const BoxView& B, int k) { return k == 0 ? B.q0 : (k == 1 ? B.q1 : B.q2); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dev_max(v, __shfl_xor(v, off, 64));
    return v;
}

// A wave's non-negative maximum v into a group's CFL word (the bit pattern
// of a non-negative double: such doubles order as their bit patterns).  One
// atomic per wave on one word per group contends -- a launch has thousands
// of items -- and only a wave that raises the maximum needs it.  The word,
// read here without ordering, is at most its true value: it only grows
// during a launch, from the value it held when the launch started (kernel
// boundaries order the clearing before it), so a stale read costs an atomic
// and never skips a needed one.
__device__ __forceinline__ void group_max(unsigned long long* word, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (v > 0.0 && b > *reinterpret_cast<volatile unsigned long long*>(word)) atomicMax(word, b);
}

// Whole-wave lane shifts (DPP wave_shr / wave_shl): lane i receives lane
// i-1 (from_left) or lane i+1 (from_right); the edge lane receives 0, by the
// bound control bit, so no register has to be cleared first.
__device__ __forceinline__ double from_left(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// XCD-aware item order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; give each XCD a contiguous run of workgroup slots so that
// row-segment neighbours (which share halo rows) hit the same L2.
__device__ __forceinline__ int remap_slot(int b, int nslots) {
    const int per = (nslots + 7) / 8;
    const int x = b & 7, k = b >> 3;
    const int it = x * per + k;
    return (it < nslots) ? it : -1;
}

// Number of workgroup slots for n items (4 waves per workgroup), padded to 8.
__host__ __device__ __forceinline__ int slots_for(int nitems) { return (((nitems + 3) / 4 + 7) / 8) * 8; }

// The fused stage.  EPI: 0 = plain stage (writes dst), 1 = stage + Heun with
// buffer `old` + end-of-step UpdateState + CFL max (order 2, second stage),
// 2 = stage + end-of-step UpdateState + CFL max (order 1 step).
// Occupancy request for the stage kernel (waves per SIMD).  A wave needs
// its VGPRs for three waves to share a SIMD's 512: at most 168.  The 4-row
// order-2 stage without Heun fits (154); the one with Heun, the end-of-step
// update and the CFL max asks for ~200 and runs at two waves unless asked
// for three, when the compiler spills ~20 registers to scratch (in the
// per-item prologue, not the row loop) -- measured at 4096^2 the third wave
// wins (0.51 -> 0.46 ms), so three are requested for it.  Rows beyond four (shallow moments) and the diffusive or recording
// variants keep what the compiler chooses.  ZAMR_STAGE_WAVES > 0 overrides
// the rule for every stage kernel.
#ifndef ZAMR_STAGE_WAVES
#define ZAMR_STAGE_WAVES 0
#endif
template <class M, int ORDER, int EPI, int SW>
constexpr int stage_min_waves() {
    if (ZAMR_STAGE_WAVES > 0) return ZAMR_STAGE_WAVES;
    return (M::n_dof_q <= 4 && M::dimension == 2 && !M::has_diffusion && ORDER == 2 && SW == 0) ? 3 : 1;
}

// SW: 0 = plain; 1 = record every cell's limiter choices into a.sw, read
// as bytes: the four 2-bit codes of the cell's x slope at byte idx and of its
// y slope at byte sc + idx (idx = the cell's index in the box layout, sc the
// component stride); 2 = replay them (single box, for derivative checks).
// WB: 0 / 1 = well-balancing compiled off / on (the driver picks the variant
// from solver.well_balanced); 2 = decided at run time by a.wb.
template <class M, class N, int ORDER, int EPI, int SW = 0, int WB = 2>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(stage_min_waves<M, ORDER, EPI, SW>()))) stage_kernel(const BoxDesc* __restrict__ boxes,
                                                   const WorkItem* __restrict__ items, int nitems,
                                                   StageArgs a, GroupState* __restrict__ gs, int xcd_remap) {
    using S = Sys<M>;
    using VQ = typename S::VQ;
    using VA = typename S::VA;
    using VP = typename S::VP;
    using VF = typename S::VF;
    using VN = typename S::VN;
    constexpr int nq = S::nq;
    constexpr int dim = S::dim;
    constexpr int HX = (ORDER == 2) ? 2 : 1;
    constexpr int TXW = WV - 2 * HX;  // cells per item row
    static_assert(SW == 0 || nq <= 4, "the switch table holds four 2-bit codes per slope");

    VP p;
#pragma unroll
    for (int k = 0; k < S::np; ++k) p(k, 0) = a.p.v[k];
    const int wb = (WB == 2) ? a.wb : WB;

    const int t = threadIdx.x & (WV - 1);
    const int wv = wave_in_block();
    const int nslots = slots_for(nitems);
    for (int blk = blockIdx.x; blk < nslots; blk += gridDim.x) {
        const int slot = xcd_remap ? remap_slot(blk, nslots) : blk;
        if (slot < 0) continue;
        const int it = slot * (NT / WV) + wv;
        if (it >= nitems) continue;
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        GroupState* G = gs + B.group;
        const int c = w.i0 - HX + t;  // this lane's column
        const bool own = (t >= HX) && (t < HX + TXW) && (c <= B.hi0);       // assembles cells
        const bool xf = (t >= HX) && (t <= HX + TXW) && (c <= B.hi0 + 1);  // computes face c-1/2
        // The column this lane loads: its own, or past the grown box the box's
        // last ghost column (the values of lanes past that column reach no
        // face of the box, and the loads stay in bounds and in the edge lane's
        // cache lines)
        const int cl = min(c, B.hi0 + HX);
        if (!G->active) {
            // A step moves the current state to another buffer (order 1:
            // src -> dst; order 2 with three buffers: Q^n in `old` -> dst); a
            // finished group carries its state across so that every group's
            // current buffer stays in step.  (Two-buffer order 2 writes back
            // into `old` itself; the MOOD redo writes the buffer stage 2 just
            // carried into.)
            int from = -1;
            if constexpr (EPI == 2) from = a.src;
            if constexpr (EPI == 1)
                if (a.mood != 2 && !a.dstp && a.dst != a.old) from = a.old;
            if (from >= 0 && own) {
                const double* __restrict__ qs = qsel(B, from);
                double* __restrict__ qd = qsel(B, a.dst);
                for (int j = w.j0; j < w.j0 + w.nj; ++j) {
                    const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
#pragma unroll
                    for (int n = 0; n < nq; ++n) qd[idx + n * B.sc] = qs[idx + n * B.sc];
                }
            }
            continue;
        }
        if (a.mood == 2 && G->troubled == 0) continue;  // the candidate stands
        const double dt = G->dt;
        const double time = G->time;
        const double rdx = 1.0 / B.dx0, rdy = 1.0 / B.dx1;  // uniform per item
        const double hgx = 0.5 * p(0, 0) * rdx, hgy = 0.5 * p(0, 0) * rdy;
        // select, never index: a runtime index would push the descriptor to scratch
        const gdouble* __restrict__ qs = to_global(qsel(B, a.src));
        gdouble* __restrict__ qd = to_global(a.dstp ? a.dstp : qsel(B, a.dst));
        const gdouble* __restrict__ qo = to_global(qsel(B, a.old));
        gdouble* __restrict__ pre = to_global(a.pre);
        unsigned char *swx = nullptr, *swy = nullptr;
        if constexpr (SW != 0) {
            swx = reinterpret_cast<unsigned char*>(a.sw);
            swy = swx + B.sc;
        }
        // lanes whose x slope a face uses, and where the table holds it
        const bool sx_used = (t >= HX - 1) && (t <= HX + TXW) && (c >= B.lo0 - 1) && (c <= B.hi0 + 1);
        const bool sx_rec = own || (t == HX - 1 && c == B.lo0 - 1) || (xf && c == B.hi0 + 1);

        const int jstart = w.j0;
        const int jend = w.j0 + w.nj;
        // y state: W0 = row j, W1 = row j+1 (order 2), sy0 = limited y slope
        // of row j, and the bottom face j-1/2 of row j, evaluated one row
        // earlier.  In two dimensions the loop starts one row below the item
        // (j0 = jstart - 1) and that row only evaluates its top face, the
        // first row's bottom face: every face is evaluated by the one code
        // path of the loop body, so its value cannot depend on where item
        // boundaries fall (with contracted FMAs, a second copy of the face
        // code could round differently).
        VQ W0, W1, sy0;
        vq_zero(W0);
        vq_zero(W1);
        vq_zero(sy0);
        VQ fyB;
        VF dyB;
        vq_zero(fyB);
#pragma unroll
        for (int n = 0; n < 2 * nq; ++n) dyB(n, 0) = 0.0;
        double hyB = 0.0;
        VQ dfyB;  // diffusive flux through the bottom face (diffusive models)
        vq_zero(dfyB);
        const int j0 = (dim == 2) ? jstart - 1 : jstart;
        // the first leading row and the prologue's rows are fetched together
        // and then finished: finish() branches (domain edges), and a load is
        // not moved across a branch, so loading row by row would wait once
        // per row
        const int jlead0 = (dim == 2) ? j0 + HX : jstart;
        typename S::Raw nxt = S::fetch(B, qs, cl, jlead0);
        if constexpr (dim == 2) {
            constexpr int NP = 2 * HX - 1;  // rows j0 - HX + 1 .. j0 + HX - 1
            typename S::Raw R[NP];
#pragma unroll
            for (int r = 0; r < NP; ++r) R[r] = S::fetch(B, qs, cl, j0 - HX + 1 + r);
            VQ P[NP];
#pragma unroll
            for (int r = 0; r < NP; ++r) P[r] = S::finish(B, R[r], cl, j0 - HX + 1 + r, p, a.clamp, time);
            if constexpr (ORDER == 2) {
                if (own) {  // the limited slope of row j0
                    unsigned cm = 0u;
                    const long long im = (long long)(c - B.gl0) + (long long)(j0 - B.gl1) * B.sy;
                    if constexpr (SW == 2) cm = swy[im];
#pragma unroll
                    for (int n = 0; n < nq; ++n)
                        sy0(n, 0) = S::template slope<SW>(P[0](n, 0), P[1](n, 0), P[2](n, 0), &cm, n);
                    if constexpr (SW == 1)
                        if (jstart == B.lo1) swy[im] = (unsigned char)cm;  // the ghost row below the box
                }
                W0 = P[1];
                W1 = P[2];
            } else {
                W0 = P[0];
            }
        }
        // the window and the first leading row enter the row loop landed
        land(W0);
        land(W1);
        land(nxt.v);
        double lam = 0.0;
        int nbad = 0;  // MOOD: troubled cells among this lane's candidate cells
        for (int j = j0; j < jend; ++j) {
            const bool cell = own && j >= jstart;  // this lane updates a cell of this row
            // The leading row, fetched one row ahead.  The fetch is
            // unconditional (past the grown box the last row is read again):
            // a conditional load keeps the old value alive beside the new one.
            const VQ lead = S::finish(B, nxt, cl, (dim == 2) ? j + HX : j, p, a.clamp, time);
            nxt = S::fetch(B, qs, cl, (dim == 2) ? min(j + HX + 1, B.hi1 + HX) : j);
            VQ qn_row;  // stage 2: Q^n of the row for the Heun combine
            if constexpr (dim == 1) W0 = lead;
            const VQ& qc = W0;
            const long long isw = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
            // x slopes of every lane's cell, then the x-face c-1/2 (left face of
            // this column) from the left lane's right face value and this
            // lane's left face value
            VQ eR, eL;
            {
                unsigned cx = 0u;
                if constexpr (SW == 2)
                    if (sx_used) cx = swx[isw];
#pragma unroll
                for (int n = 0; n < nq; ++n) {
                    if constexpr (ORDER == 2) {
                        // the forward difference qp - q of the left lane is
                        // this lane's backward difference q - qm, bit for bit
                        const double b = from_right(qc(n, 0)) - qc(n, 0);
                        const double sx = S::template slope_ab<SW>(from_left(b), b, &cx, n);
                        eR(n, 0) = qc(n, 0) + 0.5 * sx;
                        eL(n, 0) = qc(n, 0) - 0.5 * sx;
                    } else {
                        eR(n, 0) = qc(n, 0);
                        eL(n, 0) = qc(n, 0);
                    }
                }
                if constexpr (SW == 1)
                    if (sx_rec && j >= jstart) swx[isw] = (unsigned char)cx;
            }
            VQ qLx;
#pragma unroll
            for (int n = 0; n < nq; ++n) qLx(n, 0) = from_left(eR(n, 0));
            VQ fxL;
            VF dxL;
            double hxL = 0.0, hl = 0.0;
            if (xf && j >= jstart) {
                S::template face_states<0, N>(qLx, eL, S::center(B, c - 1, j), S::center(B, c, j), p, time, wb,
                                              fxL, dxL, hl, hxL);
            } else {
                vq_zero(fxL);
#pragma unroll
                for (int n = 0; n < 2 * nq; ++n) dxL(n, 0) = 0.0;
            }
            // the left cell needs this face's flux, D- and its own reconstructed
            // depth; fold them into the x part of the divergence right away
            double Fx[nq];
#pragma unroll
            for (int n = 0; n < nq; ++n) {
                const double fR = from_right(fxL(n, 0));
                const double dmR = from_right(dxL(nq + n, 0));
                Fx[n] = -(fR - fxL(n, 0)) - dmR - dxL(n, 0);  // face_divergence, make_rhs.H:142-147
            }
            const double hL_xr = from_right(hl);
            // add_diffusive_flux's x part, (Fd_xr - Fd_xl) per row: the left
            // face by this lane, the right one by the lane to the right
            double Dx[nq];
            if constexpr (M::has_diffusion) {
                VQ qml;
#pragma unroll
                for (int n = 0; n < nq; ++n) qml(n, 0) = from_left(qc(n, 0));
                VQ dfx;
                if (xf) dfx = S::template diffusive_face<0>(qml, qc, S::center(B, c - 1, j), S::center(B, c, j), B.dx0,
                                                            rdx, p, time);
                else vq_zero(dfx);
#pragma unroll
                for (int n = 0; n < nq; ++n) Dx[n] = from_right(dfx(n, 0)) - dfx(n, 0);
            }
            // top y-face j+1/2
            VQ fyT, sy1;
            VF dyT;
            double hyT = 0.0, hyTnext = 0.0;
            vq_zero(sy1);
            VQ dfyT;
            vq_zero(dfyT);
            if constexpr (dim == 2) {
                if (own) {
                    VQ qL, qR;
                    if constexpr (ORDER == 2) {
                        unsigned cy = 0u;
                        const long long iy = isw + B.sy;  // row j+1
                        if constexpr (SW == 2) cy = swy[iy];
#pragma unroll
                        for (int n = 0; n < nq; ++n) {
                            sy1(n, 0) = S::template slope<SW>(W0(n, 0), W1(n, 0), lead(n, 0), &cy, n);
                            qL(n, 0) = W0(n, 0) + 0.5 * sy0(n, 0);
                            qR(n, 0) = W1(n, 0) - 0.5 * sy1(n, 0);
                        }
                        // the next item's prologue records its first row; the
                        // last item records the ghost row above the box
                        if constexpr (SW == 1)
                            if (j + 1 < jend || jend == B.hi1 + 1) swy[iy] = (unsigned char)cy;
                    } else {
                        qL = W0;
                        qR = lead;
                    }
                    S::template face_states<1, N>(qL, qR, S::center(B, c, j), S::center(B, c, j + 1), p, time, wb,
                                                  fyT, dyT, hyT, hyTnext);
                    if constexpr (M::has_diffusion) {
                        if constexpr (ORDER == 2)
                            dfyT = S::template diffusive_face<1>(W0, W1, S::center(B, c, j), S::center(B, c, j + 1),
                                                                 B.dx1, rdy, p, time);
                        else
                            dfyT = S::template diffusive_face<1>(W0, lead, S::center(B, c, j), S::center(B, c, j + 1),
                                                                 B.dx1, rdy, p, time);
                    }
                }
            }
            // The prefetch is landed before the cell update, ahead of this row's
            // stores (fewer live registers than landing it at the stores).
            // Heun's Q^n is loaded here, its latency covered by the cell update
            // and the source solve, and used at the row's end (a load carried
            // across rows would be copied at the loop edge, and that copy waits
            // for the row's stores).
            land(nxt.v);
            if constexpr (EPI == 1) {
                if (cell) {
                    const long long i1 = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
#pragma unroll
                    for (int n = 0; n < nq; ++n) qn_row(n, 0) = qo[i1 + n * B.sc];
                }
            }
            VQ qn;
            if (cell) {
                const typename S::VX X = S::center(B, c, j);
                const VA ac = S::aux_at(B, c, j, qc, p, time, X);
                // The cell's RHS and forward-Euler update.  The explicit
                // source is added in a branch of its own (a.impl is uniform
                // over the launch): added through selects, its adds would run
                // for every cell of an implicit-source stage too.
                auto assemble = [&](auto with_src, const VQ& src) {
                    VQ out;
#pragma unroll
                    for (int n = 0; n < nq; ++n) {
                        double Fy = 0.0;
                        if constexpr (dim == 2) Fy = -(fyT(n, 0) - fyB(n, 0)) - dyT(nq + n, 0) - dyB(n, 0);
                        double rhs = kBitwise ? Fx[n] / B.dx0 + Fy / B.dx1 : Fx[n] * rdx + Fy * rdy;
                        if constexpr (decltype(with_src)::value) rhs += src(n, 0);
                        if (wb) {
                            if constexpr (kBitwise) {
                                if (n == 2) rhs += 0.5 * p(0, 0) * (hL_xr * hL_xr - hxL * hxL) / B.dx0;
                                if constexpr (dim == 2)
                                    if (n == 3) rhs += 0.5 * p(0, 0) * (hyT * hyT - hyB * hyB) / B.dx1;
                            } else {
                                if (n == 2) rhs += (hL_xr * hL_xr - hxL * hxL) * hgx;
                                if constexpr (dim == 2)
                                    if (n == 3) rhs += (hyT * hyT - hyB * hyB) * hgy;
                            }
                        }
                        if constexpr (M::has_diffusion) {  // make_rhs.H:214-227, after the bed source
                            rhs += kBitwise ? Dx[n] / B.dx0 : Dx[n] * rdx;
                            if constexpr (dim == 2)
                                rhs += kBitwise ? (dfyT(n, 0) - dfyB(n, 0)) / B.dx1 : (dfyT(n, 0) - dfyB(n, 0)) * rdy;
                        }
                        out(n, 0) = qc(n, 0) + dt * rhs;
                    }
                    return out;
                };
                if (a.impl) {
                    qn = assemble(std::false_type{}, qc);
                } else {
                    const VQ src = M::source(qc, ac, p, time, X);
                    qn = assemble(std::true_type{}, src);
                }
                const long long idx = isw;
                if (a.pre) {
#pragma unroll
                    for (int n = 0; n < nq; ++n) pre[idx + n * B.sc] = qn(n, 0);
                }
                if (a.impl == 1) qn = S::implicit_source(qn, ac, dt, p, time, X);
                if constexpr (EPI == 1) {
#pragma unroll
                    for (int n = 0; n < nq; ++n) qn(n, 0) = 0.5 * qn_row(n, 0) + 0.5 * qn(n, 0);
                    // troubled: h < 0, NaN (fails h >= 0) or +-inf (fails |h| < 1e300)
                    // (ZoomyAmr.cpp:700-711), on the candidate before UpdateState
                    if (a.mood == 1) nbad += (!(qn(1, 0) >= 0.0) || !(qn(1, 0) < 1.0e300)) ? 1 : 0;
                }
                if constexpr (EPI != 0) {
                    qn = S::update_state(qn, p, a.clamp);
                    const VA af = S::aux_at(B, c, j, qn, p, time, X);
#pragma unroll
                    for (int d = 0; d < dim; ++d) {
                        VN nh{};
                        nh(d, 0) = 1.0;
                        const double ev = N::local_max_abs_eigenvalue(qn, af, p, nh)(0, 0);
                        const double inv = kBitwise ? ev / (d == 0 ? B.dx0 : B.dx1) : ev * (d == 0 ? rdx : rdy);
                        if (inv > lam) lam = inv;
                    }
                }
            }
            if (cell) {
                const long long idx = isw;
                // A first stage's output (read by the next stage only) goes
                // through the L2, where the halves of the cache lines that
                // neighbouring strips share meet before write-back (streaming
                // stores wrote them back as two partial lines: 1.11x the bytes);
                // the step's output streams past the caches.
#pragma unroll
                for (int n = 0; n < nq; ++n) {
                    if constexpr (EPI == 0) qd[idx + n * B.sc] = qn(n, 0);
                    else __builtin_nontemporal_store(qn(n, 0), qd + idx + n * B.sc);
                }
            }
            // carry the top face and the slope down and shift the window
            if constexpr (dim == 2) {
                fyB = fyT;
                dyB = dyT;
                hyB = hyTnext;
                if constexpr (M::has_diffusion) dfyB = dfyT;
                if constexpr (ORDER == 2) {
                    sy0 = sy1;
                    W0 = W1;
                    W1 = lead;
                } else {
                    W0 = lead;
                }
            }
        }
        if constexpr (EPI != 0) {
            lam = wave_max(lam);
            if (t == 0) group_max(&G->maxinv_bits, lam);
        }
        if constexpr (EPI == 1) {
            if (a.mood == 1) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) nbad += __shfl_xor(nbad, off, 64);
                if (t == 0 && nbad > 0) atomicAdd(&G->troubled, nbad);
            }
        }
    }
}

// Lane t of a wave owns column i0 + t of its item (t < txw) in the auxiliary
// kernels below, which walk the same item list without a halo.
#define ZAMR_ITEM_LOOP(nitems)                                                            \
    const int t = threadIdx.x & (WV - 1);                                                 \
    const int wv = wave_in_block();                                                       \
    for (int it = blockIdx.x * (NT / WV) + wv; it < (nitems); it += gridDim.x * (NT / WV))

// ComputeDt's reduction (ZoomyAmr.cpp:465-489) over the state and aux AS
// STORED — used for a step whose dt must see the stored aux (the first step
// after initialisation or after the caller replaced the state).
template <class M, class N>
__global__ void __launch_bounds__(NT) max_inv_kernel(const BoxDesc* __restrict__ boxes,
                                                     const WorkItem* __restrict__ items, int nitems, int src,
                                                     Params pp, GroupState* __restrict__ gs, int txw) {
    using S = Sys<M>;
    constexpr int nq = S::nq, na = S::na, dim = S::dim;
    typename S::VP p;
    for (int k = 0; k < S::np; ++k) p(k, 0) = pp.v[k];
    ZAMR_ITEM_LOOP(nitems) {
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        GroupState* G = gs + B.group;
        const int c = w.i0 + t;
        double lam = 0.0;
        if (t < txw && c <= B.hi0) {
            // four rows per trip: their loads go out together instead of one
            // row's latency per row
#pragma unroll 4
            for (int j = w.j0; j < w.j0 + w.nj; ++j) {
                const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
                typename S::VQ q;
                typename S::VA av;
                for (int n = 0; n < nq; ++n) q(n, 0) = qsel(B, src)[idx + n * B.sc];
                for (int n = 0; n < na; ++n) av(n, 0) = B.aux[idx + n * B.sc];
                double mi = 0.0;
                for (int d = 0; d < dim; ++d) {
                    typename S::VN nh{};
                    nh(d, 0) = 1.0;
                    // same arithmetic as the stage-2 epilogue, so both paths give one dt
                    const double ev = N::local_max_abs_eigenvalue(q, av, p, nh)(0, 0);
                    const double inv = kBitwise ? ev / (d == 0 ? B.dx0 : B.dx1) : ev * (1.0 / (d == 0 ? B.dx0 : B.dx1));
                    if (inv > mi) mi = inv;
                }
                lam = dev_max(lam, mi);
            }
        }
        lam = wave_max(lam);
        if (t == 0) group_max(&G->maxinv_bits, lam);
    }
}

// Clear the CFL maxima of ngroups group states (a fill kernel, so that the
// clearing can sit in a captured graph).
__global__ void clear_maxinv_kernel(GroupState* gs, int ngroups) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroups) gs[g].maxinv_bits = 0ull;
}

// ComputeDt's step size (ZoomyAmr.cpp:494-514) from the maxima of ngroups
// group states: the largest over the groups, a wave-free domain giving dtmax,
// then clamped to [dtmin, dtmax]; written to out[0].  One thread.
__global__ void dt_from_maxinv_kernel(const GroupState* gs, int ngroups, double cfl, double divisor, double dtmin,
                                      double dtmax, double* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double m = 0.0;
    for (int g = 0; g < ngroups; ++g) m = dev_max(m, __longlong_as_double((long long)gs[g].maxinv_bits));
    double dt;
    if (m < 1e-14) {
        dt = dtmax;
    } else {
        dt = cfl / (divisor * m);
        dt = dev_max((dt < dtmax) ? dt : dtmax, dtmin);
    }
    out[0] = dt;
}

// Start of TimeStep for every group: dt from the reduced max (ZoomyAmr.cpp:
// 494-514), capped to time_end (:746-749); the loop condition of Evolve
// (:782) decides whether the group is still active.
__global__ void begin_step_kernel(GroupState* gs, int ngroups) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    GroupState& G = gs[g];
    G.active = (G.time < G.time_end) && (G.max_step < 0 || G.step < G.max_step);
    const double m = __longlong_as_double((long long)G.maxinv_bits);
    G.maxinv_bits = 0ull;
    G.troubled = 0;
    if (!G.active) return;  // dt keeps the group's last step
    double dt;
    if (m < 1e-14) {
        dt = G.dtmax;
    } else {
        dt = G.cfl / (G.divisor * m);
        dt = dev_max((dt < G.dtmax) ? dt : G.dtmax, G.dtmin);
    }
    const double remaining = G.time_end - G.time;
    if (remaining > 0.0 && dt > remaining) dt = remaining;
    G.dt = dt;
}

// End of the step: time += dt, step += 1 (ZoomyAmr.cpp:791-793); the per-step
// (dt, time) pair goes to a history ring the host prints the Step lines from.
__global__ void end_step_kernel(GroupState* gs, int ngroups, double* hist, int hist_cap) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    GroupState& G = gs[g];
    if (!G.active) return;
    G.time += G.dt;
    const int s = G.step;
    G.step = s + 1;
    if (hist && g == 0 && s < hist_cap) {
        hist[2 * s] = G.dt;
        hist[2 * s + 1] = G.time;
    }
}

// MOOD (ZoomyAmr.cpp:699-720): a group whose order-2 candidate has troubled
// cells redoes the step at order 1; its CFL max is recomputed by the redo.
__global__ void mood_reset_kernel(GroupState* gs, int ngroups) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    GroupState& G = gs[g];
    if (G.active && G.troubled > 0) {
        G.maxinv_bits = 0ull;
        G.redos += 1;
    }
}

__global__ void set_dt_kernel(GroupState* gs, int ngroups, double dt, double time) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    gs[g].dt = dt;
    gs[g].time = time;
    gs[g].active = 1;
    gs[g].maxinv_bits = 0ull;
}

// Same-level ghost exchange (FillPatchSingleLevel / FillBoundary): copy the
// owning box's valid cells into another box's ghost cells.  With `update`
// the copy is of the owner's value after UpdateState (do_stage runs
// UpdateState before FillPatch, ZoomyAmr.cpp:563-567), so that the stage
// kernel can take every ghost cell as stored.
// One thread per ghost cell over all rectangles (`start`: a rectangle's first
// cell in the flat numbering), so that many small rectangles share blocks.
struct CopyRect {
    int src_box, dst_box, ilo, jlo, ni, nj;
    long long start;
};
template <class M>
__global__ void __launch_bounds__(256) fill_boundary_kernel(const BoxDesc* __restrict__ boxes,
                                                            const CopyRect* __restrict__ rects, int nrects,
                                                            long long total, int buf, Params pp, int clamp,
                                                            int update) {
    using S = Sys<M>;
    constexpr int nq = S::nq;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    int lo = 0, hi = nrects - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rects[mid].start <= k) lo = mid;
        else hi = mid - 1;
    }
    const CopyRect r = rects[lo];
    const BoxView Sb = view(boxes, r.src_box);
    const BoxView D = view(boxes, r.dst_box);
    const int cell = (int)(k - r.start);  // < ni * nj: 32-bit division
    const int i = r.ilo + cell % r.ni, j = r.jlo + cell / r.ni;
    const long long si = (long long)(i - Sb.gl0) + (long long)(j - Sb.gl1) * Sb.sy;
    const long long di = (long long)(i - D.gl0) + (long long)(j - D.gl1) * D.sy;
    const double* __restrict__ qs = qsel(Sb, buf);
    double* __restrict__ qd = qsel(D, buf);
    typename S::VQ v;
#pragma unroll
    for (int n = 0; n < nq; ++n) v(n, 0) = qs[si + n * Sb.sc];
    if (update) {
        typename S::VP p;
        for (int q = 0; q < S::np; ++q) p(q, 0) = pp.v[q];
        v = S::update_state(v, p, clamp);
    }
#pragma unroll
    for (int n = 0; n < nq; ++n) qd[di + n * D.sc] = v(n, 0);
}

// Qaux as the reference holds it after UpdateState (ZoomyAmr.cpp:328):
// update_aux_variables(Q, Qaux_stored) over the valid cells, at `time` or,
// with gs, at the time of each box's group (the step's start when launched
// before end_step).
template <class M>
__global__ void __launch_bounds__(NT) materialize_aux_kernel(const BoxDesc* __restrict__ boxes,
                                                             const WorkItem* __restrict__ items, int nitems, int src,
                                                             Params pp, double time0, int txw,
                                                             const GroupState* __restrict__ gs) {
    using S = Sys<M>;
    typename S::VP p;
    for (int k = 0; k < S::np; ++k) p(k, 0) = pp.v[k];
    ZAMR_ITEM_LOOP(nitems) {
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        const double time = gs ? gs[B.group].time : time0;
        const int c = w.i0 + t;
        if (!(t < txw && c <= B.hi0)) continue;
        for (int j = w.j0; j < w.j0 + w.nj; ++j) {
            const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
            typename S::VQ q;
            typename S::VA av;
            for (int n = 0; n < S::nq; ++n) q(n, 0) = qsel(B, src)[idx + n * B.sc];
            for (int n = 0; n < S::na; ++n) av(n, 0) = B.aux[idx + n * B.sc];
            av = M::update_aux_variables(q, av, p, time, S::center(B, c, j));
            for (int n = 0; n < S::na; ++n) B.aux[idx + n * B.sc] = av(n, 0);
        }
    }
}

// UpdateState over the valid cells of buffer src (in place) plus the aux.
template <class M>
__global__ void __launch_bounds__(NT) update_state_kernel(const BoxDesc* __restrict__ boxes,
                                                          const WorkItem* __restrict__ items, int nitems, int src,
                                                          Params pp, double time, int clamp, int txw) {
    using S = Sys<M>;
    typename S::VP p;
    for (int k = 0; k < S::np; ++k) p(k, 0) = pp.v[k];
    ZAMR_ITEM_LOOP(nitems) {
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        const int c = w.i0 + t;
        if (!(t < txw && c <= B.hi0)) continue;
        for (int j = w.j0; j < w.j0 + w.nj; ++j) {
            const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
            typename S::VQ q;
            typename S::VA av;
            for (int n = 0; n < S::nq; ++n) q(n, 0) = qsel(B, src)[idx + n * B.sc];
            for (int n = 0; n < S::na; ++n) av(n, 0) = B.aux[idx + n * B.sc];
            q = M::update_variables(q, av, p);
            if (clamp && q(1, 0) < 0.0) q(1, 0) = 0.0;
            av = M::update_aux_variables(q, av, p, time, S::center(B, c, j));
            for (int n = 0; n < S::nq; ++n) qsel(B, src)[idx + n * B.sc] = q(n, 0);
            for (int n = 0; n < S::na; ++n) B.aux[idx + n * B.sc] = av(n, 0);
        }
    }
}

// Per-item partial sums of h over valid cells, written in item order
// (deterministic); the host adds the partials in order.  Cells flagged in
// `cover` (one byte per cell of the level, row length cover_nx) are left
// out: on an AMR level they lie under the next finer level.
template <class M>
__global__ void __launch_bounds__(NT) mass_partial_kernel(const BoxDesc* __restrict__ boxes,
                                                          const WorkItem* __restrict__ items, int nitems, int src,
                                                          double* partial, int txw,
                                                          const unsigned char* __restrict__ cover, int cover_nx) {
    ZAMR_ITEM_LOOP(nitems) {
        const WorkItem w = items[it];
        const BoxView B = view(boxes, w.box);
        const int c = w.i0 + t;
        double s = 0.0;
        if (t < txw && c <= B.hi0)
            for (int j = w.j0; j < w.j0 + w.nj; ++j) {
                if (cover && cover[(long long)j * cover_nx + c]) continue;
                const long long idx = (long long)(c - B.gl0) + (long long)(j - B.gl1) * B.sy;
                s += qsel(B, src)[idx + 1 * B.sc];
            }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (t == 0) partial[it] = s * B.dx0 * B.dx1;
    }
}

}  // namespace zamr
