#pragma once
// The AMR advance of Source/ZoomyAmr.cpp on the device, levels 0 ..
// amr.max_level <= 3 at one amr.ref_ratio of 2 or 4: ErrorEst tagging
// (:225-253), regrid with RemakeLevel / MakeNewLevelFromCoarse / ClearLevel
// (:161-219, AmrCore::regrid), the coarse-fine part of FillPatch and
// FillCoarsePatch with cell_cons_interp (:258-288), nsubsteps = ref_ratio
// (:96-106), the recursive subcycled TimeStep with average_down (:733-762)
// and the composite ComputeTotalMass (:517-536).  Level 0 is the single-level
// Solver; every refined level is a further Solver over its refined domain
// whose boxes follow the tags.  Two levels at ratio 2 take a path of their
// own that overlaps the regrid with level 0's step; deeper hierarchies and
// ratio 4 run the levels one at a time in the reference's order.
//
// The AMR machinery is AMReX's (external to the reference, not vendored);
// its published algorithms are restated here in device code and host code:
// CellConservativeLinear with linear limiting, average_down, TagBoxArray
// buffering and coarsening, MakeNewGrids' proper nesting (the coarsened
// lbase grids shrunk by n_proper, projected finer grids tagged below),
// Berger-Rigoutsos clustering (ClusterList::chop).
//
// Where the reference reads memory it never wrote, this build fixes a value:
// coarse cells beyond the domain in the interpolation stencil repeat the edge
// cell (the foextrap its BCRec declares; its fill functor is NullFill), and
// a stencil neighbour the coarse level does not cover repeats the centre;
// ErrorEst before any stage reads ghost cells as 0, and on a level made or
// remade since its last stage reads the ghosts that fill left (0 beyond the
// domain); a new or remade fine level's Qaux is update_aux_variables(Q, 0)
// (the reference leaves it unset and ComputeDt reads it next).
//
// Data movement: the coarse state is fixed while a finer level subcycles
// (FillPatchTwoLevels is handed the coarse level at t_new only), so its
// coarse-fine ghost cells are interpolated into every fine state buffer once
// per coarse step (two levels) or once per fine step after the regrid that
// reads the old ones (deeper); stage kernels never write ghost cells.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "zamr_driver.hpp"

namespace zamr {

// Level 0 is a tensor-product grid of boxes over its domain (max_grid_size
// chopping); xbox / ybox give the box column / row of each cell index.
struct CoarseGrid {
    const BoxDesc* boxes;
    const int* xbox;
    const int* ybox;
    int nbx, nx, ny;
};

__device__ __forceinline__ int coarse_box_index(const CoarseGrid& g, int ic, int jc) {
    return g.ybox[jc] * g.nbx + g.xbox[ic];
}

// Cells (ilo..ilo+ni-1, jlo..jlo+nj-1) of box `box` (and source box `src`
// for level-to-level copies); `start` is the rectangle's first cell in the
// launch's flat cell numbering.
struct CellRect {
    int box, src, ilo, jlo, ni, nj;
    long long start;
};

__device__ __forceinline__ int find_rect(const CellRect* __restrict__ r, int n, long long k) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (r[mid].start <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// All nq components of coarse cell (ic, jc), clamped into the domain.
template <int NQ>
__device__ __forceinline__ void coarse_cell(const CoarseGrid& g, int buf, int ic, int jc, double* v) {
    ic = ic < 0 ? 0 : (ic >= g.nx ? g.nx - 1 : ic);
    jc = jc < 0 ? 0 : (jc >= g.ny ? g.ny - 1 : jc);
    const BoxDesc& C = g.boxes[coarse_box_index(g, ic, jc)];
    const double* __restrict__ q = buf == 0 ? C.q[0] : (buf == 1 ? C.q[1] : C.q[2]);
    const long long idx = (long long)(ic - C.gl[0]) + (long long)(jc - C.gl[1]) * C.sy;
#pragma unroll
    for (int n = 0; n < NQ; ++n) v[n] = q[idx + n * C.sc];
}

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min

// cell_cons_interp (ratio 2) of the level-0 state into the listed fine
// cells, written to every fine state buffer in `bufmask`:
//   sx = central difference, limited by the one-sided differences times 2;
//   one factor per direction, the least over components of limited/central;
//   fine = crse + xoff * (sx * fx) + yoff * (sy * fy), xoff, yoff = -+1/4.
// With `cover` (one byte per level-0 cell, set under every fine box) cells
// that a fine box covers are skipped: FillPatch takes those from the fine
// level itself.
template <class M>
__global__ void __launch_bounds__(256) cf_interp_kernel(const BoxDesc* __restrict__ fine,
                                                        const CellRect* __restrict__ rects, int nrects,
                                                        long long total, CoarseGrid g, int cbuf, unsigned bufmask,
                                                        const unsigned char* __restrict__ cover) {
#pragma clang fp contract(off)
    constexpr int nq = M::n_dof_q;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);  // < ni * nj: 32-bit division
    const int i = R.ilo + cell % R.ni, j = R.jlo + cell / R.ni;
    const int ic = i >> 1, jc = j >> 1;
    if (cover && cover[(long long)jc * g.nx + ic]) return;
    double u0[nq], sx[nq], sy[nq], xm[nq], xp[nq], ym[nq], yp[nq];
    coarse_cell<nq>(g, cbuf, ic, jc, u0);
    coarse_cell<nq>(g, cbuf, ic - 1, jc, xm);
    coarse_cell<nq>(g, cbuf, ic + 1, jc, xp);
    coarse_cell<nq>(g, cbuf, ic, jc - 1, ym);
    coarse_cell<nq>(g, cbuf, ic, jc + 1, yp);
    double fx = 1.0, fy = 1.0;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        const double u = u0[n];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double um = d == 0 ? xm[n] : ym[n];
            const double up = d == 0 ? xp[n] : yp[n];
            const double dc = 0.5 * (up - um);
            const double df = 2.0 * (up - u);
            const double db = 2.0 * (u - um);
            double sl = (df * db >= 0.0) ? dmin(fabs(df), fabs(db)) : 0.0;
            sl = copysign(1.0, dc) * dmin(sl, fabs(dc));
            if (dc != 0.0) {
                if (d == 0) fx = dmin(fx, sl / dc);
                else fy = dmin(fy, sl / dc);
            }
            if (d == 0) sx[n] = dc;
            else sy[n] = dc;
        }
    }
    const double xoff = ((i - 2 * ic) + 0.5) / 2.0 - 0.5;
    const double yoff = ((j - 2 * jc) + 0.5) / 2.0 - 0.5;
    const BoxView F = view(fine, R.box);
    const long long idx = (long long)(i - F.gl0) + (long long)(j - F.gl1) * F.sy;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        const double v = u0[n] + xoff * (sx[n] * fx) + yoff * (sy[n] * fy);
        if (bufmask & 1u) F.q0[idx + n * F.sc] = v;
        if (bufmask & 2u) F.q1[idx + n * F.sc] = v;
        if (bufmask & 4u) F.q2[idx + n * F.sc] = v;
    }
}

// average_down (ratio 2): every level-0 cell under a fine box takes the mean
// of its four children, summed j-outer, i-inner, times 1/4.  Rectangles are
// in level-0 index space, one per fine box.
template <class M>
__global__ void __launch_bounds__(256) average_down_kernel(const BoxDesc* __restrict__ fine,
                                                           const CellRect* __restrict__ rects, int nrects,
                                                           long long total, CoarseGrid g, int fbuf, int cbuf) {
#pragma clang fp contract(off)
    constexpr int nq = M::n_dof_q;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);
    const int ic = R.ilo + cell % R.ni, jc = R.jlo + cell / R.ni;
    const BoxView F = view(fine, R.box);
    const BoxView C = view(g.boxes, coarse_box_index(g, ic, jc));
    const double* __restrict__ qf = qsel(F, fbuf);
    double* __restrict__ qc = qsel(C, cbuf);
    const long long f00 = (long long)(2 * ic - F.gl0) + (long long)(2 * jc - F.gl1) * F.sy;
    const long long ci = (long long)(ic - C.gl0) + (long long)(jc - C.gl1) * C.sy;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        const long long o = f00 + n * F.sc;
        double sum = 0.0;
        sum += qf[o];
        sum += qf[o + 1];
        sum += qf[o + F.sy];
        sum += qf[o + F.sy + 1];
        qc[ci + n * C.sc] = 0.25 * sum;
    }
}

// Level-to-level copy (RemakeLevel's FillPatch from the old fine level):
// cells of new box `box` from old box `src`, valid cells of the old level.
template <class M>
__global__ void __launch_bounds__(256) copy_levels_kernel(const BoxDesc* __restrict__ dst, int dbuf,
                                                          const BoxDesc* __restrict__ src, int sbuf,
                                                          const CellRect* __restrict__ rects, int nrects,
                                                          long long total) {
    constexpr int nq = M::n_dof_q;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);  // < ni * nj: 32-bit division
    const int i = R.ilo + cell % R.ni, j = R.jlo + cell / R.ni;
    const BoxView D = view(dst, R.box);
    const BoxView S = view(src, R.src);
    const long long di = (long long)(i - D.gl0) + (long long)(j - D.gl1) * D.sy;
    const long long si = (long long)(i - S.gl0) + (long long)(j - S.gl1) * S.sy;
#pragma unroll
    for (int n = 0; n < nq; ++n) qsel(D, dbuf)[di + n * D.sc] = qsel(S, sbuf)[si + n * S.sc];
}

// ErrorEst on level 0 (ZoomyAmr.cpp:225-253): cell (i, j) is tagged when its
// bed is below b_max and |grad h| (undivided central differences) exceeds
// the threshold or h > 1e-2.  Neighbours outside the cell's box read the
// box's ghost cells as the last stage's FillPatch left them (buffer `last`;
// S::load reproduces the stored ghosts and the physical-BC ghosts), or 0
// before any stage ran (last < 0).
template <class M>
__global__ void __launch_bounds__(256) error_est_kernel(CoarseGrid g, int cur, int last, Params pp, int clamp,
                                                        double time_last, double threshold, double b_max,
                                                        unsigned char* __restrict__ tags) {
#pragma clang fp contract(off)
    using S = Sys<M>;
    // one row per blockIdx.y; a wave whose cells lie in one box (the usual
    // case) reads that box's descriptor once, wave-uniformly
    const int j = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nx) return;
    typename S::VP p;
    for (int q = 0; q < S::np; ++q) p(q, 0) = pp.v[q];
    const int k = j * g.nx + i;  // a level's cells fit in 32 bits
    auto cell = [&](const BoxView& B) {
        auto h_at = [&](int c, int r) -> double {
            if (c >= B.lo0 && c <= B.hi0 && r >= B.lo1 && r <= B.hi1) {
                const long long idx = (long long)(c - B.gl0) + (long long)(r - B.gl1) * B.sy;
                return qsel(B, cur)[idx + 1 * B.sc];
            }
            if (last < 0) return 0.0;
            return S::load(B, qsel(B, last), c, r, p, clamp, time_last)(1, 0);
        };
        const long long idx = (long long)(i - B.gl0) + (long long)(j - B.gl1) * B.sy;
        unsigned char t = 0;
        if (qsel(B, cur)[idx] < b_max) {
            double grad = 0.0;
            const double dhdx = fabs(h_at(i + 1, j) - h_at(i - 1, j));
            grad += dhdx * dhdx;
            if constexpr (S::dim >= 2) {
                const double dhdy = fabs(h_at(i, j + 1) - h_at(i, j - 1));
                grad += dhdy * dhdy;
            }
            if (sqrt(grad) > threshold || qsel(B, cur)[idx + B.sc] > 1.0e-2) t = 1;
        }
        tags[k] = t;
    };
    const int bi = coarse_box_index(g, i, j);
    const int b0 = __builtin_amdgcn_readfirstlane(bi);
    if (__all(bi == b0)) cell(view(g.boxes, b0));
    else cell(view(g.boxes, bi));
}

// TagBoxArray buffer + coarsen: coarsened cell (I, J) of factor bf is tagged
// when a tag lies within nbuf cells (in both directions) of its bf x bf block.
// `changed` is set when the map differs from `prev` (the last one clustered).
__global__ void __launch_bounds__(256) tag_buffer_kernel(const unsigned char* __restrict__ tags, int nx, int ny,
                                                         int nbuf, int bf, unsigned char* __restrict__ out, int cnx,
                                                         int cny, const unsigned char* __restrict__ prev,
                                                         unsigned* __restrict__ changed) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;  // a level's cells fit in 32 bits
    const bool in = k0 < cnx * cny;
    const int k = in ? k0 : 0;  // lanes past the end repeat cell 0 so the wave vote stays whole
    const int I = k % cnx, J = k / cnx;
    const int i0 = max(0, I * bf - nbuf), i1 = min(nx - 1, (I + 1) * bf - 1 + nbuf);
    const int j0 = max(0, J * bf - nbuf), j1 = min(ny - 1, (J + 1) * bf - 1 + nbuf);
    unsigned t = 0;
    if (bf == 1 && nbuf == 1) {  // the common case: nine independent loads
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj) {
            const int j = min(max(J + dj, j0), j1);
#pragma unroll
            for (int di = -1; di <= 1; ++di) t |= tags[(long long)j * nx + min(max(I + di, i0), i1)];
        }
    } else {
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) t |= tags[(long long)j * nx + i];
    }
    t = t ? 1u : 0u;
    if (in) out[k] = (unsigned char)t;
    // one atomic per wave, and none once the flag is up
    if (__any(in && t != prev[k]) && (threadIdx.x & (WV - 1)) == 0 &&
        __hip_atomic_load(changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
        atomicOr(changed, 1u);
}

// ---------------- levels above 1, and ref_ratio 4 ----------------
// A refined level as the coarse side of the next one: its boxes and an owner
// map (the box of every cell of the level, -1 where no box is).
struct OwnerGrid {
    const BoxDesc* boxes;
    const int* owner;
    int nx, ny;
};

__device__ __forceinline__ int level_box(const CoarseGrid& g, int ic, int jc) { return coarse_box_index(g, ic, jc); }
__device__ __forceinline__ int level_box(const OwnerGrid& g, int ic, int jc) {
    return g.owner[(long long)jc * g.nx + ic];
}

// All nq components of cell (ic, jc) of a level (clamped into the domain);
// false where the level has no box.
template <int NQ, class G>
__device__ __forceinline__ bool level_cell(const G& g, int buf, int ic, int jc, double* v) {
    ic = ic < 0 ? 0 : (ic >= g.nx ? g.nx - 1 : ic);
    jc = jc < 0 ? 0 : (jc >= g.ny ? g.ny - 1 : jc);
    const int b = level_box(g, ic, jc);
    if (b < 0) return false;
    const BoxView C = view(g.boxes, b);
    const double* __restrict__ q = qsel(C, buf);
    const long long idx = (long long)(ic - C.gl0) + (long long)(jc - C.gl1) * C.sy;
#pragma unroll
    for (int n = 0; n < NQ; ++n) v[n] = q[idx + n * C.sc];
    return true;
}

// cell_cons_interp at ratio r (cf_interp_kernel's rule) of coarse level g
// into the listed cells of a fine level, written to every fine buffer in
// `bufmask`.  With `fine_owner` (the fine level's owner map, fnx cells a
// row) cells a fine box covers are skipped.  A stencil neighbour the coarse
// level does not cover (proper nesting leaves at most one such cell beyond
// it) repeats the stencil's centre, which is always covered; a centre that
// is not leaves the cell alone.
template <class M, class G>
__global__ void __launch_bounds__(256) interp_r_kernel(const BoxDesc* __restrict__ fine,
                                                       const CellRect* __restrict__ rects, int nrects, long long total,
                                                       G g, int r, int cbuf, unsigned bufmask,
                                                       const int* __restrict__ fine_owner, int fnx) {
#pragma clang fp contract(off)
    constexpr int nq = M::n_dof_q;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);
    const int i = R.ilo + cell % R.ni, j = R.jlo + cell / R.ni;
    if (fine_owner && fine_owner[(long long)j * fnx + i] >= 0) return;
    const int ic = i / r, jc = j / r;  // i, j >= 0: inside the domain
    double u0[nq], sx[nq], sy[nq], xm[nq], xp[nq], ym[nq], yp[nq];
    if (!level_cell<nq>(g, cbuf, ic, jc, u0)) return;
    if (!level_cell<nq>(g, cbuf, ic - 1, jc, xm))
        for (int n = 0; n < nq; ++n) xm[n] = u0[n];
    if (!level_cell<nq>(g, cbuf, ic + 1, jc, xp))
        for (int n = 0; n < nq; ++n) xp[n] = u0[n];
    if (!level_cell<nq>(g, cbuf, ic, jc - 1, ym))
        for (int n = 0; n < nq; ++n) ym[n] = u0[n];
    if (!level_cell<nq>(g, cbuf, ic, jc + 1, yp))
        for (int n = 0; n < nq; ++n) yp[n] = u0[n];
    double fx = 1.0, fy = 1.0;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        const double u = u0[n];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double um = d == 0 ? xm[n] : ym[n];
            const double up = d == 0 ? xp[n] : yp[n];
            const double dc = 0.5 * (up - um);
            const double df = 2.0 * (up - u);
            const double db = 2.0 * (u - um);
            double sl = (df * db >= 0.0) ? dmin(fabs(df), fabs(db)) : 0.0;
            sl = copysign(1.0, dc) * dmin(sl, fabs(dc));
            if (dc != 0.0) {
                if (d == 0) fx = dmin(fx, sl / dc);
                else fy = dmin(fy, sl / dc);
            }
            if (d == 0) sx[n] = dc;
            else sy[n] = dc;
        }
    }
    const double rr = (double)r;
    const double xoff = ((i - r * ic) + 0.5) / rr - 0.5;
    const double yoff = ((j - r * jc) + 0.5) / rr - 0.5;
    const BoxView F = view(fine, R.box);
    const long long idx = (long long)(i - F.gl0) + (long long)(j - F.gl1) * F.sy;
#pragma unroll
    for (int n = 0; n < nq; ++n) {
        const double v = u0[n] + xoff * (sx[n] * fx) + yoff * (sy[n] * fy);
        if (bufmask & 1u) F.q0[idx + n * F.sc] = v;
        if (bufmask & 2u) F.q1[idx + n * F.sc] = v;
        if (bufmask & 4u) F.q2[idx + n * F.sc] = v;
    }
}

// average_down at ratio r into coarse level g: the mean of the r x r
// children, summed j-outer, i-inner, times 1/r^2.  Rectangles in the coarse
// index space, one per fine box.
template <class M, class G>
__global__ void __launch_bounds__(256) average_down_r_kernel(const BoxDesc* __restrict__ fine,
                                                             const CellRect* __restrict__ rects, int nrects,
                                                             long long total, G g, int r, int fbuf, int cbuf) {
#pragma clang fp contract(off)
    constexpr int nq = M::n_dof_q;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);
    const int ic = R.ilo + cell % R.ni, jc = R.jlo + cell / R.ni;
    const int b = level_box(g, ic, jc);
    if (b < 0) return;  // proper nesting puts every fine box over coarse boxes
    const BoxView F = view(fine, R.box);
    const BoxView C = view(g.boxes, b);
    const double* __restrict__ qf = qsel(F, fbuf);
    double* __restrict__ qc = qsel(C, cbuf);
    const long long f00 = (long long)(r * ic - F.gl0) + (long long)(r * jc - F.gl1) * F.sy;
    const long long ci = (long long)(ic - C.gl0) + (long long)(jc - C.gl1) * C.sy;
    const double vol = 1.0 / (double)(r * r);
    for (int n = 0; n < nq; ++n) {
        double sum = 0.0;
        for (int jj = 0; jj < r; ++jj)
            for (int ii = 0; ii < r; ++ii) sum += qf[f00 + n * F.sc + (long long)jj * F.sy + ii];
        qc[ci + n * C.sc] = vol * sum;
    }
}

// ErrorEst on a refined level, over the listed valid cells of its boxes,
// into a tag map of nx cells a row.  Neighbours outside the cell's box:
// after a stage (last >= 0) the box's ghost cells as that stage's FillPatch
// stored them, pushed through UpdateState as the end of Advance leaves them
// (the physical-BC ghosts as S::load rebuilds them); on a level made or
// remade since its last stage (last < 0) the ghost cells inside the domain
// as the fill that made it stored them, and 0 beyond the domain.
template <class M>
__global__ void __launch_bounds__(256) error_est_boxes_kernel(const BoxDesc* __restrict__ boxes,
                                                              const CellRect* __restrict__ rects, int nrects,
                                                              long long total, int cur, int last, Params pp, int clamp,
                                                              double time_last, double threshold, double b_max,
                                                              unsigned char* __restrict__ tags, int nx) {
#pragma clang fp contract(off)
    using S = Sys<M>;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);
    const int i = R.ilo + cell % R.ni, j = R.jlo + cell / R.ni;
    typename S::VP p;
    for (int q = 0; q < S::np; ++q) p(q, 0) = pp.v[q];
    const BoxView B = view(boxes, R.box);
    auto h_at = [&](int c, int rw) -> double {
        const long long idx = (long long)(c - B.gl0) + (long long)(rw - B.gl1) * B.sy;
        if (c >= B.lo0 && c <= B.hi0 && rw >= B.lo1 && rw <= B.hi1) return qsel(B, cur)[idx + B.sc];
        if (last < 0) {
            const bool in = c >= B.dlo0 && c <= B.dhi0 && (S::dim < 2 || (rw >= B.dlo1 && rw <= B.dhi1));
            return in ? qsel(B, cur)[idx + B.sc] : 0.0;
        }
        return S::update_state(S::load(B, qsel(B, last), c, rw, p, clamp, time_last), p, clamp)(1, 0);
    };
    const long long idx = (long long)(i - B.gl0) + (long long)(j - B.gl1) * B.sy;
    unsigned char t = 0;
    if (qsel(B, cur)[idx] < b_max) {
        double grad = 0.0;
        const double dhdx = fabs(h_at(i + 1, j) - h_at(i - 1, j));
        grad += dhdx * dhdx;
        if constexpr (S::dim >= 2) {
            const double dhdy = fabs(h_at(i, j + 1) - h_at(i, j - 1));
            grad += dhdy * dhdy;
        }
        if (sqrt(grad) > threshold || qsel(B, cur)[idx + B.sc] > 1.0e-2) t = 1;
    }
    tags[(long long)j * nx + i] = t;
}

// Owner map of a level: every cell of rectangle k (a box's valid cells)
// takes the rectangle's box index.
__global__ void __launch_bounds__(256) paint_owner_kernel(const CellRect* __restrict__ rects, int nrects,
                                                          long long total, int* __restrict__ owner, int nx) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const CellRect R = rects[find_rect(rects, nrects, k)];
    const int cell = (int)(k - R.start);
    const int i = R.ilo + cell % R.ni, j = R.jlo + cell / R.ni;
    owner[(long long)j * nx + i] = R.box;
}

// ---------------- Berger-Rigoutsos on the device ----------------
// One generation = the regions still to be clustered: disjoint rectangles of
// the buffered, coarsened tag map, each holding at least one tag.  Their row
// and column tag counts sit at hist[offx + i - lo0] (column i) and
// hist[offy + j - lo1] (row j); `tile` is the first of the region's 64 x 64
// tiles in the generation's running tile count.
struct Region {
    int lo0, lo1, hi0, hi1;
    int offx, offy, tile;
};

// Control block: generation g works on parity p = g & 1.
struct BrCtl {
    int nreg[2], ntile[2], nh[2];
    int nout;      // clusters found
    int overflow;  // a capacity was exceeded: the host grows the buffers and starts again
    int gens;      // generations that had regions
};

struct BrBufs {
    Region* reg[2];
    int* hist[2];
    Region* child;  // two per region of a generation (the decide kernel's scratch)
    int* nchild;    // per region: 0 or 2
    int4* out;      // clusters (lo0, lo1, hi0, hi1)
    int reg_cap, hist_cap, out_cap;
};

// Generation 0: the whole map as one region (hist_cap >= cnx + cny).
__global__ void __launch_bounds__(256) br_init_kernel(BrCtl* ctl, BrBufs b, int cnx, int cny) {
    const int nh = cnx + cny;
    for (int k = threadIdx.x; k < nh; k += blockDim.x) b.hist[0][k] = 0;
    if (threadIdx.x == 0) {
        b.reg[0][0] = Region{0, 0, cnx - 1, cny - 1, 0, cnx, 0};
        ctl->nreg[0] = 1;
        ctl->ntile[0] = ((cnx + 63) / 64) * ((cny + 63) / 64);
        ctl->nh[0] = nh;
        ctl->nreg[1] = ctl->ntile[1] = ctl->nh[1] = 0;
        ctl->nout = 0;
        ctl->overflow = 0;
        ctl->gens = 0;
    }
}

// Row and column tag counts of generation p's regions: a block per 64 x 64
// tile (grid-stride over the generation's tiles), partial sums in LDS, one
// atomic per row and column of the tile.
__global__ void __launch_bounds__(256) br_hist_kernel(const unsigned char* __restrict__ map, int cnx,
                                                      const BrCtl* __restrict__ ctl, BrBufs b, int p) {
    __builtin_amdgcn_s_setprio(3);  // a short serial chain beside a VALU-bound stage kernel: issue first
    __shared__ int cs[64], rs[64];
    const int nreg = ctl->nreg[p], ntile = ctl->ntile[p];
    const Region* __restrict__ reg = b.reg[p];
    int* __restrict__ hist = b.hist[p];
    const int ci = threadIdx.x & 63;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        int lo = 0, hi = nreg - 1;
        while (lo < hi) {  // the region owning this tile
            const int mid = (lo + hi + 1) >> 1;
            if (reg[mid].tile <= tile) lo = mid;
            else hi = mid - 1;
        }
        const Region R = reg[lo];
        const int tw = (R.hi0 - R.lo0 + 64) / 64;
        const int t = tile - R.tile;
        const int i0 = R.lo0 + (t % tw) * 64, j0 = R.lo1 + (t / tw) * 64;
        if (threadIdx.x < 64) {
            cs[threadIdx.x] = 0;
            rs[threadIdx.x] = 0;
        }
        __syncthreads();
        const int i = i0 + ci;
        int csum = 0;
        for (int r = threadIdx.x >> 6; r < 64; r += 4) {
            const int j = j0 + r;
            if (i <= R.hi0 && j <= R.hi1 && map[(long long)j * cnx + i]) {
                ++csum;
                atomicAdd(&rs[r], 1);
            }
        }
        if (csum) atomicAdd(&cs[ci], csum);
        __syncthreads();
        if (threadIdx.x < 64) {
            if (cs[threadIdx.x]) atomicAdd(&hist[R.offx + (i0 - R.lo0) + threadIdx.x], cs[threadIdx.x]);
        } else if (threadIdx.x < 128) {
            const int r = threadIdx.x - 64;
            if (rs[r]) atomicAdd(&hist[R.offy + (j0 - R.lo1) + r], rs[r]);
        }
        __syncthreads();  // cs / rs are cleared for the next tile
    }
}

constexpr int BR_UNROLL = 8;

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = WV / 2; o; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = WV / 2; o; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int o) {
    const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, o);
    const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), o);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
    for (int o = WV / 2; o; o >>= 1) {
        const unsigned long long w = shfl_xor_u64(v, o);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = WV / 2; o; o >>= 1) {
        const unsigned long long w = shfl_xor_u64(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// FindCut (AMReX_Cluster.cpp) over hist[0..len) by one wave; status 0 hole,
// 1 inflection, 2 bisection, 3 none (len <= 1); returns the cut index (cells
// below it stay on the low side).  The cut is the zero nearest the middle
// (the lower index among equals); else, among sign changes of the second
// difference (0 at both ends) at least 2 cells from either end, the largest
// jump, nearest the middle, lower index among equals, if that jump exceeds
// 2; else the middle.  Keys pack that order (indices < 2^20, jumps < 2^21).
__device__ int br_find_cut_wave(const int* __restrict__ h, int len, int lo, int lane, int& status) {
    status = 3;
    if (len <= 1) return lo;
    const int mid = len / 2;
    // loads go out BR_UNROLL (or 4) at a time per lane: the scans are
    // latency-bound walks over counts in L2
    unsigned long long hole = ~0ull;  // (|i - mid|, i), least
    for (int base = 0; base < len; base += WV * BR_UNROLL) {
        int v[BR_UNROLL];
#pragma unroll
        for (int u = 0; u < BR_UNROLL; ++u) {
            const int i = base + u * WV + lane;
            v[u] = i < len ? h[i] : 1;
        }
#pragma unroll
        for (int u = 0; u < BR_UNROLL; ++u) {
            const int i = base + u * WV + lane;
            if (v[u] == 0) {
                const unsigned long long k = ((unsigned long long)abs(i - mid) << 32) | (unsigned)i;
                hole = k < hole ? k : hole;
            }
        }
    }
    hole = wave_min_u64(hole);
    if (hole != ~0ull) {
        status = 0;
        return lo + (int)(hole & 0xffffffffu);
    }
    unsigned long long infl = 0;  // 1 + (jump, 2^20 - 1 - |i - mid|, 2^20 - 1 - i), greatest; 0: none
    for (int base = 2; base < len - 2; base += WV * 4) {
        int m2[4], m1[4], c0[4], p1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = base + u * WV + lane;
            const bool in = i < len - 2;
            m2[u] = in ? h[i - 2] : 0;
            m1[u] = in ? h[i - 1] : 0;
            c0[u] = in ? h[i] : 0;
            p1[u] = in ? h[i + 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = base + u * WV + lane;
            const int a = c0[u] - 2 * m1[u] + m2[u];  // second difference at i - 1
            const int c = p1[u] - 2 * c0[u] + m1[u];  // at i
            if (i < len - 2 && a * c < 0) {
                const unsigned long long k = 1ull + (((unsigned long long)abs(a - c) << 40) |
                                                     ((unsigned long long)(0xFFFFF - abs(i - mid)) << 20) |
                                                     (unsigned long long)(0xFFFFF - i));
                infl = k > infl ? k : infl;
            }
        }
    }
    infl = wave_max_u64(infl);
    if (infl == 0 || (int)((infl - 1) >> 40) <= 2) {
        status = 2;
        return lo + mid;
    }
    status = 1;
    return lo + (0xFFFFF - (int)((infl - 1) & 0xFFFFF));
}

// One generation's decisions (a single 512-thread block, a wave per region;
// two waves per SIMD at its register count still fit beside the stage
// kernel it overlaps):
// a region's counts give its cluster (the nonzero row and column ranges)
// and efficiency; a cluster at grid_eff or better is output, any other is
// cut (FindCut in both directions, the better status, then the longer
// shorter side, y among equals) into two regions of the next generation,
// compacted in region order by a block scan that also lays out their tiles
// and counts.
constexpr int BR_DECIDE_THREADS = 512;
__global__ void __launch_bounds__(BR_DECIDE_THREADS) br_decide_kernel(BrCtl* ctl, BrBufs b, int p,
                                                                      double grid_eff) {
    constexpr int T = BR_DECIDE_THREADS;
    __builtin_amdgcn_s_setprio(3);
    __shared__ int s_nc[T], s_nt[T], s_nh[T];
    __shared__ int s_over;
    const int t = threadIdx.x, lane = t & (WV - 1), wave = wave_in_block();
    const int nreg = ctl->nreg[p];
    const Region* __restrict__ reg = b.reg[p];
    const int* __restrict__ hist = b.hist[p];
    if (t == 0) s_over = 0;
    __syncthreads();
    for (int r = wave; r < nreg; r += T / WV) {
        const Region R = reg[r];
        const int w = R.hi0 - R.lo0 + 1, hh = R.hi1 - R.lo1 + 1;
        const int* hx = hist + R.offx;
        const int* hy = hist + R.offy;
        int a0 = 0x7fffffff, a1 = -1, b0 = 0x7fffffff, b1 = -1;
        unsigned long long cnt = 0;
        for (int base = 0; base < w; base += WV * BR_UNROLL) {
            int v[BR_UNROLL];
#pragma unroll
            for (int u = 0; u < BR_UNROLL; ++u) {
                const int i = base + u * WV + lane;
                v[u] = i < w ? hx[i] : 0;
            }
#pragma unroll
            for (int u = 0; u < BR_UNROLL; ++u)
                if (v[u]) {
                    const int i = base + u * WV + lane;
                    a0 = min(a0, i);
                    a1 = max(a1, i);
                    cnt += (unsigned)v[u];
                }
        }
        for (int base = 0; base < hh; base += WV * BR_UNROLL) {
            int v[BR_UNROLL];
#pragma unroll
            for (int u = 0; u < BR_UNROLL; ++u) {
                const int j = base + u * WV + lane;
                v[u] = j < hh ? hy[j] : 0;
            }
#pragma unroll
            for (int u = 0; u < BR_UNROLL; ++u)
                if (v[u]) {
                    const int j = base + u * WV + lane;
                    b0 = min(b0, j);
                    b1 = max(b1, j);
                }
        }
        a0 = wave_min_i(a0);
        a1 = wave_max_i(a1);
        b0 = wave_min_i(b0);
        b1 = wave_max_i(b1);
        for (int o = WV / 2; o; o >>= 1) cnt += shfl_xor_u64(cnt, o);
        int nch = 0;
        if (a1 >= 0) {
            const int lo0 = R.lo0 + a0, hi0 = R.lo0 + a1, lo1 = R.lo1 + b0, hi1 = R.lo1 + b1;
            const double vol = (double)(hi0 - lo0 + 1) * (double)(hi1 - lo1 + 1);
            if ((double)(long long)cnt / vol >= grid_eff) {
                if (lane == 0) {
                    const int slot = atomicAdd(&ctl->nout, 1);
                    if (slot < b.out_cap) b.out[slot] = make_int4(lo0, lo1, hi0, hi1);
                    else atomicOr(&s_over, 1);
                }
            } else {
                int st[2];
                const int cut0 = br_find_cut_wave(hx + a0, a1 - a0 + 1, lo0, lane, st[0]);
                const int cut1 = br_find_cut_wave(hy + b0, b1 - b0 + 1, lo1, lane, st[1]);
                const int best = min(st[0], st[1]);
                int dir = -1, minlen = -1;
                if (st[0] == best) {
                    dir = 0;
                    minlen = min(cut0 - lo0, hi0 - cut0);
                }
                if (st[1] == best && min(cut1 - lo1, hi1 - cut1) >= minlen) dir = 1;
                Region A{lo0, lo1, hi0, hi1, 0, 0, 0}, B = A;
                if (dir == 0) {
                    A.hi0 = cut0 - 1;
                    B.lo0 = cut0;
                } else {
                    A.hi1 = cut1 - 1;
                    B.lo1 = cut1;
                }
                if (lane == 0) {
                    b.child[2 * r] = A;
                    b.child[2 * r + 1] = B;
                }
                nch = 2;
            }
        }
        if (lane == 0) b.nchild[r] = nch;
    }
    __syncthreads();
    // compaction: each thread a run of regions; inclusive scans over the runs
    const int C = (nreg + T - 1) / T;
    const int r0 = min(nreg, t * C), r1 = min(nreg, r0 + C);
    auto tiles = [](const Region& q) { return ((q.hi0 - q.lo0 + 64) / 64) * ((q.hi1 - q.lo1 + 64) / 64); };
    auto span = [](const Region& q) { return (q.hi0 - q.lo0 + 1) + (q.hi1 - q.lo1 + 1); };
    int nc = 0, nt = 0, nh = 0;
    for (int r = r0; r < r1; ++r)
        if (b.nchild[r])
            for (int c = 0; c < 2; ++c) {
                const Region q = b.child[2 * r + c];
                ++nc;
                nt += tiles(q);
                nh += span(q);
            }
    s_nc[t] = nc;
    s_nt[t] = nt;
    s_nh[t] = nh;
    __syncthreads();
    for (int o = 1; o < T; o <<= 1) {
        int x = 0, y = 0, z = 0;
        if (t >= o) {
            x = s_nc[t - o];
            y = s_nt[t - o];
            z = s_nh[t - o];
        }
        __syncthreads();
        s_nc[t] += x;
        s_nt[t] += y;
        s_nh[t] += z;
        __syncthreads();
    }
    const int NC = s_nc[T - 1], NT = s_nt[T - 1], NH = s_nh[T - 1];
    const bool over = s_over != 0 || NC > b.reg_cap || NH > b.hist_cap;
    if (!over) {
        int pc = s_nc[t] - nc, pt = s_nt[t] - nt, ph = s_nh[t] - nh;
        Region* __restrict__ nxt = b.reg[p ^ 1];
        for (int r = r0; r < r1; ++r)
            if (b.nchild[r])
                for (int c = 0; c < 2; ++c) {
                    Region q = b.child[2 * r + c];
                    q.offx = ph;
                    q.offy = ph + (q.hi0 - q.lo0 + 1);
                    q.tile = pt;
                    nxt[pc++] = q;
                    pt += tiles(q);
                    ph += span(q);
                }
        int* __restrict__ hn = b.hist[p ^ 1];
        for (int k = t; k < NH; k += T) hn[k] = 0;
    }
    if (t == 0) {
        ctl->nreg[p ^ 1] = over ? 0 : NC;
        ctl->ntile[p ^ 1] = over ? 0 : NT;
        ctl->nh[p ^ 1] = over ? 0 : NH;
        if (over) ctl->overflow = 1;
        if (nreg > 0) ctl->gens += 1;
    }
}

// The level-0 cells under the fine boxes (level-0 index space boxes).
__global__ void cover_paint_kernel(const int4* __restrict__ boxes, unsigned char* __restrict__ cover, int nx) {
    const int4 b = boxes[blockIdx.x];
    const int w = b.z - b.x + 1, n = w * (b.w - b.y + 1);
    for (int k = threadIdx.x; k < n; k += blockDim.x) cover[(long long)(b.y + k / w) * nx + b.x + k % w] = 1;
}

__global__ void clear_max_kernel(GroupState* gs, int ngroups) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroups) gs[g].maxinv_bits = 0ull;
}

// The fine level's substeps start at the coarse step's start time
// (TimeStep(lev + 1, time + (i - 1) * dt_level[lev + 1]), ZoomyAmr.cpp:756-757).
__global__ void sync_clock_kernel(const GroupState* coarse, GroupState* fine) { fine->time = coarse->time; }

// ---------------- host: the hierarchy ----------------
template <class M, class N>
class AmrDriver {
   public:
    using Sol = Solver<M, N>;
    static constexpr int nq = M::n_dof_q;
    static constexpr int na = M::n_dof_qaux;

    Sol& L0;
    std::unique_ptr<Sol> L1;
    int max_level = 1, ref_ratio = 2, blocking_factor = 8, n_error_buf = 1, n_proper = 1;
    double grid_eff = 0.7, tag_threshold = 0.01, tag_b_max = 1.0e30;
    int fine_max_grid_size = 64;

    // level-0 grid tables and tag maps
    int* d_xbox = nullptr;
    int* d_ybox = nullptr;
    int nbx = 0;
    unsigned char* d_tags = nullptr;      // ErrorEst, one byte per level-0 cell
    unsigned char* d_ctags[2] = {nullptr, nullptr};  // buffered and coarsened: current, last clustered
    unsigned* d_changed = nullptr;
    bool have_last = false;               // d_ctags[1] holds the map the current grids came from
    unsigned char* d_cover = nullptr;     // level-0 cells under level 1
    int cnx = 0, cny = 0, bf_lev = 1;
    std::vector<std::array<int, 4>> last_boxes;
    // device clustering: control block (and its pinned copy), buffers, and
    // the generations the last clustering took (the next one's first batch)
    BrCtl* d_brctl = nullptr;
    BrCtl* h_brctl = nullptr;
    int4* h_brout = nullptr;  // pinned copy of the clusters
    BrBufs br{};
    int br_depth = 8;
    unsigned* h_flag = nullptr;
    // the clustering's round trips run on a stream of their own, beside the
    // level-0 step on the main stream
    hipStream_t cstream = nullptr;
    hipEvent_t ev_tag = nullptr;  // recorded behind the tag flag's copy to the host
    int4* d_fboxes = nullptr;  // fine boxes in level-0 index space (cover painting)
    size_t fbox_cap = 0;
    BufferCache fine_cache;  // a retired fine level's device buffers, reused by the next
    std::unique_ptr<HostStage> stage_ = std::make_unique<HostStage>(4u << 20);  // a regrid's uploads

    // level-1 launch lists (device, with their cell totals)
    CellRect* d_cf = nullptr;  // coarse-fine ghost cells
    size_t cf_cap = 0, avg_cap = 0;
    int n_cf = 0;
    long long cf_total = 0;
    CellRect* d_avg = nullptr;  // level-0 cells under each fine box
    int n_avg = 0;
    long long avg_total = 0;

    // Hierarchies deeper than two levels or at ref_ratio 4 (the general
    // path): levels 2 .. finest, and per level 0 .. max_level its tag maps,
    // owner map, coverage by the next level, launch lists and clock.
    bool general = false;
    std::vector<std::unique_ptr<Sol>> up;
    struct GLev {
        int nx = 0, ny = 0, cnx = 0, cny = 0;
        unsigned char* d_tag = nullptr;   // ErrorEst (tagging levels)
        unsigned char* d_ctag = nullptr;  // buffered and coarsened
        int* d_owner = nullptr;           // box of every cell (levels >= 1), -1 where none
        unsigned char* d_cov = nullptr;   // cells under the next finer level
        CellRect* d_valid = nullptr;      // valid cells of every box
        CellRect* d_ghost = nullptr;      // in-domain ghost strips of every box
        CellRect* d_avg = nullptr;        // cells of the level below under every box
        int4* d_fb = nullptr;             // the boxes coarsened into the level below
        size_t valid_cap = 0, ghost_cap = 0, avg_cap = 0, fb_cap = 0;
        int n_valid = 0, n_ghost = 0, n_avg = 0;
        long long valid_total = 0, ghost_total = 0, avg_total = 0;
        bool max_stale = true;            // CFL maxima must be reduced from the stored state
        double t_last = 0.0;              // start time of the level's last step
    };
    std::vector<GLev> gl;

    bool l0_stage_ran = false;  // level-0 ghosts hold a FillPatch
    double t_last_stage = 0.0;  // the time that FillPatch ran at (its step's start)
    bool l0_max_stale = true;   // level-0 CFL max must be reduced from the stored state
    bool l1_max_stale = true;
    long long regrids = 0, remakes = 0;
    double regrid_ms = 0.0;  // host wall-clock inside regrid (tags to host, clustering, remake)
    int plot_step = 0;
    double next_plot_time = 0.0;

    AmrDriver(Sol& l0, const std::string& deck_text) : L0(l0) {
        Deck pp(deck_text);
        pp.query("amr.max_level", max_level);
        pp.query("amr.ref_ratio", ref_ratio);
        pp.query("amr.blocking_factor", blocking_factor);
        pp.query("amr.n_error_buf", n_error_buf);
        pp.query("amr.n_proper", n_proper);
        pp.query("amr.grid_eff", grid_eff);
        pp.query("amr.max_grid_size", fine_max_grid_size);
        pp.query("tagging.threshold", tag_threshold);
        pp.query("tagging.b_max", tag_b_max);
        if (M::dimension != 2) throw Error(5, "AMR (amr.max_level > 0) is built for two-dimensional models only");
        if (ref_ratio != 2 && ref_ratio != 4) throw Error(5, "amr.ref_ratio must be 2 or 4 in this build");
        if (max_level < 1 || max_level > 3) throw Error(5, "amr.max_level must be 1..3 in this build");
        general = max_level > 1 || ref_ratio != 2;
        if (blocking_factor < 1 || n_error_buf < 0 || n_proper < 0 || !(grid_eff > 0.0) || grid_eff > 1.0)
            throw Error(1, "bad amr.blocking_factor / amr.n_error_buf / amr.grid_eff");
        if (L0.nx % blocking_factor != 0 || L0.ny % blocking_factor != 0)
            throw Error(1, "amr.n_cell must be divisible by amr.blocking_factor");
        bf_lev = std::max(1, blocking_factor / ref_ratio);
        cnx = (L0.nx + bf_lev - 1) / bf_lev;
        cny = (L0.ny + bf_lev - 1) / bf_lev;
        // level-0 tensor tables (configure_deck builds hb y-major from chop_axis)
        auto xs = chop_axis(L0.nx, L0.max_grid_size), ys = chop_axis(L0.ny, L0.max_grid_size);
        nbx = (int)xs.size();
        std::vector<int> xb(L0.nx), yb(L0.ny);
        for (int k = 0; k < (int)xs.size(); ++k)
            for (int i = xs[k].first; i <= xs[k].second; ++i) xb[i] = k;
        for (int k = 0; k < (int)ys.size(); ++k)
            for (int j = ys[k].first; j <= ys[k].second; ++j) yb[j] = k;
        ZAMR_HIP(hipMalloc(&d_xbox, xb.size() * sizeof(int)));
        ZAMR_HIP(hipMalloc(&d_ybox, yb.size() * sizeof(int)));
        ZAMR_HIP(hipMemcpy(d_xbox, xb.data(), xb.size() * sizeof(int), hipMemcpyHostToDevice));
        ZAMR_HIP(hipMemcpy(d_ybox, yb.data(), yb.size() * sizeof(int), hipMemcpyHostToDevice));
        const size_t ncell = (size_t)L0.nx * L0.ny;
        ZAMR_HIP(hipMalloc(&d_tags, ncell));
        for (auto& p : d_ctags) {
            ZAMR_HIP(hipMalloc(&p, (size_t)cnx * cny));
            ZAMR_HIP(hipMemset(p, 0, (size_t)cnx * cny));
        }
        ZAMR_HIP(hipMalloc(&d_changed, sizeof(unsigned)));
        ZAMR_HIP(hipHostMalloc(&h_flag, sizeof(unsigned)));
        ZAMR_HIP(hipMalloc(&d_cover, ncell));
        ZAMR_HIP(hipMemset(d_cover, 0, ncell));
        ZAMR_HIP(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking));
        ZAMR_HIP(hipEventCreateWithFlags(&ev_tag, hipEventDisableTiming));
        // the device clustering's packed sort keys hold indices below 2^18
        if (cnx >= (1 << 18) || cny >= (1 << 18) || (long long)cnx * cny >= (1ll << 29))
            throw Error(1, "amr: the level-0 grid over blocking_factor / ref_ratio is too large for the clustering");
        ZAMR_HIP(hipMalloc(&d_brctl, sizeof(BrCtl)));
        ZAMR_HIP(hipHostMalloc(&h_brctl, sizeof(BrCtl)));
        // starting capacities (regions and count entries per generation,
        // clusters); ZAMR_BR_CAP="r,h,o" sets them, e.g. tiny, so that the
        // grow-and-repeat path runs
        int rcap = 4096, hcap = 32 * (cnx + cny), ocap = 4096;
        if (const char* e = std::getenv("ZAMR_BR_CAP")) {
            if (std::sscanf(e, "%d,%d,%d", &rcap, &hcap, &ocap) != 3)
                throw Error(1, "ZAMR_BR_CAP must read r,h,o (three integers)");
        }
        br_alloc(std::max(rcap, 1), std::max(hcap, cnx + cny), std::max(ocap, 1));
        if (general) g_setup();
    }

    // Device clustering buffers at the given capacities (regions per
    // generation, histogram entries per generation, clusters).
    void br_free() {
        for (void* q : {(void*)br.reg[0], (void*)br.reg[1], (void*)br.hist[0], (void*)br.hist[1], (void*)br.child,
                        (void*)br.nchild, (void*)br.out})
            if (q) hipFree(q);
        if (h_brout) hipHostFree(h_brout);
        br = BrBufs{};
        h_brout = nullptr;
    }
    void br_alloc(int reg_cap, int hist_cap, int out_cap) {
        br_free();
        for (auto& q : br.reg) ZAMR_HIP(hipMalloc(&q, (size_t)reg_cap * sizeof(Region)));
        for (auto& q : br.hist) ZAMR_HIP(hipMalloc(&q, (size_t)hist_cap * sizeof(int)));
        ZAMR_HIP(hipMalloc(&br.child, 2 * (size_t)reg_cap * sizeof(Region)));
        ZAMR_HIP(hipMalloc(&br.nchild, (size_t)reg_cap * sizeof(int)));
        ZAMR_HIP(hipMalloc(&br.out, (size_t)out_cap * sizeof(int4)));
        ZAMR_HIP(hipHostMalloc(&h_brout, (size_t)out_cap * sizeof(int4)));
        br.reg_cap = reg_cap;
        br.hist_cap = hist_cap;
        br.out_cap = out_cap;
    }

    ~AmrDriver() {
        hipDeviceSynchronize();  // queued copies from the pinned stage and launches on the buffers below
        up.clear();
        L1.reset();
        for (GLev& g : gl)
            for (void* p : {(void*)g.d_tag, (void*)g.d_ctag, (void*)g.d_owner, (void*)g.d_cov, (void*)g.d_valid,
                            (void*)g.d_ghost, (void*)g.d_avg, (void*)g.d_fb})
                if (p) hipFree(p);
        for (void* p : {(void*)d_xbox, (void*)d_ybox, (void*)d_tags, (void*)d_ctags[0], (void*)d_ctags[1],
                        (void*)d_changed, (void*)d_cover, (void*)d_cf, (void*)d_avg, (void*)d_brctl,
                        (void*)d_fboxes, (void*)d_tmp_rects})
            if (p) hipFree(p);
        br_free();
        for (void* p : {(void*)h_brctl, (void*)h_flag})
            if (p) hipHostFree(p);
        if (cstream) hipStreamDestroy(cstream);
        if (ev_tag) hipEventDestroy(ev_tag);
    }

    CoarseGrid grid() const { return CoarseGrid{L0.d_boxes, d_xbox, d_ybox, nbx, L0.nx, L0.ny}; }
    int finest_level() const { return L1 ? 1 + (int)up.size() : 0; }
    Sol& level(int l) {
        if (l == 0) return L0;
        if (l == 1 && L1) return *L1;
        if (l >= 2 && l - 2 < (int)up.size()) return *up[l - 2];
        throw Error(1, "level " + std::to_string(l) + " does not exist");
    }
    std::unique_ptr<Sol>& level_ptr(int l) { return l == 1 ? L1 : up.at(l - 2); }
    template <class F>
    void for_fine(F&& f) {
        if (L1) f(*L1);
        for (auto& u : up) f(*u);
    }
    hipStream_t stream() const { return L0.stream; }

    static unsigned grid1(long long n) { return (unsigned)((n + 255) / 256); }

    // Upload a rectangle list with its running cell offsets into d (capacity
    // cap rectangles, grown to twice the need when short), through the stage.
    void upload_rects(std::vector<CellRect>& r, CellRect*& d, size_t& cap, int& n, long long& total) {
        total = 0;
        for (auto& x : r) {
            x.start = total;
            total += (long long)x.ni * x.nj;
        }
        n = (int)r.size();
        if (!n) return;
        if (r.size() > cap) {
            ZAMR_HIP(hipStreamSynchronize(stream()));  // queued launches may still read d
            if (d) hipFree(d);
            cap = 2 * r.size();
            ZAMR_HIP(hipMalloc(&d, cap * sizeof(CellRect)));
        }
        stage_->upload(d, r.data(), r.size() * sizeof(CellRect), stream());
    }

    // A rectangle list for one launch, in a scratch buffer that only grows
    // (no allocation or free per regrid), through the stage.  The next call
    // overwrites the buffer only after the launch that reads it (stream
    // order).
    CellRect* d_tmp_rects = nullptr;
    size_t tmp_rects_cap = 0;
    CellRect* upload_scratch(std::vector<CellRect>& r, int& n, long long& total) {
        total = 0;
        for (auto& x : r) {
            x.start = total;
            total += (long long)x.ni * x.nj;
        }
        n = (int)r.size();
        if (!n) return nullptr;
        if (r.size() > tmp_rects_cap) {
            ZAMR_HIP(hipStreamSynchronize(stream()));  // the last launch may still read it
            if (d_tmp_rects) hipFree(d_tmp_rects);
            tmp_rects_cap = r.size() * 2;
            ZAMR_HIP(hipMalloc(&d_tmp_rects, tmp_rects_cap * sizeof(CellRect)));
        }
        stage_->upload(d_tmp_rects, r.data(), r.size() * sizeof(CellRect), stream());
        return d_tmp_rects;
    }

    // cell_cons_interp of level-0 buffer cbuf into the listed cells of level
    // F, into buffers bufmask; with `cover`, cells under a fine box are left
    // alone.
    void interp(Sol& F, CellRect* d, int n, long long total, int cbuf, unsigned bufmask,
                const unsigned char* cover) {
        if (!n || !total) return;
        hipLaunchKernelGGL((cf_interp_kernel<M>), dim3(grid1(total)), dim3(256), 0, stream(), F.d_boxes, d, n, total,
                           grid(), cbuf, bufmask, cover);
        ZAMR_HIP(hipGetLastError());
    }

    // Launch lists of the current level 1: its ghost strips (four per box,
    // inside the domain; the interpolation skips the cells another fine box
    // covers), the level-0 cells under each box (average_down), and the
    // coverage map.
    void build_fine_lists() {
        std::vector<CellRect> cf, avg;
        const auto& hb = L1->hb;
        const int NX = L1->nx, NY = L1->ny;
        auto strip = [&](int b, int i0, int j0, int i1, int j1) {
            i0 = std::max(i0, 0);
            j0 = std::max(j0, 0);
            i1 = std::min(i1, NX - 1);
            j1 = std::min(j1, NY - 1);
            if (i0 <= i1 && j0 <= j1) cf.push_back(CellRect{b, b, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
        };
        std::vector<int4> fb;
        for (size_t b = 0; b < hb.size(); ++b) {
            const HostBox& B = hb[b];
            const int g = B.ng;
            strip((int)b, B.lo[0] - g, B.lo[1] - g, B.hi[0] + g, B.lo[1] - 1);  // below, corners included
            strip((int)b, B.lo[0] - g, B.hi[1] + 1, B.hi[0] + g, B.hi[1] + g);  // above
            strip((int)b, B.lo[0] - g, B.lo[1], B.lo[0] - 1, B.hi[1]);          // left
            strip((int)b, B.hi[0] + 1, B.lo[1], B.hi[0] + g, B.hi[1]);          // right
            avg.push_back(CellRect{(int)b, (int)b, B.lo[0] / 2, B.lo[1] / 2, (B.hi[0] - B.lo[0] + 1) / 2,
                                   (B.hi[1] - B.lo[1] + 1) / 2, 0});
            fb.push_back(int4{B.lo[0] / 2, B.lo[1] / 2, B.hi[0] / 2, B.hi[1] / 2});
        }
        upload_rects(cf, d_cf, cf_cap, n_cf, cf_total);
        upload_rects(avg, d_avg, avg_cap, n_avg, avg_total);
        if (fb.size() > fbox_cap) {
            if (d_fboxes) hipFree(d_fboxes);
            fbox_cap = fb.size() * 2;
            ZAMR_HIP(hipMalloc(&d_fboxes, fbox_cap * sizeof(int4)));
        }
        stage_->upload(d_fboxes, fb.data(), fb.size() * sizeof(int4), stream());
        ZAMR_HIP(hipMemsetAsync(d_cover, 0, (size_t)L0.nx * L0.ny, stream()));
        hipLaunchKernelGGL(cover_paint_kernel, dim3((unsigned)fb.size()), dim3(256), 0, stream(), d_fboxes, d_cover,
                           L0.nx);
        ZAMR_HIP(hipGetLastError());
    }

    // Retire level 1; its device buffers go to fine_cache for the next one.
    void retire_fine() {
        if (!L1) return;
        L1.reset();
    }

    void clear_fine() {
        retire_fine();
        ZAMR_HIP(hipMemsetAsync(d_cover, 0, (size_t)L0.nx * L0.ny, stream()));
        last_boxes.clear();
    }

    // Berger-Rigoutsos on the device map d_ctags[1] (complete: tag_result()
    // waited for it), all on the clustering stream: generation 0 is the whole
    // map; each generation is a histogram launch and a decide launch, queued
    // in batches (the first as deep as the last clustering went) with one
    // copy of the control block to the host per batch.  A generation that
    // would exceed a capacity stops the run; the buffers grow and it starts
    // again.
    std::vector<std::array<int, 4>> gpu_cluster() { return gpu_cluster(d_ctags[1], cnx, cny); }
    std::vector<std::array<int, 4>> gpu_cluster(const unsigned char* map, int mx, int my) {
        if (br.hist_cap < mx + my) br_alloc(br.reg_cap, 32 * (mx + my), br.out_cap);
        for (;;) {
            hipLaunchKernelGGL(br_init_kernel, dim3(1), dim3(256), 0, cstream, d_brctl, br, mx, my);
            ZAMR_HIP(hipGetLastError());
            const int peek = std::min(br.out_cap, 1024);
            int p = 0, gens = 0;
            for (;;) {
                const int k = gens == 0 ? std::max(br_depth, 1) : 4;
                for (int g = 0; g < k; ++g, p ^= 1) {
                    hipLaunchKernelGGL(br_hist_kernel, dim3(1024), dim3(256), 0, cstream, map, mx, d_brctl, br, p);
                    hipLaunchKernelGGL(br_decide_kernel, dim3(1), dim3(BR_DECIDE_THREADS), 0, cstream, d_brctl, br,
                                       p, grid_eff);
                }
                ZAMR_HIP(hipGetLastError());
                gens += k;
                ZAMR_HIP(hipMemcpyAsync(h_brctl, d_brctl, sizeof(BrCtl), hipMemcpyDeviceToHost, cstream));
                ZAMR_HIP(hipMemcpyAsync(h_brout, br.out, (size_t)peek * sizeof(int4), hipMemcpyDeviceToHost, cstream));
                ZAMR_HIP(hipStreamSynchronize(cstream));
                if (h_brctl->overflow || h_brctl->nreg[p] == 0) break;
            }
            if (h_brctl->overflow) {
                br_alloc(4 * br.reg_cap, 4 * br.hist_cap, 4 * br.out_cap);
                continue;
            }
            const int n = h_brctl->nout;
            if (n > peek) {
                ZAMR_HIP(hipMemcpyAsync(h_brout, br.out, (size_t)n * sizeof(int4), hipMemcpyDeviceToHost, cstream));
                ZAMR_HIP(hipStreamSynchronize(cstream));
            }
            br_depth = h_brctl->gens;
            std::vector<std::array<int, 4>> out(n);
            for (int k = 0; k < n; ++k) out[k] = {h_brout[k].x, h_brout[k].y, h_brout[k].z, h_brout[k].w};
            return out;
        }
    }

    // Fine boxes from the clusters of the buffered, coarsened tag map:
    // refined by bf_lev * ref_ratio, cut to amr.max_grid_size in cluster units
    // (in x to a whole number of stage strips within it).
    std::vector<std::array<int, 4>> make_grids() {
        auto cl = gpu_cluster();
        std::sort(cl.begin(), cl.end(), [](const std::array<int, 4>& a, const std::array<int, 4>& b) {
            return a[1] != b[1] ? a[1] < b[1] : a[0] < b[0];
        });
        const int r = ref_ratio * bf_lev;
        const int m = std::max(1, fine_max_grid_size / r);
        // x pieces as wide as the most whole stage-kernel strips (60 cells at
        // order 2) within max_grid_size allow, when that is a whole number of
        // cluster units: only the union of the boxes enters the results, and
        // a strip-aligned width keeps every lane of a strip on a cell
        const int strip = WV - 2 * L0.hx_of(L0.spatial_order);
        const int wx = (fine_max_grid_size / strip) * strip;
        const int mx = (wx >= r && wx % r == 0) ? wx / r : m;
        std::vector<std::array<int, 4>> fine;
        for (auto& c : cl)
            for (int y0 = c[1]; y0 <= c[3]; y0 += m)
                for (int x0 = c[0]; x0 <= c[2]; x0 += mx) {
                    const int x1 = std::min(x0 + mx - 1, c[2]), y1 = std::min(y0 + m - 1, c[3]);
                    fine.push_back({x0 * r, y0 * r, (x1 + 1) * r - 1, (y1 + 1) * r - 1});
                }
        return fine;
    }

    // ErrorEst + buffer/coarsen on the device into d_ctags[0] (tag_launch);
    // tag_result waits for the flag that says whether the map differs from
    // the one the current grids were made from (only that flag crosses to the
    // host) and returns it.  Work queued on the main stream in between runs
    // on without a host round trip.
    bool tag(int mode_last, double time_last) {
        tag_launch(mode_last, time_last);
        return tag_result();
    }

    void tag_launch(int mode_last, double time_last) {
        const long long n = (long long)L0.nx * L0.ny;
        hipLaunchKernelGGL((error_est_kernel<M>), dim3((unsigned)((L0.nx + 255) / 256), (unsigned)L0.ny), dim3(256), 0,
                           stream(), grid(), L0.cur, mode_last,
                           L0.prm, L0.clamp_flag(), time_last, tag_threshold, tag_b_max, d_tags);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipMemsetAsync(d_changed, 0, sizeof(unsigned), stream()));
        hipLaunchKernelGGL(tag_buffer_kernel, dim3(grid1((long long)cnx * cny)), dim3(256), 0, stream(), d_tags,
                           L0.nx, L0.ny, n_error_buf, bf_lev, d_ctags[0], cnx, cny, d_ctags[1], d_changed);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipMemcpyAsync(h_flag, d_changed, sizeof(unsigned), hipMemcpyDeviceToHost, stream()));
        ZAMR_HIP(hipEventRecord(ev_tag, stream()));
    }

    bool tag_result() {
        ZAMR_HIP(hipEventSynchronize(ev_tag));  // the map in d_ctags[0] is complete too
        const unsigned changed = *h_flag;
        if (have_last && !changed) return false;
        std::swap(d_ctags[0], d_ctags[1]);  // [1]: the map clustered now, compared against next time
        have_last = true;
        return true;
    }

    // Qaux of a new or remade level: update_aux_variables(Q, 0) at `time`
    // (its aux storage starts zeroed).
    void fresh_aux(Sol& F, double time) {
        auto& it = F.items(F.hx_of(F.spatial_order));
        hipLaunchKernelGGL((materialize_aux_kernel<M>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, stream(), F.d_boxes,
                           it.d, it.n, F.cur, F.prm, time, WV - 2 * it.hx, (const GroupState*)nullptr);
        ZAMR_HIP(hipGetLastError());
        F.aux_stale = false;
    }

    std::unique_ptr<Sol> new_level(const std::vector<std::array<int, 4>>& boxes) {
        auto F = std::make_unique<Sol>();
        F->hist_cap = 16;  // the step history is level 0's
        F->cache = &fine_cache;
        F->stage = stage_.get();
        // room for the next regrid's level, so that most remakes reuse the pool
        // (a new pool costs ~0.4 ms of allocation at 2M fine cells)
        F->pool_headroom = 1.25;
        F->configure_fine(L0, ref_ratio, boxes);
        return F;
    }

    // FillCoarsePatch: valid and in-domain ghost cells of every box of F
    // from level-0 buffer cbuf, into F's current buffer.
    void fill_coarse_patch(Sol& F, int cbuf) {
        std::vector<CellRect> all;
        for (size_t b = 0; b < F.hb.size(); ++b) {
            const HostBox& B = F.hb[b];
            const int i0 = std::max(0, B.lo[0] - B.ng), i1 = std::min(F.nx - 1, B.hi[0] + B.ng);
            const int j0 = std::max(0, B.lo[1] - B.ng), j1 = std::min(F.ny - 1, B.hi[1] + B.ng);
            all.push_back(CellRect{(int)b, (int)b, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
        }
        int n = 0;
        long long total = 0;
        CellRect* d = upload_scratch(all, n, total);
        interp(F, d, n, total, cbuf, 1u << F.cur, nullptr);
    }

    // AmrCore::regrid(0, time) for level 1, in two parts: regrid_tags
    // (ErrorEst on level 0 as it stands; false when the tags are those the
    // current grids came from, which give the same grids) and regrid_grids
    // (clustering, then MakeNewLevelFromCoarse / RemakeLevel / ClearLevel)
    // with level 0 at `time` read from its buffer cbuf.  Work on the main
    // stream that leaves cbuf and the fine level alone may run in between.
    struct Clock {
        double& acc;
        std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
        ~Clock() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    };

    bool regrid_tags(int mode_last, double time_last) {
        if (max_level < 1) return false;
        Clock clk{regrid_ms};
        ++regrids;
        return tag(mode_last, time_last);
    }

    void regrid(int mode_last, double time_last, double time) {
        if (general) return g_regrid(0, time);
        if (regrid_tags(mode_last, time_last)) regrid_grids(time, L0.cur);
    }

    void regrid_grids(double time, int cbuf) {
        Clock clk{regrid_ms};
        stage_->reset();  // tag_result() waited on an event behind every earlier upload
        auto boxes = make_grids();
        if (boxes.empty()) {
            clear_fine();
            return;
        }
        if (L1 && boxes == last_boxes) return;
        auto F = new_level(boxes);
        fill_coarse_patch(*F, cbuf);
        if (L1) {  // RemakeLevel: the old level's valid cells where it covers
            std::vector<CellRect> cp;
            for (size_t b = 0; b < F->hb.size(); ++b) {
                const HostBox& B = F->hb[b];
                for (size_t o = 0; o < L1->hb.size(); ++o) {
                    const HostBox& O = L1->hb[o];
                    const int i0 = std::max(B.lo[0] - B.ng, O.lo[0]), i1 = std::min(B.hi[0] + B.ng, O.hi[0]);
                    const int j0 = std::max(B.lo[1] - B.ng, O.lo[1]), j1 = std::min(B.hi[1] + B.ng, O.hi[1]);
                    if (i0 > i1 || j0 > j1) continue;
                    cp.push_back(CellRect{(int)b, (int)o, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
                }
            }
            int n = 0;
            long long total = 0;
            CellRect* d = upload_scratch(cp, n, total);
            if (n)
                hipLaunchKernelGGL((copy_levels_kernel<M>), dim3(grid1(total)), dim3(256), 0, stream(), F->d_boxes,
                                   F->cur, L1->d_boxes, L1->cur, d, n, total);
            ZAMR_HIP(hipGetLastError());
            ++remakes;
        }
        fresh_aux(*F, time);
        retire_fine();
        L1 = std::move(F);
        last_boxes = boxes;
        l1_max_stale = true;
        build_fine_lists();
    }

    // The level loop of InitData (ZoomyAmr.cpp:137-155) on the current
    // level-0 state; its ghosts read as 0 there.
    void init_levels() {
        if (general) return g_init_levels();
        L0.ensure_aux();
        have_last = false;
        for (int it = 0; it < 4; ++it) {
            if (L1) fill_coarse_patch(*L1, L0.cur);
            const int old_finest = finest_level();
            regrid(-1, 0.0, 0.0);
            if (old_finest == finest_level() && it > 0) break;
        }
        if (L1) {
            fill_coarse_patch(*L1, L0.cur);
            fresh_aux(*L1, 0.0);
            ZAMR_HIP(hipStreamSynchronize(stream()));
            L1->reset_clock();
        }
        l0_stage_ran = false;
        l0_max_stale = true;
        l1_max_stale = true;
    }

    // Average-down target, CFL reductions and clocks of one coarse step
    // (ZoomyAmr::TimeStep for lev 0 and its level-1 substeps).
    // `fine_cb` runs after each fine substep (for the MOOD lines of Evolve).
    // `groups_fresh`: the caller has just downloaded level 0's group state.
    template <class CB>
    void coarse_step(int iteration, CB&& fine_cb, bool groups_fresh = false) {
        if (general) return g_time_step(0, iteration, false);
        if (!groups_fresh) L0.download_groups();
        const double t0 = L0.h_groups[0].time;
        // The regrid at t0 comes first (ZoomyAmr.cpp:736); of it only the tags
        // must precede the level-0 step, which refills the ghosts they read.
        // The level-0 stages are queued behind the tag kernels before the host
        // waits for the tag flag, and the clustering's host round trips and the
        // new level's fill from Q^n (its buffer c0 outlives the step: order 1
        // writes the other buffer, order 2 rotates three) overlap them.
        const bool tagging = iteration > 0 && max_level >= 1;
        const int c0 = L0.cur;
        const bool overlap = L0.spatial_order == 1 || L0.nbuf() == 3;
        if (tagging) {
            Clock clk{regrid_ms};
            ++regrids;
            tag_launch(l0_stage_ran ? L0.last_in : -1, t_last_stage);
            if (!overlap && tag_result()) regrid_grids(t0, c0);
        }
        t_last_stage = t0;
        hipStream_t s = stream();
        if (l0_max_stale || !L0.maxinv_fresh) {
            L0.ensure_aux();
            hipLaunchKernelGGL(clear_max_kernel, dim3(1), dim3(64), 0, s, L0.d_groups, L0.ngroups);
            auto& it = L0.items(L0.hx_of(L0.spatial_order));
            hipLaunchKernelGGL((max_inv_kernel<M, N>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, s, L0.d_boxes, it.d,
                               it.n, L0.cur, L0.prm, L0.d_groups, WV - 2 * it.hx);
            ZAMR_HIP(hipGetLastError());
        }
        L0.begin_step();
        L0.advance_kernels();
        if (tagging && overlap) {
            bool changed;
            {
                Clock clk{regrid_ms};
                changed = tag_result();
            }
            if (changed) regrid_grids(t0, c0);
        }
        // level 0's clock still reads t0 here (end_step below advances it)
        if (L1) hipLaunchKernelGGL(sync_clock_kernel, dim3(1), dim3(1), 0, s, L0.d_groups, L1->d_groups);
        if (L1) {
            // the reference's end-of-Advance UpdateState leaves Qaux = aux(Q^{n+1});
            // ComputeDt of the next step reads it beside the averaged-down Q
            auto& it = L0.items(L0.hx_of(L0.spatial_order));
            hipLaunchKernelGGL((materialize_aux_kernel<M>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, s, L0.d_boxes,
                               it.d, it.n, L0.cur, L0.prm, 0.0, WV - 2 * it.hx, (const GroupState*)L0.d_groups);
            ZAMR_HIP(hipGetLastError());
        }
        L0.end_step();
        L0.aux_stale = !L1;
        L0.maxinv_fresh = true;
        l0_stage_ran = true;
        l0_max_stale = false;
        if (L1) {
            interp(*L1, d_cf, n_cf, cf_total, L0.cur, L1->nbuf() == 3 ? 7u : 3u, d_cover);
            if (l1_max_stale) {
                hipLaunchKernelGGL(clear_max_kernel, dim3(1), dim3(64), 0, s, L1->d_groups, L1->ngroups);
                auto& it = L1->items(L1->hx_of(L1->spatial_order));
                hipLaunchKernelGGL((max_inv_kernel<M, N>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, s, L1->d_boxes,
                                   it.d, it.n, L1->cur, L1->prm, L1->d_groups, WV - 2 * it.hx);
                ZAMR_HIP(hipGetLastError());
            }
            for (int sub = 0; sub < ref_ratio; ++sub) {
                L1->begin_step();
                L1->advance_kernels();
                L1->end_step();
                fine_cb(sub);
            }
            l1_max_stale = false;
            L1->maxinv_fresh = true;
            hipLaunchKernelGGL((average_down_kernel<M>), dim3(grid1(avg_total)), dim3(256), 0, s, L1->d_boxes, d_avg,
                               n_avg, avg_total, grid(), L1->cur, L0.cur);
            ZAMR_HIP(hipGetLastError());
            l0_max_stale = true;
        }
    }

    // ================= the general path: levels 2..3, ref_ratio 4 =================
    // The same steps as the two-level path above, one level at a time in the
    // reference's order (no regrid beside a step): TimeStep(lev) regrids the
    // levels above lev (AmrCore::regrid(lev): MakeNewGrids with proper
    // nesting, then Remake / MakeNewLevelFromCoarse / Clear), steps lev,
    // subcycles lev + 1 ref_ratio times and averages it down.  A refined
    // level's coarse-fine ghosts are interpolated at the start of each of its
    // steps, after the regrid whose ErrorEst reads them as the last step left
    // them.

    int lev_nx(int l) const {
        int v = L0.nx;
        for (int k = 0; k < l; ++k) v *= ref_ratio;
        return v;
    }
    int lev_ny(int l) const {
        int v = L0.ny;
        for (int k = 0; k < l; ++k) v *= ref_ratio;
        return v;
    }

    void g_setup() {
        gl.assign(max_level + 1, GLev{});
        for (int l = 0; l <= max_level; ++l) {
            GLev& g = gl[l];
            g.nx = lev_nx(l);
            g.ny = lev_ny(l);
            const long long cells = (long long)g.nx * g.ny;
            if (cells >= (1ll << 31)) throw Error(1, "amr: a refined level's index space exceeds 2^31 cells");
            g.cnx = (g.nx + bf_lev - 1) / bf_lev;
            g.cny = (g.ny + bf_lev - 1) / bf_lev;
            if (l < max_level) {
                if (g.cnx >= (1 << 18) || g.cny >= (1 << 18) || (long long)g.cnx * g.cny >= (1ll << 29))
                    throw Error(1, "amr: a level's grid over blocking_factor / ref_ratio is too large for the clustering");
                ZAMR_HIP(hipMalloc(&g.d_tag, (size_t)cells));
                ZAMR_HIP(hipMalloc(&g.d_ctag, (size_t)g.cnx * g.cny));
                ZAMR_HIP(hipMalloc(&g.d_cov, (size_t)cells));
                ZAMR_HIP(hipMemset(g.d_cov, 0, (size_t)cells));
            }
            if (l >= 1) {
                ZAMR_HIP(hipMalloc(&g.d_owner, (size_t)cells * sizeof(int)));
                ZAMR_HIP(hipMemset(g.d_owner, 0xff, (size_t)cells * sizeof(int)));
            }
        }
    }

    // Upload a rectangle list into a GLev buffer (grown when short).
    void g_upload(std::vector<CellRect>& r, CellRect*& d, size_t& cap, int& n, long long& total) {
        upload_rects(r, d, cap, n, total);
    }

    // Launch lists, owner map and the coverage map below, of level l >= 1.
    void g_build_lists(int l) {
        Sol& F = level(l);
        GLev& g = gl[l];
        const int r = ref_ratio;
        std::vector<CellRect> valid, ghost, avg;
        std::vector<int4> fb;
        auto strip = [&](int b, int i0, int j0, int i1, int j1) {
            i0 = std::max(i0, 0);
            j0 = std::max(j0, 0);
            i1 = std::min(i1, g.nx - 1);
            j1 = std::min(j1, g.ny - 1);
            if (i0 <= i1 && j0 <= j1) ghost.push_back(CellRect{b, b, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
        };
        for (size_t b = 0; b < F.hb.size(); ++b) {
            const HostBox& B = F.hb[b];
            const int n = B.ng;
            valid.push_back(CellRect{(int)b, (int)b, B.lo[0], B.lo[1], B.hi[0] - B.lo[0] + 1, B.hi[1] - B.lo[1] + 1, 0});
            strip((int)b, B.lo[0] - n, B.lo[1] - n, B.hi[0] + n, B.lo[1] - 1);
            strip((int)b, B.lo[0] - n, B.hi[1] + 1, B.hi[0] + n, B.hi[1] + n);
            strip((int)b, B.lo[0] - n, B.lo[1], B.lo[0] - 1, B.hi[1]);
            strip((int)b, B.hi[0] + 1, B.lo[1], B.hi[0] + n, B.hi[1]);
            avg.push_back(CellRect{(int)b, (int)b, B.lo[0] / r, B.lo[1] / r, (B.hi[0] - B.lo[0] + 1) / r,
                                   (B.hi[1] - B.lo[1] + 1) / r, 0});
            fb.push_back(int4{B.lo[0] / r, B.lo[1] / r, B.hi[0] / r, B.hi[1] / r});
        }
        g_upload(valid, g.d_valid, g.valid_cap, g.n_valid, g.valid_total);
        g_upload(ghost, g.d_ghost, g.ghost_cap, g.n_ghost, g.ghost_total);
        g_upload(avg, g.d_avg, g.avg_cap, g.n_avg, g.avg_total);
        if (fb.size() > g.fb_cap) {
            ZAMR_HIP(hipStreamSynchronize(stream()));
            if (g.d_fb) hipFree(g.d_fb);
            g.fb_cap = fb.size() * 2;
            ZAMR_HIP(hipMalloc(&g.d_fb, g.fb_cap * sizeof(int4)));
        }
        stage_->upload(g.d_fb, fb.data(), fb.size() * sizeof(int4), stream());
        ZAMR_HIP(hipMemsetAsync(g.d_owner, 0xff, (size_t)g.nx * g.ny * sizeof(int), stream()));
        if (g.n_valid)
            hipLaunchKernelGGL(paint_owner_kernel, dim3(grid1(g.valid_total)), dim3(256), 0, stream(), g.d_valid,
                               g.n_valid, g.valid_total, g.d_owner, g.nx);
        ZAMR_HIP(hipMemsetAsync(gl[l - 1].d_cov, 0, (size_t)gl[l - 1].nx * gl[l - 1].ny, stream()));
        if (!fb.empty())
            hipLaunchKernelGGL(cover_paint_kernel, dim3((unsigned)fb.size()), dim3(256), 0, stream(), g.d_fb,
                               gl[l - 1].d_cov, gl[l - 1].nx);
        ZAMR_HIP(hipGetLastError());
        g.max_stale = true;
    }

    // cell_cons_interp from level l - 1's current buffer into the listed
    // cells of level l's box set F (buffers in mask).
    void g_interp(int l, Sol& F, CellRect* d, int n, long long total, unsigned mask, const int* fine_owner) {
        if (!n || !total) return;
        const Sol& C = level(l - 1);
        if (l == 1) {
            hipLaunchKernelGGL((interp_r_kernel<M, CoarseGrid>), dim3(grid1(total)), dim3(256), 0, stream(),
                               F.d_boxes, d, n, total, grid(), ref_ratio, C.cur, mask, fine_owner, gl[l].nx);
        } else {
            const OwnerGrid og{C.d_boxes, gl[l - 1].d_owner, gl[l - 1].nx, gl[l - 1].ny};
            hipLaunchKernelGGL((interp_r_kernel<M, OwnerGrid>), dim3(grid1(total)), dim3(256), 0, stream(),
                               F.d_boxes, d, n, total, og, ref_ratio, C.cur, mask, fine_owner, gl[l].nx);
        }
        ZAMR_HIP(hipGetLastError());
    }

    // FillCoarsePatch of a level-l box set: valid and in-domain ghost cells.
    void g_coarse_patch(int l, Sol& F) {
        std::vector<CellRect> all;
        for (size_t b = 0; b < F.hb.size(); ++b) {
            const HostBox& B = F.hb[b];
            const int i0 = std::max(0, B.lo[0] - B.ng), i1 = std::min(F.nx - 1, B.hi[0] + B.ng);
            const int j0 = std::max(0, B.lo[1] - B.ng), j1 = std::min(F.ny - 1, B.hi[1] + B.ng);
            all.push_back(CellRect{(int)b, (int)b, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
        }
        int n = 0;
        long long total = 0;
        CellRect* d = upload_scratch(all, n, total);
        g_interp(l, F, d, n, total, 1u << F.cur, nullptr);
    }

    // The coarse-fine ghost cells of level l before one of its steps.
    void g_cf_ghosts(int l) {
        Sol& F = level(l);
        GLev& g = gl[l];
        g_interp(l, F, g.d_ghost, g.n_ghost, g.ghost_total, F.nbuf() == 3 ? 7u : 3u, g.d_owner);
    }

    // average_down of level l + 1 into level l.
    void g_average_down(int l) {
        Sol& F = level(l + 1);
        Sol& C = level(l);
        const GLev& g = gl[l + 1];
        if (!g.n_avg || !g.avg_total) return;
        if (l == 0) {
            hipLaunchKernelGGL((average_down_r_kernel<M, CoarseGrid>), dim3(grid1(g.avg_total)), dim3(256), 0,
                               stream(), F.d_boxes, g.d_avg, g.n_avg, g.avg_total, grid(), ref_ratio, F.cur, C.cur);
        } else {
            const OwnerGrid og{C.d_boxes, gl[l].d_owner, gl[l].nx, gl[l].ny};
            hipLaunchKernelGGL((average_down_r_kernel<M, OwnerGrid>), dim3(grid1(g.avg_total)), dim3(256), 0,
                               stream(), F.d_boxes, g.d_avg, g.n_avg, g.avg_total, og, ref_ratio, F.cur, C.cur);
        }
        ZAMR_HIP(hipGetLastError());
        gl[l].max_stale = true;
    }

    // ErrorEst of level l, buffered and coarsened, on the host.
    std::vector<unsigned char> g_tags(int l) {
        GLev& g = gl[l];
        if (l == 0) {
            hipLaunchKernelGGL((error_est_kernel<M>), dim3((unsigned)((L0.nx + 255) / 256), (unsigned)L0.ny),
                               dim3(256), 0, stream(), grid(), L0.cur, l0_stage_ran ? L0.last_in : -1, L0.prm,
                               L0.clamp_flag(), g.t_last, tag_threshold, tag_b_max, g.d_tag);
        } else {
            Sol& F = level(l);
            ZAMR_HIP(hipMemsetAsync(g.d_tag, 0, (size_t)g.nx * g.ny, stream()));
            if (g.n_valid)
                hipLaunchKernelGGL((error_est_boxes_kernel<M>), dim3(grid1(g.valid_total)), dim3(256), 0, stream(),
                                   F.d_boxes, g.d_valid, g.n_valid, g.valid_total, F.cur, F.last_in, F.prm,
                                   F.clamp_flag(), g.t_last, tag_threshold, tag_b_max, g.d_tag, g.nx);
        }
        ZAMR_HIP(hipGetLastError());
        // the tag map itself stands in for the comparison map (only read)
        hipLaunchKernelGGL(tag_buffer_kernel, dim3(grid1((long long)g.cnx * g.cny)), dim3(256), 0, stream(), g.d_tag,
                           g.nx, g.ny, n_error_buf, bf_lev, g.d_ctag, g.cnx, g.cny, g.d_tag, d_changed);
        ZAMR_HIP(hipGetLastError());
        std::vector<unsigned char> ct((size_t)g.cnx * g.cny);
        ZAMR_HIP(hipMemcpyAsync(ct.data(), g.d_ctag, ct.size(), hipMemcpyDeviceToHost, stream()));
        ZAMR_HIP(hipStreamSynchronize(stream()));
        return ct;
    }

    // Proper-nesting domains (1 = allowed) of levels lbase .. max_crse in
    // cluster units: at lbase the coarsened grids shrunk by n_proper inside
    // the domain; above, the complement refined and grown by n_proper again.
    std::vector<std::vector<unsigned char>> g_nesting(int lbase, int max_crse) {
        const int bf = bf_lev, np = n_proper;
        std::vector<std::vector<unsigned char>> comp(max_crse + 1), pn(max_crse + 1);
        for (int l = lbase; l <= max_crse; ++l) {
            const int cnx_l = gl[l].cnx, cny_l = gl[l].cny;
            std::vector<unsigned char> base((size_t)cnx_l * cny_l, 0);
            if (l == lbase) {
                std::vector<unsigned char> cov(base.size(), 0);
                for (const HostBox& b : level(l).hb)
                    for (int J = b.lo[1] / bf; J <= b.hi[1] / bf; ++J)
                        for (int I = b.lo[0] / bf; I <= b.hi[0] / bf; ++I) cov[(size_t)J * cnx_l + I] = 1;
                for (size_t k = 0; k < base.size(); ++k) base[k] = cov[k] ? 0 : 1;
            } else {
                const int pcnx = gl[l - 1].cnx, rr = ref_ratio;
                for (int J = 0; J < cny_l; ++J)
                    for (int I = 0; I < cnx_l; ++I)
                        base[(size_t)J * cnx_l + I] = comp[l - 1][(size_t)(J / rr) * pcnx + I / rr];
            }
            comp[l].assign(base.size(), 0);
            for (int J = 0; J < cny_l; ++J)
                for (int I = 0; I < cnx_l; ++I) {
                    if (!base[(size_t)J * cnx_l + I]) continue;
                    for (int jj = std::max(0, J - np); jj <= std::min(cny_l - 1, J + np); ++jj)
                        for (int ii = std::max(0, I - np); ii <= std::min(cnx_l - 1, I + np); ++ii)
                            comp[l][(size_t)jj * cnx_l + ii] = 1;
                }
            pn[l].resize(base.size());
            for (size_t k = 0; k < pn[l].size(); ++k) pn[l][k] = comp[l][k] ? 0 : 1;
        }
        return pn;
    }

    // A cluster box cut by a mask into rectangles: each row's runs of set
    // cells, consecutive rows' runs that start and end alike merged.
    static std::vector<std::array<int, 4>> mask_rects(const std::vector<unsigned char>& mask, int w,
                                                     const std::array<int, 4>& c) {
        std::vector<std::array<int, 4>> done, open;
        for (int j = c[1]; j <= c[3] + 1; ++j) {
            std::vector<std::pair<int, int>> runs;
            if (j <= c[3])
                for (int i = c[0]; i <= c[2];) {
                    if (!mask[(size_t)j * w + i]) {
                        ++i;
                        continue;
                    }
                    int e = i;
                    while (e + 1 <= c[2] && mask[(size_t)j * w + e + 1]) ++e;
                    runs.push_back({i, e});
                    i = e + 1;
                }
            std::vector<std::array<int, 4>> next;
            for (const auto& o : open) {
                bool cont = false;
                for (auto& r : runs) cont = cont || (r.first == o[0] && r.second == o[2]);
                if (cont) next.push_back({o[0], o[1], o[2], j});
                else done.push_back(o);
            }
            for (auto& r : runs) {
                bool cont = false;
                for (const auto& o : open) cont = cont || (r.first == o[0] && r.second == o[2]);
                if (!cont) next.push_back({r.first, j, r.second, j});
            }
            open.swap(next);
        }
        std::sort(done.begin(), done.end(), [](const std::array<int, 4>& a, const std::array<int, 4>& b) {
            return a[1] != b[1] ? a[1] < b[1] : a[0] < b[0];
        });
        return done;
    }

    // Boxes of level l + 1 from a coarsened tag map of level l: clusters,
    // cut by the nesting mask (empty: none), cut to amr.max_grid_size (x: whole
    // stage strips), refined.
    std::vector<std::array<int, 4>> g_grids(int l, const std::vector<unsigned char>& ct,
                                            const std::vector<unsigned char>& pn) {
        GLev& g = gl[l];
        stage_->reset();
        stage_->upload(g.d_ctag, ct.data(), ct.size(), stream());
        ZAMR_HIP(hipStreamSynchronize(stream()));  // the clustering runs on its own stream
        auto cl = gpu_cluster(g.d_ctag, g.cnx, g.cny);
        std::sort(cl.begin(), cl.end(), [](const std::array<int, 4>& a, const std::array<int, 4>& b) {
            return a[1] != b[1] ? a[1] < b[1] : a[0] < b[0];
        });
        std::vector<std::array<int, 4>> parts;
        for (auto& c : cl) {
            if (pn.empty()) {
                parts.push_back(c);
                continue;
            }
            for (auto& q : mask_rects(pn, g.cnx, c)) parts.push_back(q);
        }
        const int r = ref_ratio * bf_lev;
        const int m = std::max(1, fine_max_grid_size / r);
        const int strip = WV - 2 * L0.hx_of(L0.spatial_order);
        const int wx = (fine_max_grid_size / strip) * strip;
        const int mx = (wx >= r && wx % r == 0) ? wx / r : m;
        std::vector<std::array<int, 4>> fine;
        for (auto& c : parts)
            for (int y0 = c[1]; y0 <= c[3]; y0 += m)
                for (int x0 = c[0]; x0 <= c[2]; x0 += mx) {
                    const int x1 = std::min(x0 + mx - 1, c[2]), y1 = std::min(y0 + m - 1, c[3]);
                    fine.push_back({x0 * r, y0 * r, (x1 + 1) * r - 1, (y1 + 1) * r - 1});
                }
        return fine;
    }

    std::vector<std::array<int, 4>> boxes_of(int l) {
        std::vector<std::array<int, 4>> v;
        for (const HostBox& b : level(l).hb) v.push_back({b.lo[0], b.lo[1], b.hi[0], b.hi[1]});
        return v;
    }

    double level_time(int l) {
        Sol& S = level(l);
        S.download_groups();
        return S.h_groups[0].time;
    }

    // A new box set of level l, filled from level l - 1 (valid and in-domain
    // ghost cells) and, where `old` covers, from the old level (RemakeLevel's
    // FillPatch), with fresh Qaux.
    std::unique_ptr<Sol> g_make(int l, const std::vector<std::array<int, 4>>& boxes, Sol* old, double time) {
        auto F = std::make_unique<Sol>();
        F->hist_cap = 16;
        F->cache = &fine_cache;
        F->stage = stage_.get();
        F->pool_headroom = 1.25;
        F->configure_fine(level(l - 1), ref_ratio, boxes);
        F->use_graph = L0.use_graph;
        g_coarse_patch(l, *F);
        if (old) {
            std::vector<CellRect> cp;
            for (size_t b = 0; b < F->hb.size(); ++b) {
                const HostBox& B = F->hb[b];
                for (size_t o = 0; o < old->hb.size(); ++o) {
                    const HostBox& O = old->hb[o];
                    const int i0 = std::max(B.lo[0] - B.ng, O.lo[0]), i1 = std::min(B.hi[0] + B.ng, O.hi[0]);
                    const int j0 = std::max(B.lo[1] - B.ng, O.lo[1]), j1 = std::min(B.hi[1] + B.ng, O.hi[1]);
                    if (i0 > i1 || j0 > j1) continue;
                    cp.push_back(CellRect{(int)b, (int)o, i0, j0, i1 - i0 + 1, j1 - j0 + 1, 0});
                }
            }
            int n = 0;
            long long total = 0;
            CellRect* d = upload_scratch(cp, n, total);
            if (n)
                hipLaunchKernelGGL((copy_levels_kernel<M>), dim3(grid1(total)), dim3(256), 0, stream(), F->d_boxes,
                                   F->cur, old->d_boxes, old->cur, d, n, total);
            ZAMR_HIP(hipGetLastError());
            ++remakes;
        }
        fresh_aux(*F, time);
        return F;
    }

    void g_set_level(int l, std::unique_ptr<Sol> F) {
        if (l == 1) {
            L1 = std::move(F);
        } else {
            if ((int)up.size() < l - 2) throw Error(4, "internal: a level made above a missing one");
            if ((int)up.size() == l - 2) up.push_back(std::move(F));
            else up[l - 2] = std::move(F);
        }
        g_build_lists(l);
    }

    void g_clear_above(int l) {
        while (finest_level() > l) {
            const int f = finest_level();
            if (f >= 2) up.pop_back();
            else L1.reset();
            ZAMR_HIP(hipMemsetAsync(gl[f - 1].d_cov, 0, (size_t)gl[f - 1].nx * gl[f - 1].ny, stream()));
        }
    }

    // AmrMesh::MakeNewGrids(lbase) + AmrCore::regrid(lbase, time).
    void g_regrid(int lbase, double time) {
        if (lbase >= max_level) return;
        Clock clk{regrid_ms};
        ++regrids;
        const int max_crse = std::min(finest_level(), max_level - 1);
        const auto pn = g_nesting(lbase, max_crse);
        const int bf = bf_lev, r = ref_ratio, np = n_proper;
        std::vector<std::vector<std::array<int, 4>>> ng(max_crse + 2);
        int new_finest = lbase;
        for (int levc = max_crse; levc >= lbase; --levc) {
            const int levf = levc + 1;
            const GLev& g = gl[levc];
            auto ct = g_tags(levc);
            if (levf < new_finest)
                for (auto& b : ng[levf + 1]) {
                    int lo[2], hi[2];
                    for (int d = 0; d < 2; ++d) {
                        lo[d] = std::max(b[d] / r - np, 0) / r / bf;
                        hi[d] = (b[d + 2] / r + np) / r / bf;
                    }
                    for (int J = lo[1]; J <= std::min(hi[1], g.cny - 1); ++J)
                        for (int I = lo[0]; I <= std::min(hi[0], g.cnx - 1); ++I) ct[(size_t)J * g.cnx + I] = 1;
                }
            bool full = true;
            for (unsigned char v : pn[levc]) full = full && v;
            for (size_t k = 0; k < ct.size(); ++k)
                if (!pn[levc][k]) ct[k] = 0;
            bool any = false;
            for (unsigned char v : ct) any = any || v;
            ng[levf] = any ? g_grids(levc, ct, full ? std::vector<unsigned char>{} : pn[levc])
                           : std::vector<std::array<int, 4>>{};
            if (!ng[levf].empty()) new_finest = std::max(new_finest, levf);
        }
        bool coarse_changed = false;
        for (int lev = lbase + 1; lev <= new_finest; ++lev) {
            if (lev <= finest_level()) {
                const bool changed = ng[lev] != boxes_of(lev);
                if (changed || coarse_changed) {
                    auto& slot = level_ptr(lev);
                    auto F = g_make(lev, changed ? ng[lev] : boxes_of(lev), slot.get(), time);
                    g_set_level(lev, std::move(F));
                }
                coarse_changed = changed;
            } else {
                g_set_level(lev, g_make(lev, ng[lev], nullptr, time));
            }
        }
        g_clear_above(new_finest);
    }

    // The level loop of InitData on the current level-0 state.
    void g_init_levels() {
        L0.ensure_aux();
        l0_stage_ran = false;
        for (int it = 0; it < 4; ++it) {
            for (int l = 1; l <= finest_level(); ++l) g_coarse_patch(l, level(l));
            const int old_finest = finest_level();
            g_regrid(0, 0.0);
            if (old_finest == finest_level() && it > 0) break;
        }
        for (int l = 1; l <= finest_level(); ++l) {
            g_coarse_patch(l, level(l));
            fresh_aux(level(l), 0.0);
        }
        ZAMR_HIP(hipStreamSynchronize(stream()));
        for (int l = 1; l <= finest_level(); ++l) level(l).reset_clock();
        for (GLev& g : gl) g.max_stale = true;
    }

    // ZoomyAmr::TimeStep(lev, time, iteration), the time on the level's clock.
    // `lines`: print Evolve's MOOD line after each level's Advance.
    void g_time_step(int lev, int iteration, bool lines) {
        if (lev < max_level && iteration > 0) g_regrid(lev, level_time(lev));
        Sol& S = level(lev);
        GLev& g = gl[lev];
        hipStream_t st = stream();
        if (g.max_stale || !S.maxinv_fresh) {
            if (lev == 0) S.ensure_aux();
            hipLaunchKernelGGL(clear_max_kernel, dim3(1), dim3(64), 0, st, S.d_groups, S.ngroups);
            auto& it = S.items(S.hx_of(S.spatial_order));
            hipLaunchKernelGGL((max_inv_kernel<M, N>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, st, S.d_boxes, it.d,
                               it.n, S.cur, S.prm, S.d_groups, WV - 2 * it.hx);
            ZAMR_HIP(hipGetLastError());
        }
        if (lev > 0) g_cf_ghosts(lev);
        g.t_last = level_time(lev);
        S.begin_step();
        S.advance_kernels();
        const bool finer = lev < finest_level();
        if (finer) {
            hipLaunchKernelGGL(sync_clock_kernel, dim3(1), dim3(1), 0, st, S.d_groups, level(lev + 1).d_groups);
            // Qaux = aux(Q^{n+1}) as Advance's last UpdateState leaves it; the
            // next ComputeDt reads it beside the averaged-down Q
            auto& it = S.items(S.hx_of(S.spatial_order));
            hipLaunchKernelGGL((materialize_aux_kernel<M>), dim3(Sol::aux_grid(it.n)), dim3(NT), 0, st, S.d_boxes,
                               it.d, it.n, S.cur, S.prm, 0.0, WV - 2 * it.hx, (const GroupState*)S.d_groups);
            ZAMR_HIP(hipGetLastError());
        }
        S.end_step();
        S.aux_stale = !finer;
        S.maxinv_fresh = true;
        g.max_stale = false;
        if (lev == 0) {
            l0_stage_ran = true;
            t_last_stage = g.t_last;
        }
        if (lines && S.mood()) {
            S.download_groups();
            if (S.h_groups[0].troubled > 0)
                std::cout << "  [MOOD] lev " << lev << " troubled cells: " << (long)S.h_groups[0].troubled
                          << " -> whole-step order-1 redo\n";
        }
        if (finer) {
            for (int i = 1; i <= ref_ratio; ++i) g_time_step(lev + 1, i, lines);
            g_average_down(lev);
        }
    }

    // ZoomyAmr::TimeStep(0, time, iteration) at a caller-given time; returns dt.
    double time_step(double time, int iteration) {
        L0.download_groups();
        if (L0.h_groups[0].time != time) {
            L0.h_groups[0].time = time;
            L0.upload_groups();
        }
        coarse_step(iteration, [](int) {});
        L0.download_groups();
        return L0.h_groups[0].dt;
    }

    // ErrorEst of level 0 as it stands (ghosts as the last stage left them,
    // or 0 before any), one byte per cell in raster order.
    void error_est_raw(int mode_last, double time_last, unsigned char* host) {
        const long long n = (long long)L0.nx * L0.ny;
        hipLaunchKernelGGL((error_est_kernel<M>), dim3((unsigned)((L0.nx + 255) / 256), (unsigned)L0.ny), dim3(256), 0,
                           stream(), grid(), L0.cur, mode_last,
                           L0.prm, L0.clamp_flag(), time_last, tag_threshold, tag_b_max, d_tags);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipMemcpyAsync(host, d_tags, (size_t)n, hipMemcpyDeviceToHost, stream()));
        ZAMR_HIP(hipStreamSynchronize(stream()));
    }

    bool running_host() const {  // on level 0's group state as last downloaded
        const GroupState& g = L0.h_groups[0];
        return g.time < L0.time_end && (L0.max_step < 0 || g.step < L0.max_step);
    }

    void run_steps(int nsteps, int* done, double* time) {
        L0.download_groups();
        const int step0 = L0.h_groups[0].step;
        for (int k = 0; k < nsteps; ++k) {
            if (k > 0) L0.download_groups();
            if (!running_host()) break;
            coarse_step(L0.h_groups[0].step, [](int) {}, true);
        }
        L0.download_groups();
        if (done) *done = L0.h_groups[0].step - step0;
        if (time) *time = L0.h_groups[0].time;
    }

    double total_mass() {
        if (general) {
            double m = 0.0;
            for (int l = 0; l <= finest_level(); ++l) m += level(l).total_mass(l < finest_level() ? gl[l].d_cov : nullptr);
            return m;
        }
        double m = L0.total_mass(L1 ? d_cover : nullptr);
        if (L1) m += L1->total_mass();
        return m;
    }

    // Dense (ncomp, ny_lev, nx_lev) copy of a level's valid cells; cells no
    // box covers are left as the caller filled them.
    void copy_level(int lev, int field, double* host, size_t n, bool to_device) {
        if (lev == 0) {
            L0.copy_level(0, field, host, n, to_device);
            if (to_device) {
                l0_max_stale = true;
                L0.maxinv_fresh = false;
                if (general) gl[0].max_stale = true;
            }
            return;
        }
        Sol& F = level(lev);
        const int nc = field == 0 ? nq : na;
        if (field != 0 && field != 1) throw Error(1, "field must be 0 (Q) or 1 (Qaux)");
        if (n != (size_t)nc * F.nx * F.ny) throw Error(1, "array size mismatch: expected ncomp*ny*nx doubles");
        if (field == 1 || to_device) F.ensure_aux();
        for (const HostBox& B : F.hb) {
            const size_t cells = (size_t)B.gx * B.gy;
            std::vector<double> tmp(cells * nc);
            double* dptr = F.d_pool + (field == 0 ? B.off_q[F.cur] : B.off_aux);
            ZAMR_HIP(hipMemcpyAsync(tmp.data(), dptr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
            ZAMR_HIP(hipStreamSynchronize(stream()));
            for (int c = 0; c < nc; ++c)
                for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                    for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                        const size_t li = (size_t)(i - B.lo[0] + B.ng) + (size_t)(j - B.lo[1] + B.ng) * B.gx +
                                          (size_t)c * cells;
                        const size_t gi = ((size_t)c * F.ny + j) * F.nx + i;
                        if (to_device) tmp[li] = host[gi];
                        else host[gi] = tmp[li];
                    }
            if (to_device) {
                ZAMR_HIP(hipMemcpyAsync(dptr, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                                        stream()));
                ZAMR_HIP(hipStreamSynchronize(stream()));
            }
        }
        if (to_device) {
            l1_max_stale = true;
            F.maxinv_fresh = false;
            if (general) gl[lev].max_stale = true;
        }
    }

    void reset_clock() {
        L0.reset_clock();
        if (L1) L1->reset_clock();
        for (auto& u : up) u->reset_clock();
        plot_step = 0;
        next_plot_time = 0.0;
    }

    void write_plotfile(const std::string& dir, int step, double time);

    void evolve() {  // ZoomyAmr::Evolve over the hierarchy
        reset_clock();
        double time = 0.0;
        int step = 0;
        next_plot_time = 0.0;
        const auto tstart = std::chrono::steady_clock::now();
        const double mass0 = total_mass();
        std::cout << "Initial total mass = " << std::setprecision(12) << mass0 << "\n";
        while (time < L0.time_end && (L0.max_step < 0 || step < L0.max_step)) {
            if (time >= next_plot_time) {
                write_plotfile("", plot_step, time);
                next_plot_time += L0.plot_dt_interval;
                plot_step++;
            }
            bool l0_printed = false;
            auto mood_lines = [&](int) {
                if (!l0_printed) {  // level 0's Advance came first
                    l0_printed = true;
                    L0.download_groups();
                    if (L0.mood() && L0.h_groups[0].troubled > 0)
                        std::cout << "  [MOOD] lev 0 troubled cells: " << (long)L0.h_groups[0].troubled
                                  << " -> whole-step order-1 redo\n";
                }
                L1->download_groups();
                if (L1->mood() && L1->h_groups[0].troubled > 0)
                    std::cout << "  [MOOD] lev 1 troubled cells: " << (long)L1->h_groups[0].troubled
                              << " -> whole-step order-1 redo\n";
            };
            if (general) {
                g_time_step(0, step, true);
                l0_printed = true;
            } else {
                coarse_step(step, mood_lines);
            }
            L0.download_groups();
            if (!l0_printed && L0.mood() && L0.h_groups[0].troubled > 0)
                std::cout << "  [MOOD] lev 0 troubled cells: " << (long)L0.h_groups[0].troubled
                          << " -> whole-step order-1 redo\n";
            const double dt0 = L0.h_groups[0].dt;
            time = L0.h_groups[0].time;
            step++;
            const double mass = total_mass();
            const double drift = (mass0 != 0.0) ? (mass - mass0) / mass0 : (mass - mass0);
            std::cout << "Step " << step << " dt: " << std::setprecision(12) << dt0 << " time: "
                      << std::setprecision(12) << time << "s"
                      << " levels: " << finest_level() + 1 << " mass: " << std::setprecision(14) << mass
                      << " drift: " << std::setprecision(4) << drift << "\n";
        }
        write_plotfile("", plot_step, time);
        const double evo = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
        std::cout << "Total evolution time = " << evo << " seconds\n";
        std::cout.flush();
    }
};

}  // namespace zamr
