#pragma once
// CPU ORACLE — test infrastructure only, never part of the product path.
//
// The two-level AMR advance of Source/ZoomyAmr.cpp restated on the oracle's
// Solver levels: InitData's level loop (:137-155), ErrorEst (:225-253),
// regrid with RemakeLevel / MakeNewLevelFromCoarse / ClearLevel (:161-219),
// FillPatch / FillCoarsePatch (:258-288), the subcycled TimeStep with
// average_down (:733-762), ComputeTotalMass over the composite (:517-536).
//
// The AMR machinery itself lives in AMReX, an external dependency the
// reference does not vendor (Exec/GNUmakefile: AMREX_HOME; no version pin,
// the container's /opt/amrex).  Its published algorithms are restated here:
//   * cell_cons_interp = CellConservativeLinear with linear limiting
//     (AMReX_MFInterp_C.H, mf_cell_cons_lin_interp_llslope + _lin_interp):
//     central slopes, one limiting factor per direction shared by all
//     components, fine = crse + xoff * sx + yoff * sy, xoff = -+1/4;
//   * average_down (AMReX_MultiFabUtil_2D_C.H, amrex_avgdown): the mean of
//     the ratio^2 children, summed j-outer i-inner, times 1/ratio^2;
//   * MakeNewGrids (AMReX_AmrMesh.cpp) for level 0 -> 1: tags buffered by
//     n_error_buf, coarsened by blocking_factor / ref_ratio, clustered by
//     Berger-Rigoutsos (AMReX_Cluster.cpp: ClusterList::chop, Cluster::chop,
//     FindCut) to efficiency grid_eff, refined back.
//
// Conventions where the reference reads memory it never wrote (each is
// undefined there; this restatement and the device code share the choice):
//   * coarse cells outside the domain in the interpolation stencil take the
//     nearest interior value (first-order extrapolation, the BCRec foextrap
//     ZoomyAmr.cpp:88-94 declares; the reference's fill functor is NullFill);
//   * ErrorEst before any FillPatch (during InitData) reads ghost cells as 0;
//   * the Qaux of a new or remade fine level is update_aux_variables(Q, 0)
//     on its valid cells (RemakeLevel / MakeNewLevelFromCoarse leave it unset
//     and ComputeDt reads it next).
// Only the union of the fine boxes enters the results, so the fine level's
// decomposition into boxes (AMReX's simplify / maxSize order) is free.
//
// Pinned by: interpolation conservation (children average to the parent
// exactly in exact arithmetic, to round-off here) and exactness on linear
// data, average_down(interp) == identity on linear data, clustering
// invariants (every tag covered, disjoint boxes, efficiency >= grid_eff or
// unsplittable), decomposition invariance (fine max_grid_size), and the
// no-tag run equal to the single-level run bit for bit.
#include <array>
#include <memory>

#include "zamr_oracle.hpp"

namespace zorc {

struct AmrParams {
    int max_level = 1;           // amr.max_level (this restatement: 0 or 1)
    int ref_ratio = 2;           // amr.ref_ratio
    int blocking_factor = 2;     // amr.blocking_factor (build.py:92-95 writes 2 with AMR)
    int n_error_buf = 1;         // amr.n_error_buf (AMReX default)
    Real grid_eff = 0.7;         // amr.grid_eff (AMReX default)
    int fine_max_grid_size = 64; // decomposition of the fine level (results do not depend on it)
    Real tag_threshold = 0.01;   // tagging.threshold (ZoomyAmr.H:147)
    Real tag_b_max = 1.0e30;     // tagging.b_max (ZoomyAmr.H:148)
};

// ---- Berger-Rigoutsos clustering (AMReX_Cluster.cpp) ----
using IV = std::array<int, 2>;

inline Box bounding_box(const std::vector<IV>& pts) {
    Box b{{pts[0][0], pts[0][1]}, {pts[0][0], pts[0][1]}};
    for (const IV& q : pts)
        for (int d = 0; d < 2; ++d) {
            b.lo[d] = std::min(b.lo[d], q[d]);
            b.hi[d] = std::max(b.hi[d], q[d]);
        }
    return b;
}

enum CutStatus { hole_cut = 0, gradient_cut = 1, bisect_cut = 2, invalid_cut = 3 };

// FindCut: the hole nearest the middle, else the largest sign change of the
// second difference (at least MINOFF from either end, above CUT_THRESH),
// nearest the middle among equals, else the middle.
inline int find_cut(const std::vector<int>& hist, int lo, CutStatus& status) {
    const int MINOFF = 2, CUT_THRESH = 2;
    const int len = (int)hist.size();
    status = invalid_cut;
    if (len <= 1) return lo;
    const int mid = len / 2;
    int cutpoint = -1;
    for (int i = 0; i < len; ++i) {
        if (hist[i] == 0) {
            status = hole_cut;
            if (std::abs(cutpoint - mid) > std::abs(i - mid)) {
                cutpoint = i;
                if (i > mid) break;
            }
        }
    }
    if (status == hole_cut) return lo + cutpoint;
    std::vector<int> d2(len, 0);
    for (int i = 1; i < len - 1; ++i) d2[i] = hist[i + 1] - 2 * hist[i] + hist[i - 1];
    int locmax = -1;
    for (int i = MINOFF; i < len - MINOFF; ++i) {
        const int a = d2[i - 1], b = d2[i];
        const int dif = std::abs(a - b);
        if (a * b < 0 && dif >= locmax) {
            if (dif > locmax) {
                status = gradient_cut;
                cutpoint = i;
                locmax = dif;
            } else if (std::abs(i - mid) < std::abs(cutpoint - mid)) {
                cutpoint = i;
            }
        }
    }
    if (locmax <= CUT_THRESH) {
        cutpoint = mid;
        status = bisect_cut;
    }
    return lo + cutpoint;
}

// Cluster::chop: cut the point set in the direction of the best cut status
// (ties: the larger shorter side, later direction on equality); the points
// below the cut stay, the rest are returned.
inline std::vector<IV> chop_cluster(std::vector<IV>& pts) {
    const Box b = bounding_box(pts);
    std::vector<int> hist[2];
    for (int d = 0; d < 2; ++d) hist[d].assign(b.len(d), 0);
    for (const IV& q : pts)
        for (int d = 0; d < 2; ++d) hist[d][q[d] - b.lo[d]]++;
    CutStatus st[2];
    int cut[2];
    CutStatus best = invalid_cut;
    for (int d = 0; d < 2; ++d) {
        cut[d] = find_cut(hist[d], b.lo[d], st[d]);
        if (st[d] < best) best = st[d];
    }
    int dir = -1, minlen = -1;
    for (int d = 0; d < 2; ++d) {
        if (st[d] != best) continue;
        const int l = std::min(cut[d] - b.lo[d], b.hi[d] - cut[d]);
        if (l >= minlen) {
            dir = d;
            minlen = l;
        }
    }
    std::vector<IV> lower, upper;
    for (const IV& q : pts) (q[dir] < cut[dir] ? lower : upper).push_back(q);
    pts.swap(lower);
    return upper;
}

inline double cluster_eff(const std::vector<IV>& pts) {
    const Box b = bounding_box(pts);
    return (double)pts.size() / ((double)b.len(0) * (double)b.len(1));
}

// ClusterList::chop: split every cluster below the efficiency until none is.
inline std::vector<Box> cluster_tags(const std::vector<IV>& tags, Real grid_eff) {
    std::vector<Box> out;
    if (tags.empty()) return out;
    std::vector<std::vector<IV>> lst{tags};
    for (size_t k = 0; k < lst.size();) {
        if (cluster_eff(lst[k]) < grid_eff) {
            std::vector<IV> hi = chop_cluster(lst[k]);
            lst.push_back(std::move(hi));
        } else {
            ++k;
        }
    }
    for (auto& c : lst) out.push_back(bounding_box(c));
    return out;
}

template <class Model, class Numerics>
struct Amr {
    using Lev = Solver<Model, Numerics>;
    static constexpr int nq = Lev::nq;
    static constexpr int na = Lev::na;
    using VQ = typename Lev::VQ;
    using VA = typename Lev::VA;

    Settings s0;
    std::vector<Real> prm;
    AmrParams ap;
    std::unique_ptr<Lev> L0, L1;
    Real dt_level[2] = {1.0e100, 1.0e100};
    int nthreads = 1;
    long regrids = 0, remakes = 0;

    Amr(const Settings& s, const Real* params, const AmrParams& a) : s0(s), prm(params, params + Lev::np), ap(a) {
        L0 = std::make_unique<Lev>(s0, prm.data());
    }

    int finest_level() const { return L1 ? 1 : 0; }
    Lev& level(int l) { return l == 0 ? *L0 : *L1; }

    // ---------- coarse data access ----------
    Real coarse_at(int ic, int jc, int n) const {
        ic = std::min(std::max(ic, 0), s0.nx - 1);  // foextrap beyond the domain
        jc = std::min(std::max(jc, 0), s0.ny - 1);
        for (auto& f : L0->Q.fabs)
            if (ic >= f.vb.lo[0] && ic <= f.vb.hi[0] && jc >= f.vb.lo[1] && jc <= f.vb.hi[1]) return f(ic, jc, n);
        return 0.0;  // unreachable: level 0 covers the domain
    }
    Fab& coarse_fab(int ic, int jc) {
        for (auto& f : L0->Q.fabs)
            if (ic >= f.vb.lo[0] && ic <= f.vb.hi[0] && jc >= f.vb.lo[1] && jc <= f.vb.hi[1]) return f;
        return L0->Q.fabs[0];
    }

    // cell_cons_interp at fine cell (i, j) of level 1.
    void interp(int i, int j, Real* out) const {
        const int r = ap.ref_ratio;
        const int ic = i / r, jc = j / r;  // i, j >= 0 inside the domain
        Real sfx = 1.0, sfy = 1.0;
        Real slx[nq], sly[nq], u0[nq];
        for (int n = 0; n < nq; ++n) {
            const Real u = coarse_at(ic, jc, n);
            u0[n] = u;
            for (int d = 0; d < 2; ++d) {
                const Real um = d == 0 ? coarse_at(ic - 1, jc, n) : coarse_at(ic, jc - 1, n);
                const Real up = d == 0 ? coarse_at(ic + 1, jc, n) : coarse_at(ic, jc + 1, n);
                const Real dc = 0.5 * (up - um);
                const Real df = 2.0 * (up - u);
                const Real db = 2.0 * (u - um);
                Real sl = (df * db >= 0.0) ? std::min(std::abs(df), std::abs(db)) : 0.0;
                sl = std::copysign(1.0, dc) * std::min(sl, std::abs(dc));
                Real& sf = d == 0 ? sfx : sfy;
                if (dc != 0.0) sf = std::min(sf, sl / dc);
                (d == 0 ? slx : sly)[n] = dc;
            }
        }
        const Real xoff = ((i - ic * r) + 0.5) / r - 0.5;
        const Real yoff = ((j - jc * r) + 0.5) / r - 0.5;
        for (int n = 0; n < nq; ++n) out[n] = u0[n] + xoff * (slx[n] * sfx) + yoff * (sly[n] * sfy);
    }

    // FillCoarsePatch (InterpFromCoarseLevel): valid and in-domain ghost cells.
    void fill_coarse_patch(Lev& F) {
        Real v[nq];
        for (auto& f : F.Q.fabs)
            for (int j = f.gl[1]; j <= f.gh[1]; ++j)
                for (int i = f.gl[0]; i <= f.gh[0]; ++i) {
                    if (i < 0 || i >= F.s.nx || j < 0 || j >= F.s.ny) continue;
                    interp(i, j, v);
                    for (int n = 0; n < nq; ++n) f(i, j, n) = v[n];
                }
    }

    // Qaux of a new or remade level: update_aux_variables(Q, 0) on valid cells.
    void fresh_aux(Lev& F, Real time) {
        for (size_t b = 0; b < F.boxes.size(); ++b) {
            Fab& fq = F.Q.fabs[b];
            Fab& fa = F.Qaux.fabs[b];
            for (int j = fq.vb.lo[1]; j <= fq.vb.hi[1]; ++j)
                for (int i = fq.vb.lo[0]; i <= fq.vb.hi[0]; ++i) {
                    VQ q;
                    VA a{};
                    for (int n = 0; n < nq; ++n) q(n, 0) = fq(i, j, n);
                    for (int n = 0; n < na; ++n) a(n, 0) = 0.0;
                    a = Model::update_aux_variables(q, a, F.p, time, F.cell_center(i, j));
                    for (int n = 0; n < na; ++n) fa(i, j, n) = a(n, 0);
                }
        }
    }

    // average_down (TimeStep, ZoomyAmr.cpp:759-760).
    void average_down() {
        const int r = ap.ref_ratio;
        const Real vol = 1.0 / (Real)(r * r);
        for (auto& f : L1->Q.fabs)
            for (int jc = f.vb.lo[1] / r; jc <= f.vb.hi[1] / r; ++jc)
                for (int ic = f.vb.lo[0] / r; ic <= f.vb.hi[0] / r; ++ic) {
                    Fab& c = coarse_fab(ic, jc);
                    for (int n = 0; n < nq; ++n) {
                        Real sum = 0.0;
                        for (int jj = 0; jj < r; ++jj)
                            for (int ii = 0; ii < r; ++ii) sum += f(ic * r + ii, jc * r + jj, n);
                        c(ic, jc, n) = vol * sum;
                    }
                }
    }

    // ErrorEst (ZoomyAmr.cpp:225-253) on level 0, ghosts as they stand.
    std::vector<unsigned char> error_est() const {
        std::vector<unsigned char> tag((size_t)s0.nx * s0.ny, 0);
        for (auto& f : L0->Q.fabs)
            for (int j = f.vb.lo[1]; j <= f.vb.hi[1]; ++j)
                for (int i = f.vb.lo[0]; i <= f.vb.hi[0]; ++i) {
                    if (f(i, j, idx_b) >= ap.tag_b_max) continue;
                    Real grad = 0.0;
                    const Real dhdx = std::abs(f(i + 1, j, idx_h) - f(i - 1, j, idx_h));
                    grad += dhdx * dhdx;
                    if (Lev::dim >= 2) {
                        const Real dhdy = std::abs(f(i, j + 1, idx_h) - f(i, j - 1, idx_h));
                        grad += dhdy * dhdy;
                    }
                    if (std::sqrt(grad) > ap.tag_threshold || f(i, j, idx_h) > 1.0e-2)
                        tag[(size_t)j * s0.nx + i] = 1;
                }
        return tag;
    }

    // MakeNewGrids for level 1: buffer, coarsen, cluster, refine, chop.
    std::vector<Box> make_grids(const std::vector<unsigned char>& tag) const {
        const int nx = s0.nx, ny = s0.ny, nb = ap.n_error_buf;
        std::vector<unsigned char> buf((size_t)nx * ny, 0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (!tag[(size_t)j * nx + i]) continue;
                for (int jj = std::max(0, j - nb); jj <= std::min(ny - 1, j + nb); ++jj)
                    for (int ii = std::max(0, i - nb); ii <= std::min(nx - 1, i + nb); ++ii) buf[(size_t)jj * nx + ii] = 1;
            }
        const int bf = std::max(1, ap.blocking_factor / ap.ref_ratio);
        std::vector<IV> pts;
        const int cnx = (nx + bf - 1) / bf, cny = (ny + bf - 1) / bf;
        for (int J = 0; J < cny; ++J)
            for (int I = 0; I < cnx; ++I) {
                bool any = false;
                for (int j = J * bf; j < std::min(ny, (J + 1) * bf) && !any; ++j)
                    for (int i = I * bf; i < std::min(nx, (I + 1) * bf) && !any; ++i) any = buf[(size_t)j * nx + i] != 0;
                if (any) pts.push_back(IV{I, J});
            }
        std::vector<Box> cl = cluster_tags(pts, ap.grid_eff);
        // chop in cluster units, then refine: every fine box stays coarsenable
        std::vector<Box> fine;
        const int r = ap.ref_ratio * bf;
        const int m = std::max(1, ap.fine_max_grid_size / r);
        for (const Box& c : cl)
            for (int y0 = c.lo[1]; y0 <= c.hi[1]; y0 += m)
                for (int x0 = c.lo[0]; x0 <= c.hi[0]; x0 += m) {
                    const int x1 = std::min(x0 + m - 1, c.hi[0]), y1 = std::min(y0 + m - 1, c.hi[1]);
                    fine.push_back(Box{{x0 * r, y0 * r}, {(x1 + 1) * r - 1, (y1 + 1) * r - 1}});
                }
        return fine;
    }

    Settings fine_settings() const {
        Settings s1 = s0;
        s1.nx = s0.nx * ap.ref_ratio;
        s1.ny = s0.ny * ap.ref_ratio;
        s1.time_end = 1.0e300;
        s1.max_step = -1;
        return s1;
    }

    std::unique_ptr<Lev> new_level(const std::vector<Box>& bxs) {
        auto F = std::make_unique<Lev>(fine_settings(), prm.data(), bxs);
        F->nthreads = nthreads;
        F->coarse_fill = [this](int i, int j, Real* v) { interp(i, j, v); };
        return F;
    }

    static bool same_boxes(const std::vector<Box>& a, const std::vector<Box>& b) {
        if (a.size() != b.size()) return false;
        for (size_t k = 0; k < a.size(); ++k)
            for (int d = 0; d < 2; ++d)
                if (a[k].lo[d] != b[k].lo[d] || a[k].hi[d] != b[k].hi[d]) return false;
        return true;
    }

    // AmrCore::regrid(0, time) with RemakeLevel / MakeNewLevelFromCoarse /
    // ClearLevel.
    void regrid(Real time) {
        if (ap.max_level < 1) return;
        ++regrids;
        std::vector<Box> nb = make_grids(error_est());
        if (nb.empty()) {
            L1.reset();
            return;
        }
        if (L1 && same_boxes(nb, L1->boxes)) return;
        auto F = new_level(nb);
        fill_coarse_patch(*F);
        if (L1) {  // RemakeLevel: FillPatch from the old level where it covers
            for (auto& f : F->Q.fabs)
                for (auto& g : L1->Q.fabs) {
                    const int ilo = std::max(f.gl[0], g.vb.lo[0]), ihi = std::min(f.gh[0], g.vb.hi[0]);
                    const int jlo = std::max(f.gl[1], g.vb.lo[1]), jhi = std::min(f.gh[1], g.vb.hi[1]);
                    for (int n = 0; n < nq; ++n)
                        for (int j = jlo; j <= jhi; ++j)
                            for (int i = ilo; i <= ihi; ++i) f(i, j, n) = g(i, j, n);
                }
            ++remakes;
        }
        fresh_aux(*F, time);
        L1 = std::move(F);
    }

    // The level loop of InitData (ZoomyAmr.cpp:137-155) on the current level-0
    // state, whose ghosts read as 0 (no FillPatch has run yet).
    void init_levels() {
        for (auto& f : L0->Q.fabs)
            for (int n = 0; n < nq; ++n)
                for (int j = f.gl[1]; j <= f.gh[1]; ++j)
                    for (int i = f.gl[0]; i <= f.gh[0]; ++i)
                        if (i < f.vb.lo[0] || i > f.vb.hi[0] || j < f.vb.lo[1] || j > f.vb.hi[1]) f(i, j, n) = 0.0;
        for (int it = 0; it < 4; ++it) {
            if (L1) fill_coarse_patch(*L1);
            const int old_finest = finest_level();
            regrid(0.0);
            if (old_finest == finest_level() && it > 0) break;
        }
        if (L1) {
            fill_coarse_patch(*L1);
            fresh_aux(*L1, 0.0);
        }
        for (int l = 0; l <= finest_level(); ++l) dt_level[l] = level(l).compute_dt();
    }

    // ZoomyAmr::TimeStep (ZoomyAmr.cpp:733-762).
    void time_step(int lev, Real time, int iteration) {
        if (lev < ap.max_level && iteration > 0) regrid(time);
        Lev& L = level(lev);
        dt_level[lev] = L.compute_dt();
        if (lev == 0) {
            const Real remaining = s0.time_end - time;
            if (remaining > 0.0 && dt_level[lev] > remaining) dt_level[lev] = remaining;
        }
        L.advance(time, dt_level[lev]);
        if (lev < finest_level()) {
            for (int i = 1; i <= ap.ref_ratio; ++i) time_step(lev + 1, time + (i - 1) * dt_level[lev + 1], i);
            average_down();
        }
    }

    // ComputeTotalMass: level-0 cells under level 1 are left out.
    Real total_mass() const {
        Real tot = 0.0;
        {
            Real s = 0.0;
            for (auto& f : L0->Q.fabs)
                for (int j = f.vb.lo[1]; j <= f.vb.hi[1]; ++j)
                    for (int i = f.vb.lo[0]; i <= f.vb.hi[0]; ++i) {
                        bool covered = false;
                        if (L1)
                            for (const Box& B : L1->boxes)
                                covered = covered || (i >= B.lo[0] / ap.ref_ratio && i <= B.hi[0] / ap.ref_ratio &&
                                                      j >= B.lo[1] / ap.ref_ratio && j <= B.hi[1] / ap.ref_ratio);
                        if (!covered) s += f(i, j, idx_h);
                    }
            tot += s * (L0->dx[0] * L0->dx[1]);
        }
        if (L1) {
            Real s = 0.0;
            for (auto& f : L1->Q.fabs)
                for (int j = f.vb.lo[1]; j <= f.vb.hi[1]; ++j)
                    for (int i = f.vb.lo[0]; i <= f.vb.hi[0]; ++i) s += f(i, j, idx_h);
            tot += s * (L1->dx[0] * L1->dx[1]);
        }
        return tot;
    }

    // The Evolve loop without plotfiles; returns the number of coarse steps.
    int evolve(Real* time_out, std::vector<Real>* dts, std::vector<Real>* masses, std::vector<int>* nlev,
               int step_limit) {
        Real time = 0.0;
        int step = 0;
        while (time < s0.time_end && (s0.max_step < 0 || step < s0.max_step)) {
            if (step_limit >= 0 && step >= step_limit) break;
            time_step(0, time, step);
            time += dt_level[0];
            ++step;
            if (dts) dts->push_back(dt_level[0]);
            if (masses) masses->push_back(total_mass());
            if (nlev) nlev->push_back(finest_level() + 1);
        }
        if (time_out) *time_out = time;
        return step;
    }
};

}  // namespace zorc
