#pragma once
// CPU ORACLE — test infrastructure only, never part of the product path.
//
// The AMR advance of Source/ZoomyAmr.cpp restated on the oracle's Solver
// levels (0 .. amr.max_level <= 3, one amr.ref_ratio of 2 or 4 between every
// pair): nsubsteps = ref_ratio (:96-106), InitData's level loop (:137-155),
// ErrorEst (:225-253), regrid with RemakeLevel / MakeNewLevelFromCoarse /
// ClearLevel (:161-219), FillPatch / FillCoarsePatch (:258-288), the
// recursive subcycled TimeStep with average_down (:733-762), ComputeTotalMass
// over the composite (:517-536).
//
// The AMR machinery itself lives in AMReX, an external dependency the
// reference does not vendor (Exec/GNUmakefile: AMREX_HOME; no version pin,
// the container's /opt/amrex).  Its published algorithms are restated here:
//   * cell_cons_interp = CellConservativeLinear with linear limiting
//     (AMReX_MFInterp_C.H, mf_cell_cons_lin_interp_llslope + _lin_interp):
//     central slopes, one limiting factor per direction shared by all
//     components, fine = crse + xoff * sx + yoff * sy, xoff = (i - r ic +
//     1/2) / r - 1/2;
//   * average_down (AMReX_MultiFabUtil_2D_C.H, amrex_avgdown): the mean of
//     the ratio^2 children, summed j-outer i-inner, times 1/ratio^2;
//   * AmrMesh::MakeNewGrids / AmrCore::regrid (AMReX_AmrMesh.cpp,
//     AMReX_AmrCore.cpp): at most one new level per regrid; from the finest
//     level that may tag down to lbase, tags buffered by n_error_buf,
//     coarsened by blocking_factor / ref_ratio, the next finer new grids
//     projected down (coarsened, grown by n_proper) and tagged, cells outside
//     the proper-nesting domain (the coarsened lbase grids shrunk by n_proper
//     inside the domain, refined and shrunk again per level above) cleared,
//     Berger-Rigoutsos clusters (AMReX_Cluster.cpp: ClusterList::chop,
//     Cluster::chop, FindCut) to efficiency grid_eff cut by that domain,
//     refined back; a level is remade when its grids or those below changed.
//
// Conventions where the reference reads memory it never wrote (each is
// undefined there; this restatement and the device code share the choice):
//   * coarse cells outside the domain in the interpolation stencil take the
//     nearest interior value (first-order extrapolation, the BCRec foextrap
//     ZoomyAmr.cpp:88-94 declares; the reference's fill functor is NullFill);
//     a stencil neighbour inside the domain that the coarse level does not
//     cover (one cell beyond it, where proper nesting leaves a single cell)
//     repeats the stencil's centre;
//   * ErrorEst before any FillPatch (level 0 during InitData) reads ghost
//     cells as 0; on a level created or remade since its last stage, ghost
//     cells inside the domain as that fill left them and beyond the domain 0;
//   * the Qaux of a new or remade fine level is update_aux_variables(Q, 0)
//     on its valid cells (RemakeLevel / MakeNewLevelFromCoarse leave it unset
//     and ComputeDt reads it next).
// The boxes of a refined level are cut from the clusters to
// amr.max_grid_size, in x to whole stage strips of the device kernel
// (AmrParams::strip), as the device code cuts them.  With two levels only
// their union enters the results; with three, ErrorEst on level 1 reads its
// ghost cells as the last fill left them, so the level-1 cut enters level
// 2's tags (AMReX's own cut, BoxList simplify + maxSize, would shift which
// cells read a neighbour's fresh value and which a stale ghost).
//
// Pinned by: interpolation conservation (children average to the parent
// exactly in exact arithmetic, to round-off here) and exactness on linear
// data at ratios 2 and 4, average_down(interp) == identity on linear data,
// clustering invariants (every tag covered, disjoint boxes, efficiency >=
// grid_eff or unsplittable), proper nesting after every step and the exact
// n_proper erosion of a regrid from level 1, decomposition invariance (fine
// max_grid_size, two levels), the no-tag run equal to the single-level run
// bit for bit, and the all-tagged finest level of 2 and 3 levels at ratios 2
// and 4 equal to the single-level run at that resolution bit for bit.
#include <algorithm>
#include <array>
#include <memory>
#include <stdexcept>

#include "zamr_oracle.hpp"

namespace zorc {

struct AmrParams {
    int max_level = 1;           // amr.max_level (this restatement: 0..3)
    int ref_ratio = 2;           // amr.ref_ratio (2 or 4, the same between every pair of levels)
    int n_proper = 1;            // amr.n_proper (AMReX default)
    int strip = 0;               // x cut of fine boxes: whole strips of this many cells (0: as y)
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

// A region of cells given by a byte mask (nx x ny, raster order) cut into
// rectangles: each row's runs of set cells, with runs of consecutive rows
// that start and end alike merged into one rectangle.  Used for a cluster
// box intersected with the proper-nesting domain (ClusterList::intersect);
// only the union of the rectangles enters the results.
inline std::vector<Box> mask_rects(const std::vector<unsigned char>& mask, int nx, const Box& c) {
    std::vector<Box> done, open;
    for (int j = c.lo[1]; j <= c.hi[1] + 1; ++j) {
        std::vector<std::pair<int, int>> runs;
        if (j <= c.hi[1])
            for (int i = c.lo[0]; i <= c.hi[0];) {
                if (!mask[(size_t)j * nx + i]) {
                    ++i;
                    continue;
                }
                int e = i;
                while (e + 1 <= c.hi[0] && mask[(size_t)j * nx + e + 1]) ++e;
                runs.push_back({i, e});
                i = e + 1;
            }
        std::vector<Box> next;
        for (const Box& o : open) {
            bool cont = false;
            for (auto& r : runs)
                if (r.first == o.lo[0] && r.second == o.hi[0]) cont = true;
            if (cont) {
                Box g = o;
                g.hi[1] = j;
                next.push_back(g);
            } else {
                done.push_back(o);
            }
        }
        for (auto& r : runs) {
            bool cont = false;
            for (const Box& o : open)
                if (r.first == o.lo[0] && r.second == o.hi[0]) cont = true;
            if (!cont) next.push_back(Box{{r.first, j}, {r.second, j}});
        }
        open.swap(next);
    }
    std::sort(done.begin(), done.end(), [](const Box& x, const Box& y) {
        return x.lo[1] != y.lo[1] ? x.lo[1] < y.lo[1] : x.lo[0] < y.lo[0];
    });
    return done;
}

inline bool same_boxes(const std::vector<Box>& a, const std::vector<Box>& b) {
    if (a.size() != b.size()) return false;
    for (size_t k = 0; k < a.size(); ++k)
        for (int d = 0; d < 2; ++d)
            if (a[k].lo[d] != b[k].lo[d] || a[k].hi[d] != b[k].hi[d]) return false;
    return true;
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
    std::vector<std::unique_ptr<Lev>> L;      // levels 0 .. finest
    std::vector<std::vector<int>> owner;      // per level >= 1: box index of each cell, -1 where none
    std::vector<char> fresh;                  // per level: created or remade since its last stage
    std::vector<Real> dt_level;
    int nthreads = 1;
    long regrids = 0, remakes = 0;

    Amr(const Settings& s, const Real* params, const AmrParams& a) : s0(s), prm(params, params + Lev::np), ap(a) {
        if (ap.max_level < 0 || ap.max_level > 3) throw std::invalid_argument("amr.max_level must be 0..3");
        if (ap.ref_ratio != 2 && ap.ref_ratio != 4) throw std::invalid_argument("amr.ref_ratio must be 2 or 4");
        L.push_back(std::make_unique<Lev>(s0, prm.data()));
        owner.resize(1);
        fresh.assign(1, 0);
        dt_level.assign(ap.max_level + 1, 1.0e100);
    }

    int finest_level() const { return (int)L.size() - 1; }
    Lev& level(int l) { return *L.at(l); }
    const Lev& level(int l) const { return *L.at(l); }
    int nx_of(int l) const { return s0.nx * ipow(ap.ref_ratio, l); }
    int ny_of(int l) const { return s0.ny * ipow(ap.ref_ratio, l); }
    static int ipow(int r, int l) {
        int v = 1;
        while (l-- > 0) v *= r;
        return v;
    }
    int bf_lev() const { return std::max(1, ap.blocking_factor / ap.ref_ratio); }

    // ---------- level data access ----------
    // The box of level l holding valid cell (i, j), or -1.
    int box_at(int l, int i, int j) const {
        const int nx = nx_of(l), ny = ny_of(l);
        if (i < 0 || j < 0 || i >= nx || j >= ny) return -1;
        if (l == 0) {
            for (size_t b = 0; b < L[0]->Q.fabs.size(); ++b) {
                const Box& v = L[0]->Q.fabs[b].vb;
                if (i >= v.lo[0] && i <= v.hi[0] && j >= v.lo[1] && j <= v.hi[1]) return (int)b;
            }
            return -1;
        }
        return owner[l][(size_t)j * nx + i];
    }
    void build_owner(int l) {
        const int nx = nx_of(l), ny = ny_of(l);
        owner[l].assign((size_t)nx * ny, -1);
        const auto& bx = L[l]->boxes;
        for (size_t b = 0; b < bx.size(); ++b)
            for (int j = bx[b].lo[1]; j <= bx[b].hi[1]; ++j)
                for (int i = bx[b].lo[0]; i <= bx[b].hi[0]; ++i) owner[l][(size_t)j * nx + i] = (int)b;
    }

    // cell_cons_interp of level l-1 at level-l cell (i, j) (inside the
    // domain).  Stencil cells beyond the domain repeat the edge cell; a
    // stencil neighbour that level l-1 does not cover (the centre always is,
    // by proper nesting) repeats the centre.
    void interp(int l, int i, int j, Real* out) const {
        const int r = ap.ref_ratio, lc = l - 1;
        const int cnx = nx_of(lc), cny = ny_of(lc);
        const int ic = i / r, jc = j / r;
        const int b0 = box_at(lc, ic, jc);
        if (b0 < 0) throw std::logic_error("interp: coarse centre not covered (proper nesting violated)");
        const Fab& f0 = level(lc).Q.fabs[b0];
        auto at = [&](int ii, int jj, int n) {
            ii = std::min(std::max(ii, 0), cnx - 1);
            jj = std::min(std::max(jj, 0), cny - 1);
            const int b = box_at(lc, ii, jj);
            if (b < 0) return f0(ic, jc, n);
            return level(lc).Q.fabs[b](ii, jj, n);
        };
        Real sfx = 1.0, sfy = 1.0;
        Real slx[nq], sly[nq], u0[nq];
        for (int n = 0; n < nq; ++n) {
            const Real u = f0(ic, jc, n);
            u0[n] = u;
            for (int d = 0; d < 2; ++d) {
                const Real um = d == 0 ? at(ic - 1, jc, n) : at(ic, jc - 1, n);
                const Real up = d == 0 ? at(ic + 1, jc, n) : at(ic, jc + 1, n);
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
    void fill_coarse_patch(int l) {
        Lev& F = level(l);
        Real v[nq];
        for (auto& f : F.Q.fabs)
            for (int j = f.gl[1]; j <= f.gh[1]; ++j)
                for (int i = f.gl[0]; i <= f.gh[0]; ++i) {
                    if (i < 0 || i >= F.s.nx || j < 0 || j >= F.s.ny) continue;
                    interp(l, i, j, v);
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

    // average_down of level l+1 into level l (TimeStep, ZoomyAmr.cpp:759-760).
    void average_down(int l) {
        const int r = ap.ref_ratio;
        const Real vol = 1.0 / (Real)(r * r);
        for (auto& f : level(l + 1).Q.fabs)
            for (int jc = f.vb.lo[1] / r; jc <= f.vb.hi[1] / r; ++jc)
                for (int ic = f.vb.lo[0] / r; ic <= f.vb.hi[0] / r; ++ic) {
                    const int b = box_at(l, ic, jc);
                    if (b < 0) throw std::logic_error("average_down: coarse cell not covered");
                    Fab& c = level(l).Q.fabs[b];
                    for (int n = 0; n < nq; ++n) {
                        Real sum = 0.0;
                        for (int jj = 0; jj < r; ++jj)
                            for (int ii = 0; ii < r; ++ii) sum += f(ic * r + ii, jc * r + jj, n);
                        c(ic, jc, n) = vol * sum;
                    }
                }
    }
    void average_down() { average_down(0); }

    // ErrorEst (ZoomyAmr.cpp:225-253) on level l, the neighbours read from
    // the box's ghost cells as they stand: the last stage's FillPatch; on a
    // level created or remade since its last stage, the fill that made it
    // (inside the domain; 0 beyond); on level 0 before any stage, 0.
    std::vector<unsigned char> error_est(int l = 0) const {
        const int nx = nx_of(l), ny = ny_of(l);
        std::vector<unsigned char> tag((size_t)nx * ny, 0);
        for (auto& f : level(l).Q.fabs)
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
                    if (std::sqrt(grad) > ap.tag_threshold || f(i, j, idx_h) > 1.0e-2) tag[(size_t)j * nx + i] = 1;
                }
        return tag;
    }

    // TagBoxArray buffer (n_error_buf, clipped to the domain) and coarsen by
    // bf: the tag map of level l in cluster units.
    std::vector<unsigned char> buffer_coarsen(const std::vector<unsigned char>& tag, int l) const {
        const int nx = nx_of(l), ny = ny_of(l), nb = ap.n_error_buf, bf = bf_lev();
        std::vector<unsigned char> buf((size_t)nx * ny, 0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (!tag[(size_t)j * nx + i]) continue;
                for (int jj = std::max(0, j - nb); jj <= std::min(ny - 1, j + nb); ++jj)
                    for (int ii = std::max(0, i - nb); ii <= std::min(nx - 1, i + nb); ++ii) buf[(size_t)jj * nx + ii] = 1;
            }
        const int cnx = (nx + bf - 1) / bf, cny = (ny + bf - 1) / bf;
        std::vector<unsigned char> c((size_t)cnx * cny, 0);
        for (int J = 0; J < cny; ++J)
            for (int I = 0; I < cnx; ++I) {
                bool any = false;
                for (int j = J * bf; j < std::min(ny, (J + 1) * bf) && !any; ++j)
                    for (int i = I * bf; i < std::min(nx, (I + 1) * bf) && !any; ++i) any = buf[(size_t)j * nx + i] != 0;
                c[(size_t)J * cnx + I] = any ? 1 : 0;
            }
        return c;
    }

    // Fine boxes (level l+1) from the clusters of a coarsened tag map of
    // level l: each cluster cut by the proper-nesting mask pn (cluster
    // units; empty = everywhere allowed) into rectangles, each cut to
    // amr.max_grid_size in cluster units (in x, when it is a whole number of
    // cluster units, to the most whole stage strips of ap.strip cells that
    // fit), refined by bf * ref_ratio.
    std::vector<Box> grids_from(const std::vector<unsigned char>& ctag, int l, const std::vector<unsigned char>& pn) const {
        const int bf = bf_lev();
        const int cnx = (nx_of(l) + bf - 1) / bf, cny = (ny_of(l) + bf - 1) / bf;
        std::vector<IV> pts;
        for (int J = 0; J < cny; ++J)
            for (int I = 0; I < cnx; ++I)
                if (ctag[(size_t)J * cnx + I]) pts.push_back(IV{I, J});
        std::vector<Box> cl = cluster_tags(pts, ap.grid_eff);
        std::sort(cl.begin(), cl.end(), [](const Box& x, const Box& y) {
            return x.lo[1] != y.lo[1] ? x.lo[1] < y.lo[1] : x.lo[0] < y.lo[0];
        });
        std::vector<Box> parts;
        for (const Box& c : cl) {
            if (pn.empty()) {
                parts.push_back(c);
                continue;
            }
            for (const Box& q : mask_rects(pn, cnx, c)) parts.push_back(q);
        }
        const int r = ap.ref_ratio * bf;
        const int m = std::max(1, ap.fine_max_grid_size / r);
        int mx = m;
        if (ap.strip > 0) {
            const int wx = (ap.fine_max_grid_size / ap.strip) * ap.strip;
            if (wx >= r && wx % r == 0) mx = wx / r;
        }
        std::vector<Box> fine;
        for (const Box& c : parts)
            for (int y0 = c.lo[1]; y0 <= c.hi[1]; y0 += m)
                for (int x0 = c.lo[0]; x0 <= c.hi[0]; x0 += mx) {
                    const int x1 = std::min(x0 + mx - 1, c.hi[0]), y1 = std::min(y0 + m - 1, c.hi[1]);
                    fine.push_back(Box{{x0 * r, y0 * r}, {(x1 + 1) * r - 1, (y1 + 1) * r - 1}});
                }
        return fine;
    }

    // The level-0 tag map to level-1 boxes (MakeNewGrids with lbase 0 and
    // no finer level to nest).
    std::vector<Box> make_grids(const std::vector<unsigned char>& tag) const {
        return grids_from(buffer_coarsen(tag, 0), 0, {});
    }

    // Proper-nesting domains of MakeNewGrids (AMReX_AmrMesh.cpp) in cluster
    // units of levels lbase .. max_crse: at lbase the cells of the coarsened
    // level-lbase grids farther than n_proper from any domain cell they do
    // not cover; above, the complement refined by ref_ratio and grown by
    // n_proper again.  Masks: 1 = allowed.
    std::vector<std::vector<unsigned char>> nesting(int lbase, int max_crse) const {
        const int bf = bf_lev(), np = ap.n_proper;
        std::vector<std::vector<unsigned char>> comp(max_crse + 1), pn(max_crse + 1);
        for (int l = lbase; l <= max_crse; ++l) {
            const int cnx = (nx_of(l) + bf - 1) / bf, cny = (ny_of(l) + bf - 1) / bf;
            std::vector<unsigned char> base((size_t)cnx * cny, 0);
            if (l == lbase) {
                std::vector<unsigned char> cov((size_t)cnx * cny, 0);
                for (const Box& b : level(l).boxes)
                    for (int J = b.lo[1] / bf; J <= b.hi[1] / bf; ++J)
                        for (int I = b.lo[0] / bf; I <= b.hi[0] / bf; ++I) cov[(size_t)J * cnx + I] = 1;
                for (size_t k = 0; k < base.size(); ++k) base[k] = cov[k] ? 0 : 1;
            } else {
                const int pcnx = (nx_of(l - 1) + bf - 1) / bf;
                const int rr = ap.ref_ratio;  // (ref_ratio * bf) / bf
                for (int J = 0; J < cny; ++J)
                    for (int I = 0; I < cnx; ++I) base[(size_t)J * cnx + I] = comp[l - 1][(size_t)(J / rr) * pcnx + I / rr];
            }
            comp[l].assign((size_t)cnx * cny, 0);
            for (int J = 0; J < cny; ++J)
                for (int I = 0; I < cnx; ++I) {
                    if (!base[(size_t)J * cnx + I]) continue;
                    for (int jj = std::max(0, J - np); jj <= std::min(cny - 1, J + np); ++jj)
                        for (int ii = std::max(0, I - np); ii <= std::min(cnx - 1, I + np); ++ii)
                            comp[l][(size_t)jj * cnx + ii] = 1;
                }
            pn[l].resize(comp[l].size());
            for (size_t k = 0; k < pn[l].size(); ++k) pn[l][k] = comp[l][k] ? 0 : 1;
        }
        return pn;
    }

    Settings level_settings(int l) const {
        Settings s = s0;
        s.nx = nx_of(l);
        s.ny = ny_of(l);
        s.time_end = 1.0e300;
        s.max_step = -1;
        return s;
    }

    std::unique_ptr<Lev> new_level(int l, const std::vector<Box>& bxs) {
        auto F = std::make_unique<Lev>(level_settings(l), prm.data(), bxs);
        F->nthreads = nthreads;
        F->coarse_fill = [this, l](int i, int j, Real* v) { interp(l, i, j, v); };
        return F;
    }

    void set_level(int l, std::unique_ptr<Lev> F) {
        if ((int)L.size() <= l) {
            L.resize(l + 1);
            owner.resize(l + 1);
            fresh.resize(l + 1, 0);
        }
        L[l] = std::move(F);
        build_owner(l);
        fresh[l] = 1;
    }

    void clear_above(int l) {
        while (finest_level() > l) {
            L.pop_back();
            owner.pop_back();
            fresh.pop_back();
        }
    }

    // AmrMesh::MakeNewGrids(lbase) + AmrCore::regrid(lbase, time) with
    // RemakeLevel / MakeNewLevelFromCoarse / ClearLevel (ZoomyAmr.cpp:161-219):
    // from the finest level that can tag (at most one new level) down to
    // lbase, each level's tags, buffered and coarsened, with the next finer
    // new grids projected down (coarsened, grown by n_proper) and set, cells
    // outside the proper-nesting domain cleared, clustered, cut by that
    // domain; then every level above lbase remade when its grids or the
    // grids below changed, made when new, cleared when gone.
    void regrid(int lbase, Real time) {
        if (lbase >= ap.max_level) return;
        ++regrids;
        const int max_crse = std::min(finest_level(), ap.max_level - 1);
        const auto pn = nesting(lbase, max_crse);
        const int bf = bf_lev(), r = ap.ref_ratio, np = ap.n_proper;
        std::vector<std::vector<Box>> ng(max_crse + 2);
        int new_finest = lbase;
        for (int levc = max_crse; levc >= lbase; --levc) {
            const int levf = levc + 1;
            auto ct = buffer_coarsen(error_est(levc), levc);
            const int cnx = (nx_of(levc) + bf - 1) / bf, cny = (ny_of(levc) + bf - 1) / bf;
            if (levf < new_finest) {
                for (const Box& g : ng[levf + 1]) {
                    // level levf+1 -> levf (coarsen by r), grow by n_proper,
                    // -> levc (coarsen by r), -> cluster units (coarsen by bf)
                    int lo[2], hi[2];
                    for (int d = 0; d < 2; ++d) {
                        lo[d] = std::max(g.lo[d] / r - np, 0) / r / bf;  // clipped to the domain
                        hi[d] = (g.hi[d] / r + np) / r / bf;
                    }
                    for (int J = lo[1]; J <= std::min(hi[1], cny - 1); ++J)
                        for (int I = lo[0]; I <= std::min(hi[0], cnx - 1); ++I) ct[(size_t)J * cnx + I] = 1;
                }
            }
            const auto& allowed = pn[levc];
            bool full = true;
            for (unsigned char v : allowed) full = full && v;
            for (size_t k = 0; k < ct.size(); ++k)
                if (!allowed[k]) ct[k] = 0;
            auto bx = grids_from(ct, levc, full ? std::vector<unsigned char>{} : allowed);
            if (!bx.empty()) new_finest = std::max(new_finest, levf);
            ng[levf] = bx;
        }
        bool coarse_changed = false;
        for (int lev = lbase + 1; lev <= new_finest; ++lev) {
            if (lev <= finest_level()) {
                const bool changed = !same_boxes(ng[lev], level(lev).boxes);
                if (changed || coarse_changed) remake(lev, time, changed ? ng[lev] : level(lev).boxes);
                coarse_changed = changed;
            } else {
                auto F = new_level(lev, ng[lev]);
                set_level(lev, std::move(F));
                fill_coarse_patch(lev);
                fresh_aux(level(lev), time);
            }
        }
        clear_above(new_finest);
    }
    void regrid(Real time) { regrid(0, time); }

    // RemakeLevel: FillPatch from the old level where it covers, else from
    // the level below.
    void remake(int lev, Real time, const std::vector<Box>& bxs) {
        auto F = new_level(lev, bxs);
        std::unique_ptr<Lev> old = std::move(L[lev]);
        std::vector<int> old_owner = std::move(owner[lev]);
        set_level(lev, std::move(F));
        fill_coarse_patch(lev);
        for (auto& f : level(lev).Q.fabs)
            for (auto& g : old->Q.fabs) {
                const int ilo = std::max(f.gl[0], g.vb.lo[0]), ihi = std::min(f.gh[0], g.vb.hi[0]);
                const int jlo = std::max(f.gl[1], g.vb.lo[1]), jhi = std::min(f.gh[1], g.vb.hi[1]);
                for (int n = 0; n < nq; ++n)
                    for (int j = jlo; j <= jhi; ++j)
                        for (int i = ilo; i <= ihi; ++i) f(i, j, n) = g(i, j, n);
            }
        ++remakes;
        fresh_aux(level(lev), time);
    }

    // The level loop of InitData (ZoomyAmr.cpp:137-155) on the current level-0
    // state, whose ghosts read as 0 (no FillPatch has run yet).
    void init_levels() {
        for (auto& f : L[0]->Q.fabs)
            for (int n = 0; n < nq; ++n)
                for (int j = f.gl[1]; j <= f.gh[1]; ++j)
                    for (int i = f.gl[0]; i <= f.gh[0]; ++i)
                        if (i < f.vb.lo[0] || i > f.vb.hi[0] || j < f.vb.lo[1] || j > f.vb.hi[1]) f(i, j, n) = 0.0;
        for (int it = 0; it < 4; ++it) {
            for (int l = 1; l <= finest_level(); ++l) fill_coarse_patch(l);
            const int old_finest = finest_level();
            regrid(0, 0.0);
            if (old_finest == finest_level() && it > 0) break;
        }
        for (int l = 1; l <= finest_level(); ++l) {
            fill_coarse_patch(l);
            fresh_aux(level(l), 0.0);
        }
        for (int l = 0; l <= finest_level(); ++l) dt_level[l] = level(l).compute_dt();
    }

    // ZoomyAmr::TimeStep (ZoomyAmr.cpp:733-762).
    void time_step(int lev, Real time, int iteration) {
        if (lev < ap.max_level && iteration > 0) regrid(lev, time);
        Lev& Lv = level(lev);
        dt_level[lev] = Lv.compute_dt();
        if (lev == 0) {
            const Real remaining = s0.time_end - time;
            if (remaining > 0.0 && dt_level[lev] > remaining) dt_level[lev] = remaining;
        }
        Lv.advance(time, dt_level[lev]);
        fresh[lev] = 0;
        if (lev < finest_level()) {
            for (int i = 1; i <= ap.ref_ratio; ++i) time_step(lev + 1, time + (i - 1) * dt_level[lev + 1], i);
            average_down(lev);
        }
    }

    // ComputeTotalMass: cells of a level under the next finer one are left out.
    Real total_mass() const {
        Real tot = 0.0;
        const int r = ap.ref_ratio;
        for (int l = 0; l <= finest_level(); ++l) {
            const Lev& Lv = level(l);
            Real s = 0.0;
            for (auto& f : Lv.Q.fabs)
                for (int j = f.vb.lo[1]; j <= f.vb.hi[1]; ++j)
                    for (int i = f.vb.lo[0]; i <= f.vb.hi[0]; ++i) {
                        bool covered = false;
                        if (l < finest_level())
                            for (const Box& B : level(l + 1).boxes)
                                covered = covered || (i >= B.lo[0] / r && i <= B.hi[0] / r && j >= B.lo[1] / r &&
                                                      j <= B.hi[1] / r);
                        if (!covered) s += f(i, j, idx_h);
                    }
            tot += s * (Lv.dx[0] * Lv.dx[1]);
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
