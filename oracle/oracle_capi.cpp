// CPU ORACLE — C-ABI over zamr_oracle.hpp for the tests and the CPU-baseline
// leg of bench.py (test infrastructure only; see zamr_oracle.hpp).
//
// Compiled once per system (-I <system dir> picks Model.H / Numerics.H).
// Arrays cross the ABI as dense float64 (ncomp, ny, nx), i fastest.
//
// iset: [0] nx [1] ny [2] max_grid_size [3] spatial_order [4] implicit_source
//       [5] well_balanced [6] clamp_positivity [7] mood [8] face_based_rhs
//       [9] max_step [10..13] bc tag index x_lo,x_hi,y_lo,y_hi
//       [14] c_cfl_dimension [15] c_cfl_degree_factor [16] OpenMP threads
//       [17] implicit_global (JFNK source solve)
// dset: [0] prob_lo x [1] prob_lo y [2] prob_hi x [3] prob_hi y [4] cfl
//       [5] dtmin [6] dtmax [7] time_end
//
// AMR (orc_amr_*): ai: [0] max_level [1] ref_ratio [2] blocking_factor
//       [3] n_error_buf [4] fine max_grid_size; ad: [0] grid_eff
//       [1] tagging.threshold [2] tagging.b_max.
#include "Model.H"
#include "Numerics.H"
#include "zamr_oracle.hpp"
#include "zamr_oracle_amr.hpp"

using OSolver = zorc::Solver<Model, Numerics>;
using OAmr = zorc::Amr<Model, Numerics>;

static zorc::Settings settings_from(const int* is, const double* ds) {
    zorc::Settings s;
    s.nx = is[0];
    s.ny = is[1];
    s.max_grid_size = is[2];
    s.spatial_order = is[3];
    s.implicit_source = is[4] != 0;
    s.well_balanced = is[5] != 0;
    s.clamp_positivity = is[6] != 0;
    s.mood = is[7] != 0;
    s.face_based_rhs = is[8] != 0;
    s.max_step = is[9];
    for (int k = 0; k < 4; ++k) s.bc_idx[k] = is[10 + k];
    s.cfl_dimension = is[14];
    s.cfl_degree_factor = is[15];
    s.implicit_global = is[17] != 0;
    s.plo[0] = ds[0];
    s.plo[1] = ds[1];
    s.phi[0] = ds[2];
    s.phi[1] = ds[3];
    s.cfl = ds[4];
    s.dtmin = ds[5];
    s.dtmax = ds[6];
    s.time_end = ds[7];
    return s;
}

extern "C" {

int orc_nq() { return Model::n_dof_q; }
int orc_na() { return Model::n_dof_qaux; }
int orc_np() { return Model::n_parameters; }
int orc_dim() { return Model::dimension; }

int orc_ntags() { return (int)Model::get_boundary_tags().size(); }
// Copies tag k (sorted, as the model declares them) into buf; returns its length.
int orc_tag_name(int k, char* buf, int cap) {
    auto t = Model::get_boundary_tags();
    if (k < 0 || k >= (int)t.size()) return -1;
    int n = (int)t[k].size();
    if (n >= cap) n = cap - 1;
    std::memcpy(buf, t[k].data(), n);
    buf[n] = 0;
    return n;
}

// Copies parameter k's name (Model::parameter_names) into buf; returns its length.
int orc_param_name(int k, char* buf, int cap) {
    auto t = Model::parameter_names();
    if (k < 0 || k >= (int)t.size()) return -1;
    int n = (int)t[k].size();
    if (n >= cap) n = cap - 1;
    std::memcpy(buf, t[k].data(), n);
    buf[n] = 0;
    return n;
}

void orc_default_params(double* out) {
    auto v = Model::default_parameters();
    for (int i = 0; i < Model::n_parameters; ++i) out[i] = v[i];
}

void* orc_create(const int* is, const double* ds, const double* params) {
    const zorc::Settings s = settings_from(is, ds);
    auto* o = new OSolver(s, params);
    o->nthreads = is[16] > 0 ? is[16] : 1;
    return o;
}

void orc_destroy(void* h) { delete static_cast<OSolver*>(h); }
int orc_nboxes(void* h) { return (int)static_cast<OSolver*>(h)->boxes.size(); }

void orc_init_solution(void* h) {
    auto* o = static_cast<OSolver*>(h);
    o->init_solution(o->Q);
    o->init_solution(o->Qaux);
}
void orc_set_state(void* h, const double* a) { auto* o = static_cast<OSolver*>(h); o->scatter(o->Q, a, OSolver::nq); }
void orc_get_state(void* h, double* a) { auto* o = static_cast<OSolver*>(h); o->gather(o->Q, a, OSolver::nq); }
void orc_set_aux(void* h, const double* a) { auto* o = static_cast<OSolver*>(h); o->scatter(o->Qaux, a, OSolver::na); }
void orc_get_aux(void* h, double* a) { auto* o = static_cast<OSolver*>(h); o->gather(o->Qaux, a, OSolver::na); }

double orc_max_inv(void* h) { return static_cast<OSolver*>(h)->compute_max_inv(); }
double orc_compute_dt(void* h) { return static_cast<OSolver*>(h)->compute_dt(); }
void orc_update_state(void* h, double time) { static_cast<OSolver*>(h)->update_state(time); }
void orc_stage(void* h, int ord, double time, double dt) { static_cast<OSolver*>(h)->do_stage(ord, time, dt); }
void orc_advance(void* h, double time, double dt) { static_cast<OSolver*>(h)->advance(time, dt); }
double orc_total_mass(void* h) { return static_cast<OSolver*>(h)->total_mass(); }
long orc_mood_redos(void* h) { return static_cast<OSolver*>(h)->mood_redos; }
// Newton and GMRES iterations of the last global source solve.
void orc_jfnk_iters(void* h, long* newton, long* gmres) {
    auto* o = static_cast<OSolver*>(h);
    *newton = o->jfnk_newton;
    *gmres = o->jfnk_gmres;
}
void orc_set_tape(void* h, int mode) { static_cast<OSolver*>(h)->set_tape(mode); }

// Runs the Evolve loop; dts/masses (capacity cap) receive per-step values.
int orc_evolve(void* h, double* time_out, double* dts, double* masses, int cap, int step_limit) {
    auto* o = static_cast<OSolver*>(h);
    std::vector<double> d, m;
    int n = o->evolve(time_out, &d, &m, step_limit);
    for (int k = 0; k < n && k < cap; ++k) {
        if (dts) dts[k] = d[k];
        if (masses) masses[k] = m[k];
    }
    return n;
}

// The point-implicit Gaussian elimination alone (make_rhs.H:467-499) on an
// n x n system, n in {3, 4, 8}; returns 1 on success, 0 on the fallback path.
int orc_ge_solve(int n, const double* A, const double* rhs, double* x) {
    auto run = [&](auto tag) -> int {
        constexpr int N = decltype(tag)::value;
        amrex::SmallMatrix<double, N, N> M;
        amrex::SmallMatrix<double, N, 1> r, dq;
        for (int i = 0; i < N; ++i) {
            r(i, 0) = rhs[i];
            for (int j = 0; j < N; ++j) M(i, j) = A[i * N + j];
        }
        bool ok = zorc::ge_solve<N>(M, r, dq);
        for (int i = 0; i < N; ++i) x[i] = dq(i, 0);
        return ok ? 1 : 0;
    };
    if (n == 3) return run(std::integral_constant<int, 3>{});
    if (n == 4) return run(std::integral_constant<int, 4>{});
    if (n == 8) return run(std::integral_constant<int, 8>{});
    return -1;
}

// ComputeSourceField (imex_solver.H:21-59) on a dense (nq, ny, nx) state:
// aux = update_aux_variables(q, 0), then Model::source, cell by cell.
void orc_source_field(void* h, const double* q, double* S, double time) {
    auto* o = static_cast<OSolver*>(h);
    const long nx = o->s.nx, ny = o->s.ny, nc = nx * ny;
    for (long j = 0; j < ny; ++j)
        for (long i = 0; i < nx; ++i) {
            OSolver::VQ qq;
            OSolver::VA a;
            for (int n = 0; n < OSolver::nq; ++n) qq(n, 0) = q[n * nc + j * nx + i];
            for (int n = 0; n < OSolver::na; ++n) a(n, 0) = 0.0;
            const auto X = o->cell_center((int)i, (int)j);
            a = Model::update_aux_variables(qq, a, o->p, time, X);
            const auto Sc = Model::source(qq, a, o->p, time, X);
            for (int n = 0; n < OSolver::nq; ++n) S[n * nc + j * nx + i] = Sc(n, 0);
        }
}

// GMRES (the global source solve's linear solver) on a dense n x n system
// A x = b, row-major A; returns the iteration count.
long orc_gmres_dense(int n, const double* A, const double* b, double rtol, double* x) {
    std::vector<double> bv(b, b + n), xv;
    auto op = [&](const std::vector<double>& v, std::vector<double>& Av) {
        Av.assign(n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Av[i] += A[(size_t)i * n + j] * v[j];
    };
    const long its = zorc::gmres(xv, bv, rtol, 0.0, op);
    for (int i = 0; i < n; ++i) x[i] = xv[i];
    return its;
}

// ---------------- AMR ----------------
void* orc_amr_create(const int* is, const double* ds, const double* params, const int* ai, const double* ad) {
    zorc::AmrParams a;
    a.max_level = ai[0];
    a.ref_ratio = ai[1];
    a.blocking_factor = ai[2];
    a.n_error_buf = ai[3];
    a.fine_max_grid_size = ai[4];
    a.n_proper = ai[5];
    a.strip = ai[6];
    a.grid_eff = ad[0];
    a.tag_threshold = ad[1];
    a.tag_b_max = ad[2];
    OAmr* o = nullptr;
    try {
        o = new OAmr(settings_from(is, ds), params, a);
    } catch (const std::exception&) {
        return nullptr;  // amr.max_level or amr.ref_ratio out of range
    }
    o->nthreads = is[16] > 0 ? is[16] : 1;
    o->L[0]->nthreads = o->nthreads;
    return o;
}
void orc_amr_destroy(void* h) { delete static_cast<OAmr*>(h); }
void orc_amr_init_solution(void* h) {
    auto* o = static_cast<OAmr*>(h);
    o->L[0]->init_solution(o->L[0]->Q);
    o->L[0]->init_solution(o->L[0]->Qaux);
}
void orc_amr_init_levels(void* h) { static_cast<OAmr*>(h)->init_levels(); }
int orc_amr_finest(void* h) { return static_cast<OAmr*>(h)->finest_level(); }
// Boxes of level lev as (lo0, lo1, hi0, hi1) quadruples; returns the count.
int orc_amr_boxes(void* h, int lev, int* out, int cap) {
    auto* o = static_cast<OAmr*>(h);
    if (lev > o->finest_level()) return 0;
    auto& bx = o->level(lev).boxes;
    for (int k = 0; k < (int)bx.size() && k < cap; ++k) {
        out[4 * k + 0] = bx[k].lo[0];
        out[4 * k + 1] = bx[k].lo[1];
        out[4 * k + 2] = bx[k].hi[0];
        out[4 * k + 3] = bx[k].hi[1];
    }
    return (int)bx.size();
}
// Dense (ncomp, ny_lev, nx_lev) copy of a level's valid cells (field 0 = Q,
// 1 = Qaux); cells no box covers are left as the caller filled them.
static void amr_copy(OAmr* o, int lev, int field, double* a, bool to_level) {
    auto& L = o->level(lev);
    auto& mf = field == 0 ? L.Q : L.Qaux;
    const int nc = field == 0 ? OSolver::nq : OSolver::na;
    const long nx = L.s.nx, ny = L.s.ny;
    for (auto& f : mf.fabs)
        for (int n = 0; n < nc; ++n)
            for (int j = f.vb.lo[1]; j <= f.vb.hi[1]; ++j)
                for (int i = f.vb.lo[0]; i <= f.vb.hi[0]; ++i) {
                    double& g = a[((long)n * ny + j) * nx + i];
                    if (to_level) f(i, j, n) = g;
                    else g = f(i, j, n);
                }
}
void orc_amr_get(void* h, int lev, int field, double* a) { amr_copy(static_cast<OAmr*>(h), lev, field, a, false); }
void orc_amr_set(void* h, int lev, int field, const double* a) {
    amr_copy(static_cast<OAmr*>(h), lev, field, const_cast<double*>(a), true);
}
void orc_amr_time_step(void* h, double time, int iteration, double* dt) {
    auto* o = static_cast<OAmr*>(h);
    o->time_step(0, time, iteration);
    if (dt) *dt = o->dt_level[0];
}
double orc_amr_total_mass(void* h) { return static_cast<OAmr*>(h)->total_mass(); }
void orc_amr_regrid(void* h, int lbase, double time) { static_cast<OAmr*>(h)->regrid(lbase, time); }
void orc_amr_average_down(void* h, int lev) { static_cast<OAmr*>(h)->average_down(lev); }
long orc_amr_counts(void* h, int which) {
    auto* o = static_cast<OAmr*>(h);
    if (which == 0) return o->regrids;
    if (which == 1) return o->remakes;
    const int lev = which - 1;  // 2: level 1, 3: level 2, ...
    return lev <= o->finest_level() ? o->level(lev).mood_redos : 0;
}
// ErrorEst of level lev as it stands, one byte per cell (raster order).
void orc_amr_error_est(void* h, int lev, unsigned char* tags) {
    auto t = static_cast<OAmr*>(h)->error_est(lev);
    std::memcpy(tags, t.data(), t.size());
}
// Level-1 boxes MakeNewGrids derives from a level-0 tag map.
int orc_amr_make_grids(void* h, const unsigned char* tags, int* out, int cap) {
    auto* o = static_cast<OAmr*>(h);
    std::vector<unsigned char> t(tags, tags + (size_t)o->s0.nx * o->s0.ny);
    auto bx = o->make_grids(t);
    for (int k = 0; k < (int)bx.size() && k < cap; ++k) {
        out[4 * k + 0] = bx[k].lo[0];
        out[4 * k + 1] = bx[k].lo[1];
        out[4 * k + 2] = bx[k].hi[0];
        out[4 * k + 3] = bx[k].hi[1];
    }
    return (int)bx.size();
}
// cell_cons_interp of level lev-1 at level-lev cell (i, j).
void orc_amr_interp(void* h, int lev, int i, int j, double* out) { static_cast<OAmr*>(h)->interp(lev, i, j, out); }
// Berger-Rigoutsos clusters of n points (x, y pairs) at efficiency eff.
int orc_cluster(const int* pts, int n, double eff, int* out, int cap) {
    std::vector<zorc::IV> v(n);
    for (int k = 0; k < n; ++k) v[k] = zorc::IV{pts[2 * k], pts[2 * k + 1]};
    auto bx = zorc::cluster_tags(v, eff);
    for (int k = 0; k < (int)bx.size() && k < cap; ++k) {
        out[4 * k + 0] = bx[k].lo[0];
        out[4 * k + 1] = bx[k].lo[1];
        out[4 * k + 2] = bx[k].hi[0];
        out[4 * k + 3] = bx[k].hi[1];
    }
    return (int)bx.size();
}
// Runs the Evolve loop over the hierarchy; per-step dt, mass and level count.
int orc_amr_evolve(void* h, double* time_out, double* dts, double* masses, int* nlev, int cap, int step_limit) {
    auto* o = static_cast<OAmr*>(h);
    std::vector<double> d, m;
    std::vector<int> l;
    int n = o->evolve(time_out, &d, &m, &l, step_limit);
    for (int k = 0; k < n && k < cap; ++k) {
        if (dts) dts[k] = d[k];
        if (masses) masses[k] = m[k];
        if (nlev) nlev[k] = l[k];
    }
    return n;
}
}
