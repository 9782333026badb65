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
// dset: [0] prob_lo x [1] prob_lo y [2] prob_hi x [3] prob_hi y [4] cfl
//       [5] dtmin [6] dtmax [7] time_end
#include "Model.H"
#include "Numerics.H"
#include "zamr_oracle.hpp"

using OSolver = zorc::Solver<Model, Numerics>;

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

void orc_default_params(double* out) {
    auto v = Model::default_parameters();
    for (int i = 0; i < Model::n_parameters; ++i) out[i] = v[i];
}

void* orc_create(const int* is, const double* ds, const double* params) {
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
    s.plo[0] = ds[0];
    s.plo[1] = ds[1];
    s.phi[0] = ds[2];
    s.phi[1] = ds[3];
    s.cfl = ds[4];
    s.dtmin = ds[5];
    s.dtmax = ds[6];
    s.time_end = ds[7];
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
}
