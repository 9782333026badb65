// C-ABI of include/zamr.h for one generated system.  Compiled once per system
// with -I <system dir> so that "Model.H" / "Numerics.H" are that system's
// printer output; the MarchConstants values come from a per-variant object
// (zamr_march.cpp) linked beside this one.
#include "Model.H"
#include "Numerics.H"
#include "zamr.h"
#include "zamr_amr.hpp"
#include "zamr_driver.hpp"
#include "zamr_plotfile.hpp"
#include <memory>
#include <mutex>
#if __has_include("Adjoint.H")
#include "Adjoint.H"
#include "zamr_adjoint.hpp"
#define ZAMR_HAS_ADJOINT 1
#else
#define ZAMR_HAS_ADJOINT 0
#endif

extern "C" void zamr_internal_march(int* dim, int* deg);

using Sol = zamr::Solver<Model, Numerics>;
using Amr = zamr::AmrDriver<Model, Numerics>;
#if ZAMR_HAS_ADJOINT
using AdjDrv = zamr::AdjointDriver<Model, Numerics, Adjoint>;
#endif

struct zamr_ctx {
    Sol s;
    std::unique_ptr<Amr> amr;  // amr.max_level >= 1: the hierarchy over s (level 0)
    std::string err;
#if ZAMR_HAS_ADJOINT
    std::unique_ptr<AdjDrv> adj;
#endif
    ~zamr_ctx() {
#if ZAMR_HAS_ADJOINT
        adj.reset();  // before the solver frees its stream
#endif
        amr.reset();
    }
    Sol& level(int lev) { return amr ? amr->level(lev) : s; }
};

namespace {

template <class F>
int guard(zamr_ctx* c, F&& f) {
    try {
        f();
        return ZAMR_OK;
    } catch (const zamr::Error& e) {
        if (c) c->err = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        if (c) c->err = e.what();
        return ZAMR_EINVAL;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return ZAMR_EHIP;
    }
}

int copy_name(const std::vector<std::string>& v, int k, char* buf, int cap) {
    if (k < 0 || k >= (int)v.size() || !buf || cap <= 0) return ZAMR_EINVAL;
    const int n = std::min<int>((int)v[k].size(), cap - 1);
    std::memcpy(buf, v[k].data(), n);
    buf[n] = 0;
    return ZAMR_OK;
}

}  // namespace

template <class M, class N>
void zamr::Solver<M, N>::write_plotfile(const std::string& dir, int step, double time) {
    ensure_aux();
    std::vector<std::string> names;
    for (int n = 0; n < nq; ++n) names.push_back("var" + std::to_string(n));
    for (int n = 0; n < na; ++n) names.push_back("aux" + std::to_string(n));
    const int ncomp = nq + na;
    std::vector<PlotBox> boxes;
    std::vector<double> Q((size_t)nq * nx * ny), A((size_t)na * nx * ny);
    copy_level(0, 0, Q.data(), Q.size(), false);
    copy_level(0, 1, A.data(), A.size(), false);
    for (auto& b : hb) {
        PlotBox pb;
        const int bx = b.hi[0] - b.lo[0] + 1, by = b.hi[1] - b.lo[1] + 1;
        pb.lo[0] = b.lo[0];
        pb.lo[1] = b.lo[1];
        pb.hi[0] = b.hi[0];
        pb.hi[1] = b.hi[1];
        pb.data.resize((size_t)ncomp * bx * by);
        for (int c = 0; c < ncomp; ++c)
            for (int j = 0; j < by; ++j)
                for (int i = 0; i < bx; ++i) {
                    const size_t g = ((size_t)(j + b.lo[1])) * nx + (i + b.lo[0]);
                    const double v = c < nq ? Q[(size_t)c * nx * ny + g] : A[(size_t)(c - nq) * nx * ny + g];
                    pb.data[((size_t)c * by + j) * bx + i] = v;
                }
        boxes.push_back(std::move(pb));
    }
    const std::string path = (dir.empty() ? std::string() : dir + "/") + plotfile_name(identifier, step);
    const double dx[2] = {(phi[0] - plo[0]) / nx, (phi[1] - plo[1]) / ny};
    if (!write_plotfile_level0(path, names, time, plo, phi, nx, ny, dx, step, boxes))
        throw Error(ZAMR_EIO, "cannot write plotfile " + path);
}

namespace {
// Every box of level `lev` as a plot box of Q then Qaux rows.
void plot_boxes(Sol& L, int lev, int ncomp, std::vector<zamr::PlotBox>& boxes, Amr* amr) {
    const int nq = Model::n_dof_q, na = Model::n_dof_qaux;
    std::vector<double> Q((size_t)nq * L.nx * L.ny), A((size_t)na * L.nx * L.ny);
    if (amr) {
        amr->copy_level(lev, 0, Q.data(), Q.size(), false);
        amr->copy_level(lev, 1, A.data(), A.size(), false);
    } else {
        L.copy_level(0, 0, Q.data(), Q.size(), false);
        L.copy_level(0, 1, A.data(), A.size(), false);
    }
    for (auto& b : L.hb) {
        zamr::PlotBox pb;
        const int bx = b.hi[0] - b.lo[0] + 1, by = b.hi[1] - b.lo[1] + 1;
        pb.lo[0] = b.lo[0];
        pb.lo[1] = b.lo[1];
        pb.hi[0] = b.hi[0];
        pb.hi[1] = b.hi[1];
        pb.data.resize((size_t)ncomp * bx * by);
        for (int cc = 0; cc < ncomp; ++cc)
            for (int j = 0; j < by; ++j)
                for (int i = 0; i < bx; ++i) {
                    const size_t g = ((size_t)(j + b.lo[1])) * L.nx + (i + b.lo[0]);
                    const double v = cc < nq ? Q[(size_t)cc * L.nx * L.ny + g]
                                             : A[(size_t)(cc - nq) * L.nx * L.ny + g];
                    pb.data[((size_t)cc * by + j) * bx + i] = v;
                }
        boxes.push_back(std::move(pb));
    }
}
}  // namespace

template <class M, class N>
void zamr::AmrDriver<M, N>::write_plotfile(const std::string& dir, int step, double time) {
    std::vector<std::string> names;
    for (int n = 0; n < nq; ++n) names.push_back("var" + std::to_string(n));
    for (int n = 0; n < na; ++n) names.push_back("aux" + std::to_string(n));
    std::vector<zamr::PlotLevel> lv(finest_level() + 1);
    for (int l = 0; l <= finest_level(); ++l) {
        Sol& L = level(l);
        lv[l].nx = L.nx;
        lv[l].ny = L.ny;
        lv[l].dx[0] = (L.phi[0] - L.plo[0]) / L.nx;
        lv[l].dx[1] = (L.phi[1] - L.plo[1]) / L.ny;
        plot_boxes(L, l, (int)names.size(), lv[l].boxes, this);
    }
    const std::string path = (dir.empty() ? std::string() : dir + "/") + plotfile_name(L0.identifier, step);
    if (!write_plotfile_levels(path, names, time, L0.plo, L0.phi, ref_ratio, step, lv))
        throw Error(ZAMR_EIO, "cannot write plotfile " + path);
}

namespace {
// Device scratch of zamr_compute_dt_noalloc: the group states its CFL
// reduction accumulates into.  Every GPU the library runs on holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
constexpr int kScratchGroups = 4096;
static __device__ __align__(256) unsigned char scratch[kScratchGroups * sizeof(zamr::GroupState)];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_dev[kMaxDevices] = {};

// Address of this library's scratch on the current device, looked up once
// per device.
zamr::GroupState* scratch_here() {
    int dev = -1;
    ZAMR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw zamr::Error(ZAMR_EINVAL, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void* ptr = nullptr;
        ZAMR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = ptr;
    }
    return static_cast<zamr::GroupState*>(scratch_dev[dev]);
}
}  // namespace

extern "C" {

int zamr_model_info(int* nq, int* na, int* np, int* dim, int* ntags) {
    if (nq) *nq = Model::n_dof_q;
    if (na) *na = Model::n_dof_qaux;
    if (np) *np = Model::n_parameters;
    if (dim) *dim = Model::dimension;
    if (ntags) *ntags = (int)Model::get_boundary_tags().size();
    return ZAMR_OK;
}

int zamr_tag_name(int k, char* buf, int cap) { return copy_name(Model::get_boundary_tags(), k, buf, cap); }
int zamr_param_name(int k, char* buf, int cap) { return copy_name(Model::parameter_names(), k, buf, cap); }
int zamr_default_param(int k, double* v) {
    auto d = Model::default_parameters();
    if (k < 0 || k >= (int)d.size() || !v) return ZAMR_EINVAL;
    *v = d[k];
    return ZAMR_OK;
}
int zamr_march_constants(int* d, int* k) {
    zamr_internal_march(d, k);
    return ZAMR_OK;
}

int zamr_create(const char* text, zamr_ctx** out) {
    if (!text || !out) return ZAMR_EINVAL;
    *out = nullptr;
    auto* c = new zamr_ctx();
    zamr_internal_march(&c->s.c_cfl_dimension, &c->s.c_cfl_degree_factor);
    int rc = guard(c, [&] {
        c->s.configure_deck(text);
        if (c->s.max_level >= 1) c->amr = std::make_unique<Amr>(c->s, text);
    });
    if (rc != ZAMR_OK) {
        c->amr.reset();
        // keep the context so the caller can read the message, but nothing else works
        c->s.release();
    }
    *out = c;
    return rc;
}

int zamr_create_boxlist(const zamr_box* boxes, int nbox, const char* text, zamr_ctx** out) {
    if (!boxes || !out) return ZAMR_EINVAL;
    *out = nullptr;
    auto* c = new zamr_ctx();
    zamr_internal_march(&c->s.c_cfl_dimension, &c->s.c_cfl_degree_factor);
    int rc = guard(c, [&] { c->s.configure_boxlist(boxes, nbox, text ? text : ""); });
    if (rc != ZAMR_OK) c->s.release();
    *out = c;
    return rc;
}

int zamr_destroy(zamr_ctx* c) {
    delete c;
    return ZAMR_OK;
}

const char* zamr_last_error(const zamr_ctx* c) { return c ? c->err.c_str() : "null context"; }

int zamr_init_data(zamr_ctx* c) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.init_data();
        if (c->amr) {
            c->amr->init_levels();
            c->amr->reset_clock();
        }
    });
}

int zamr_level_shape(const zamr_ctx* cc, int lev, int* nx, int* ny, int* nboxes) {
    auto* c = const_cast<zamr_ctx*>(cc);
    if (!c) return ZAMR_EINVAL;
    if (c->amr) {
        if (lev < 0 || lev > c->amr->max_level) return ZAMR_EINVAL;
        int r = 1;
        for (int k = 0; k < lev; ++k) r *= c->amr->ref_ratio;
        if (nx) *nx = c->s.nx * r;
        if (ny) *ny = c->s.ny * r;
        if (nboxes) *nboxes = lev <= c->amr->finest_level() ? (int)c->amr->level(lev).hb.size() : 0;
        return ZAMR_OK;
    }
    if (c->s.batched) {
        if (lev < 0 || lev >= (int)c->s.hb.size()) return ZAMR_EINVAL;
        auto& b = c->s.hb[lev];
        if (nx) *nx = b.hi[0] - b.lo[0] + 1;
        if (ny) *ny = b.hi[1] - b.lo[1] + 1;
    } else {
        if (lev != 0) return ZAMR_EINVAL;
        if (nx) *nx = c->s.nx;
        if (ny) *ny = c->s.ny;
    }
    if (nboxes) *nboxes = (int)c->s.hb.size();
    return ZAMR_OK;
}

int zamr_set_state(zamr_ctx* c, int lev, int field, const double* host, size_t n) {
    if (!c || !host || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) c->amr->copy_level(lev, field, const_cast<double*>(host), n, true);
        else c->s.copy_level(lev, field, const_cast<double*>(host), n, true);
    });
}
int zamr_get_state(zamr_ctx* c, int lev, int field, double* host, size_t n) {
    if (!c || !host || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) c->amr->copy_level(lev, field, host, n, false);
        else c->s.copy_level(lev, field, host, n, false);
    });
}

int zamr_compute_dt(zamr_ctx* c, int lev, double* dt) {
    if (!c || !dt || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) {
            *dt = c->amr->level(lev).compute_dt_host();
            return;
        }
        c->s.level_box_check(lev);
        *dt = c->s.compute_dt_host();
    });
}
int zamr_compute_dt_noalloc(zamr_ctx* c, int lev, double* dt_device) {
    if (!c || !dt_device || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        Sol& L = c->amr ? c->amr->level(lev) : c->s;
        if (!c->amr) c->s.level_box_check(lev);
        if (L.ngroups > kScratchGroups)
            throw zamr::Error(ZAMR_EINVAL, "zamr_compute_dt_noalloc: " + std::to_string(L.ngroups) +
                                               " time-step groups exceed the scratch capacity of " +
                                               std::to_string(kScratchGroups));
        L.compute_dt_device(scratch_here(), dt_device);
    });
}
int zamr_update_state(zamr_ctx* c, int lev, double time) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.level_box_check(lev);
        c->s.update_state_host(time);
    });
}
int zamr_fill_patch(zamr_ctx* c, int lev, double time) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.level_box_check(lev);
        (void)time;
        c->s.fill_boundary(c->s.cur, false);
        ZAMR_HIP(hipStreamSynchronize(c->s.stream));
    });
}
int zamr_stage(zamr_ctx* c, int lev, int order, double time, double dt) {
    if (!c || !c->s.d_pool || (order != 1 && order != 2)) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.level_box_check(lev);
        c->s.stage_host(order, time, dt);
    });
}
int zamr_advance(zamr_ctx* c, int lev, double time, double dt) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.level_box_check(lev);
        c->s.advance_host(time, dt);
    });
}
int zamr_time_step(zamr_ctx* c, int lev, double time, int iteration, double* dt_taken) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (c->amr) {
        return guard(c, [&] {
            if (lev != 0) throw zamr::Error(ZAMR_EINVAL, "TimeStep is driven from level 0");
            const double dt = c->amr->time_step(time, iteration);
            if (dt_taken) *dt_taken = dt;
        });
    }
    (void)iteration;
    return guard(c, [&] {
        c->s.level_box_check(lev);
        const double dt = c->s.time_step_host(time);
        if (dt_taken) *dt_taken = dt;
    });
}
int zamr_run_steps(zamr_ctx* c, int nsteps, int* done, double* time) {
    if (!c || !c->s.d_pool || nsteps < 0) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) c->amr->run_steps(nsteps, done, time);
        else c->s.run_steps(nsteps, done, time);
    });
}
int zamr_evolve(zamr_ctx* c) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) c->amr->evolve();
        else c->s.evolve();
    });
}
int zamr_total_mass(zamr_ctx* c, double* m) {
    if (!c || !m || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] { *m = c->amr ? c->amr->total_mass() : c->s.total_mass(); });
}
int zamr_get_time(const zamr_ctx* cc, int lev, double* time, int* step, double* dt) {
    auto* c = const_cast<zamr_ctx*>(cc);
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        Sol& L = c->amr ? c->amr->level(lev) : c->s;
        L.download_groups();
        const int g = c->s.batched ? lev : 0;
        if (g < 0 || g >= L.ngroups) throw zamr::Error(ZAMR_EINVAL, "group index out of range");
        if (time) *time = L.h_groups[g].time;
        if (step) *step = L.h_groups[g].step;
        if (dt) *dt = L.h_groups[g].dt;
    });
}
int zamr_reset_clock(zamr_ctx* c) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) c->amr->reset_clock();
        else c->s.reset_clock();
    });
}
int zamr_write_plotfile(zamr_ctx* c, const char* dir, int step, double time) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->s.batched) throw zamr::Error(ZAMR_EUNSUP, "plotfiles are written for single-level runs only");
        if (c->amr) c->amr->write_plotfile(dir ? dir : "", step, time);
        else c->s.write_plotfile(dir ? dir : "", step, time);
    });
}
int zamr_device_ptr(zamr_ctx* c, int lev, int box, int field, void** ptr, long long* rs, long long* cs, int* ng,
                    void** stream) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        int b = box;
        Sol* L = &c->s;
        if (c->s.batched) {
            b = lev;
        } else if (c->amr) {
            L = &c->amr->level(lev);
        } else if (lev != 0) {
            throw zamr::Error(ZAMR_EINVAL, "only level 0 exists in this build");
        }
        if (b < 0 || b >= (int)L->hb.size()) throw zamr::Error(ZAMR_EINVAL, "box index out of range");
        if (field != 0 && field != 1) throw zamr::Error(ZAMR_EINVAL, "field must be 0 (Q) or 1 (Qaux)");
        auto& B = L->hb[b];
        if (field == 1) L->ensure_aux();
        if (ptr) *ptr = L->d_pool + (field == 0 ? B.off_q[L->cur] : B.off_aux);
        if (rs) *rs = B.gx;
        if (cs) *cs = (long long)B.gx * B.gy;
        if (ng) *ng = B.ng;
        if (stream) *stream = (void*)L->stream;
    });
}
int zamr_set_stream(zamr_ctx* c, void* stream) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        ZAMR_HIP(hipStreamSynchronize(c->s.stream));
        if (c->s.own_stream) hipStreamDestroy(c->s.stream);
        c->s.own_stream = false;
        c->s.stream = (hipStream_t)stream;
        c->s.invalidate_graph();
        if (c->amr) c->amr->for_fine([&](Sol& F) { F.stream = (hipStream_t)stream; });
    });
}
int zamr_set_tuning(zamr_ctx* c, int rows, int per_box, int xcd) {
    if (!c || rows < 0) return ZAMR_EINVAL;
    c->s.rows_per_item = rows;
    c->s.per_box_launch = per_box;
    c->s.xcd_remap = xcd;
    c->s.invalidate_graph();
    if (c->amr)
        c->amr->for_fine([&](Sol& F) {
            F.rows_per_item = rows;
            F.per_box_launch = per_box;
            F.xcd_remap = xcd;
        });
    return ZAMR_OK;
}
int zamr_set_graph(zamr_ctx* c, int on) {
    if (!c) return ZAMR_EINVAL;
    c->s.use_graph = on != 0;
    c->s.invalidate_graph();
    if (c->amr) c->amr->for_fine([&](Sol& F) { F.use_graph = on != 0; });
    return ZAMR_OK;
}
int zamr_set_launch_shape(zamr_ctx* c, const int* class_rows, const int* grid_blocks) {
    if (!c || !class_rows || !grid_blocks) return ZAMR_EINVAL;
    for (int k = 0; k < Sol::kSizeClasses; ++k)
        if (class_rows[k] < 0 || class_rows[k] > 4096) return ZAMR_EINVAL;
    if (grid_blocks[0] < 0 || grid_blocks[1] < 0) return ZAMR_EINVAL;
    if (c->amr) {
        c->err = "launch shapes are set on single-level and box-list contexts; AMR levels keep the analytic rule";
        return ZAMR_EUNSUP;
    }
    for (int k = 0; k < Sol::kSizeClasses; ++k) c->s.class_rows[k] = class_rows[k];
    c->s.grid_blocks[0] = grid_blocks[0];
    c->s.grid_blocks[1] = grid_blocks[1];
    c->s.rows_per_item = 0;
    ++c->s.tune_gen;
    c->s.invalidate_graph();
    return ZAMR_OK;
}
int zamr_autotune(zamr_ctx* c, int reps, int* class_rows, int* grid_blocks, double* ms) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (c->amr) {
        c->err = "the tuner runs on single-level and box-list contexts; AMR levels keep the analytic rule";
        return ZAMR_EUNSUP;
    }
    return guard(c, [&] {
        const auto t = c->s.autotune(reps);
        if (class_rows)
            for (int k = 0; k < Sol::kSizeClasses; ++k) class_rows[k] = c->s.class_rows[k];
        if (grid_blocks) {
            grid_blocks[0] = c->s.grid_blocks[0];
            grid_blocks[1] = c->s.grid_blocks[1];
        }
        if (ms) {
            ms[0] = t.first;
            ms[1] = t.second;
        }
    });
}
int zamr_set_profiling(zamr_ctx* c, int on) {
    if (!c) return ZAMR_EINVAL;
    c->s.profiling = on != 0;
    c->s.invalidate_graph();
    return ZAMR_OK;
}
int zamr_kernel_times(zamr_ctx* c, double* out, int* n) {
    if (!c || !out) return ZAMR_EINVAL;
    const int k = c->s.prof_steps;
    for (int i = 0; i < 3; ++i) out[i] = k ? c->s.prof_ms[i] / k : 0.0;
    if (n) *n = k;
    return ZAMR_OK;
}

int zamr_mood_stats(zamr_ctx* c, int group, int* troubled, long long* redos) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        Sol& L = c->amr ? c->amr->level(group) : c->s;
        const int g = c->amr ? 0 : group;
        if (g < 0 || g >= L.ngroups) throw zamr::Error(ZAMR_EINVAL, "group index out of range");
        L.download_groups();
        if (troubled) *troubled = L.h_groups[g].troubled;
        if (redos) *redos = L.h_groups[g].redos;
    });
}

int zamr_dt_history(zamr_ctx* c, int cap, double* dts, double* times, int* n) {
    if (!c || !c->s.d_pool || cap < 0) return ZAMR_EINVAL;
    return guard(c, [&] {
        c->s.download_groups();
        const int k = std::min({c->s.h_groups[0].step, c->s.hist_cap, cap});
        std::vector<double> h(2 * (size_t)std::max(k, 1));
        if (k > 0) {
            ZAMR_HIP(hipMemcpyAsync(h.data(), c->s.d_hist, 2 * (size_t)k * sizeof(double), hipMemcpyDeviceToHost,
                                    c->s.stream));
            ZAMR_HIP(hipStreamSynchronize(c->s.stream));
        }
        for (int i = 0; i < k; ++i) {
            if (dts) dts[i] = h[2 * i];
            if (times) times[i] = h[2 * i + 1];
        }
        if (n) *n = k;
    });
}

int zamr_finest_level(zamr_ctx* c, int* finest) {
    if (!c || !finest || !c->s.d_pool) return ZAMR_EINVAL;
    *finest = c->amr ? c->amr->finest_level() : 0;
    return ZAMR_OK;
}

int zamr_level_boxes(zamr_ctx* c, int lev, int* lohi, int cap, int* nboxes) {
    if (!c || !c->s.d_pool || cap < 0) return ZAMR_EINVAL;
    return guard(c, [&] {
        const bool exists = c->amr ? lev >= 0 && lev <= c->amr->finest_level() : lev == 0;
        if (!exists) {
            if (c->amr && lev >= 0 && lev <= c->amr->max_level) {
                if (nboxes) *nboxes = 0;
                return;
            }
            throw zamr::Error(ZAMR_EINVAL, "level index out of range");
        }
        Sol& L = c->level(lev);
        const int n = (int)L.hb.size();
        for (int k = 0; k < n && k < cap && lohi; ++k) {
            lohi[4 * k + 0] = L.hb[k].lo[0];
            lohi[4 * k + 1] = L.hb[k].lo[1];
            lohi[4 * k + 2] = L.hb[k].hi[0];
            lohi[4 * k + 3] = L.hb[k].hi[1];
        }
        if (nboxes) *nboxes = n;
    });
}

int zamr_init_levels(zamr_ctx* c) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->amr) {
        c->err = "zamr_init_levels: amr.max_level is 0";
        return ZAMR_EUNSUP;
    }
    return guard(c, [&] {
        c->amr->init_levels();
        c->amr->reset_clock();
    });
}

int zamr_regrid(zamr_ctx* c, double time) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->amr) {
        c->err = "zamr_regrid: amr.max_level is 0";
        return ZAMR_EUNSUP;
    }
    return guard(c, [&] {
        auto& A = *c->amr;
        A.regrid(A.l0_stage_ran ? c->s.last_in : -1, A.t_last_stage, time);
    });
}

int zamr_error_est(zamr_ctx* c, unsigned char* tags, size_t n) {
    if (!c || !tags || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->amr) {
        c->err = "zamr_error_est: amr.max_level is 0";
        return ZAMR_EUNSUP;
    }
    return guard(c, [&] {
        auto& A = *c->amr;
        if (n != (size_t)c->s.nx * c->s.ny) throw zamr::Error(ZAMR_EINVAL, "tag array must hold nx*ny bytes");
        A.error_est_raw(A.l0_stage_ran ? c->s.last_in : -1, A.t_last_stage, tags);
    });
}

int zamr_jfnk_stats(zamr_ctx* c, long long* solves, long long* newton, long long* gmres) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    const auto* j = c->s.jfnk.get();
    if (solves) *solves = j ? j->total_solves : 0;
    if (newton) *newton = j ? j->total_newton : 0;
    if (gmres) *gmres = j ? j->total_gmres : 0;
    return ZAMR_OK;
}

int zamr_amr_stats(zamr_ctx* c, long long* regrids, long long* remakes, double* regrid_ms) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (regrids) *regrids = c->amr ? c->amr->regrids : 0;
    if (remakes) *remakes = c->amr ? c->amr->remakes : 0;
    if (regrid_ms) *regrid_ms = c->amr ? c->amr->regrid_ms : 0.0;
    return ZAMR_OK;
}

#if ZAMR_HAS_ADJOINT
int zamr_adj_setup(zamr_ctx* c, int nsteps, int seg_steps, const double* dts, int ngauges, const int* gi,
                   const int* gj) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    return guard(c, [&] {
        if (c->amr) throw zamr::Error(ZAMR_EUNSUP, "adjoint: single-level runs only (amr.max_level = 0)");
        if (!c->adj) c->adj = std::make_unique<AdjDrv>(c->s);
        c->adj->setup(nsteps, seg_steps, dts, ngauges, gi, gj);
    });
}
int zamr_adj_forward(zamr_ctx* c, double* gauge_h) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->adj) {
        c->err = "adjoint: call zamr_adj_setup first";
        return ZAMR_EINVAL;
    }
    return guard(c, [&] { c->adj->forward(gauge_h); });
}
int zamr_adj_backward(zamr_ctx* c, const double* gauge_hbar) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->adj) {
        c->err = "adjoint: call zamr_adj_setup first";
        return ZAMR_EINVAL;
    }
    return guard(c, [&] { c->adj->backward(gauge_hbar); });
}
int zamr_adj_gradient(zamr_ctx* c, double* qbar0, size_t nq_cells, double* auxbar, size_t ncells) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->adj) {
        c->err = "adjoint: call zamr_adj_setup first";
        return ZAMR_EINVAL;
    }
    return guard(c, [&] { c->adj->gradient(qbar0, nq_cells, auxbar, ncells); });
}
int zamr_adj_device_io(zamr_ctx* c, int on) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->adj) {
        c->err = "adjoint: call zamr_adj_setup first";
        return ZAMR_EINVAL;
    }
    return guard(c, [&] {
        ZAMR_HIP(hipStreamSynchronize(c->s.stream));
        c->adj->device_io = on != 0;
    });
}
int zamr_adj_freeze_limiter(zamr_ctx* c, int on) {
    if (!c || !c->s.d_pool) return ZAMR_EINVAL;
    if (!c->adj) {
        c->err = "adjoint: call zamr_adj_setup first";
        return ZAMR_EINVAL;
    }
    return guard(c, [&] { c->adj->freeze_limiter(on != 0); });
}
#else
static int no_adjoint(zamr_ctx* c) {
    if (c) c->err = "this system has no adjoint (no Adjoint.H beside its Model.H)";
    return ZAMR_EUNSUP;
}
int zamr_adj_setup(zamr_ctx* c, int, int, const double*, int, const int*, const int*) { return no_adjoint(c); }
int zamr_adj_forward(zamr_ctx* c, double*) { return no_adjoint(c); }
int zamr_adj_backward(zamr_ctx* c, const double*) { return no_adjoint(c); }
int zamr_adj_gradient(zamr_ctx* c, double*, size_t, double*, size_t) { return no_adjoint(c); }
int zamr_adj_device_io(zamr_ctx* c, int) { return no_adjoint(c); }
int zamr_adj_freeze_limiter(zamr_ctx* c, int) { return no_adjoint(c); }
#endif

}  // extern "C"
