#pragma once
// Host driver: the single-level stage/step logic of ZoomyAmr
// (Source/ZoomyAmr.cpp) over device-resident SoA boxes, with dt, time and the
// loop condition kept on the device so a step is four kernel launches that
// can be captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "zamr_deck.hpp"
#include "zamr_error.hpp"
#include "zamr_jfnk.hpp"
#include "zamr_kernels.hpp"

namespace zamr {

struct HostBox {
    int lo[2], hi[2];
    int ng;
    int dlo[2], dhi[2];
    double dx[2], plo[2];
    int group;
    int bc[4];
    size_t off_q[3], off_aux;  // element offsets into the pool (a third state buffer with MOOD)
    int gx, gy;                // grown extents
};

// AMReX BoxList::maxSize chopping (max_grid_size): ceil(L/m) chunks per axis,
// the first L % chunks of them one cell longer.
inline std::vector<std::pair<int, int>> chop_axis(int L, int m) {
    std::vector<std::pair<int, int>> r;
    const int nb = (L + m - 1) / m;
    const int sz = L / nb, extra = L % nb;
    int s = 0;
    for (int k = 0; k < nb; ++k) {
        const int l = sz + (k < extra ? 1 : 0);
        r.push_back({s, s + l - 1});
        s += l;
    }
    return r;
}


// Pinned host memory that a regrid's small uploads are copied through, so
// that each is queued on the stream without waiting (a copy from pageable
// memory returns only once the stream has reached it, draining the device
// first).  Reset at the start of each regrid, when every copy queued from it
// has completed; a request that does not fit is copied from the caller's
// memory instead.
struct HostStage {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    explicit HostStage(size_t bytes) {
        ZAMR_HIP(hipHostMalloc(&base, bytes));
        cap = bytes;
    }
    ~HostStage() {
        if (base) hipHostFree(base);
    }
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    void reset() { used = 0; }
    void upload(void* dst, const void* src, size_t n, hipStream_t s) {
        if (!n) return;
        const size_t off = (used + 255) & ~size_t(255);
        if (off + n <= cap) {
            std::memcpy(base + off, src, n);
            used = off + n;
            ZAMR_HIP(hipMemcpyAsync(dst, base + off, n, hipMemcpyHostToDevice, s));
        } else {
            ZAMR_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
            ZAMR_HIP(hipStreamSynchronize(s));
        }
    }
};

// Device buffers that a retired solver hands to the next one, one per role:
// an AMR level is remade every few coarse steps, and once the buffers have
// grown to what the level needs a remake allocates and frees nothing.  Used
// from one host thread; the levels share a stream, so a buffer's last kernels
// for the old level run before its first for the new one.
struct BufferCache {
    enum Role { POOL, BOXES, GROUPS, HIST, RECTS, ITEMS0, ITEMS1, ITEMS2, PARTIAL, NROLES };
    void* p[NROLES] = {};
    size_t bytes[NROLES] = {};
    // the role's buffer if it holds at least `need` bytes (its size in got)
    void* take(int r, size_t need, size_t& got) {
        if (!p[r] || bytes[r] < need) return nullptr;
        void* q = p[r];
        got = bytes[r];
        p[r] = nullptr;
        bytes[r] = 0;
        return q;
    }
    // keep the larger of the held and the given buffer (the other is freed
    // once the device is idle: launches already queued may still use it)
    void give(int r, void* q, size_t b) {
        if (!q) return;
        if (p[r] && bytes[r] >= b) {
            hipDeviceSynchronize();
            hipFree(q);
            return;
        }
        if (p[r]) {
            hipDeviceSynchronize();
            hipFree(p[r]);
        }
        p[r] = q;
        bytes[r] = b;
    }
    ~BufferCache() {
        for (int r = 0; r < NROLES; ++r)
            if (p[r]) hipFree(p[r]);
    }
};

template <class M, class N>
class Solver {
   public:
    using S = Sys<M>;
    static constexpr int nq = M::n_dof_q;
    static constexpr int na = M::n_dof_qaux;
    static constexpr int np = M::n_parameters;
    static constexpr int dim = M::dimension;

    // --- ZoomyAmr members (ZoomyAmr.H:93-159) ---
    int nx = 0, ny = 0, max_grid_size = 64, max_level = 0;
    double plo[2] = {0, 0}, phi[2] = {1, 1};
    double time_end = 1.0;
    int max_step = -1;
    double cfl = 0.5, dtmin = 1.0e-7, dtmax = 1.0e-2;
    int spatial_order = 1;
    bool implicit_source = true, implicit_global = false, well_balanced = false, clamp_positivity = true;
    std::string positivity = "none";
    int face_based_rhs = 1;
    double plot_dt_interval = 0.1;
    int identifier = 0;
    std::string bc_name[4];
    std::string dem_file, release_file, friction_file;
    std::vector<std::string> state_rasters;
    Params prm{};
    int c_cfl_dimension = 1, c_cfl_degree_factor = 1;

    bool batched = false;
    std::vector<HostBox> hb;
    std::unique_ptr<Jfnk<M>> jfnk;  // solver.implicit_global: the global source solve
    int ngroups = 1;

    // --- device state ---
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double* d_pool = nullptr;
    size_t pool_elems = 0;
    size_t pool_cap = 0;             // elements allocated (>= pool_elems)
    BufferCache* cache = nullptr;    // where device buffers come from and go back to (AMR level 1)
    HostStage* stage = nullptr;      // pinned staging of uploads (AMR level 1; else direct copies)
    size_t cap_bytes[BufferCache::NROLES] = {};  // size of each held buffer
    double pool_headroom = 1.0;      // a new pool holds this many times what the boxes need
    BoxDesc* d_boxes = nullptr;
    GroupState* d_groups = nullptr;
    std::vector<GroupState> h_groups;
    CopyRect* d_rects = nullptr;
    int nrects = 0;
    long long rect_cells = 0;  // ghost cells over all copy rectangles
    double* d_hist = nullptr;
    int hist_cap = 1 << 16;
    double* d_partial = nullptr;
    int partial_cap = 0;
    int cur = 0;                    // buffer holding the current state
    int last_in = -1;               // input buffer of the last stage run (its ghosts hold that
                                    // stage's FillPatch), -1 before the first
    bool maxinv_fresh = false;      // group maxinv_bits already hold this state's reduction
    bool aux_stale = false;         // stored Qaux lags the last UpdateState (steps ran since)
    int rows_per_item = 0;          // > 0: rows per work item in every box; 0: by size class
    int per_box_launch = 0;
    int xcd_remap = 1;
    // Stage-kernel launch shape, per box size class (the larger side of a
    // box: <= 32, 64, 128, 256, 512, 1024, 2048, larger): rows per work item
    // (0 = the analytic rule of rows_for), and the grid of the first-stage
    // and of the step-ending stage kernels (0 = one block per slot of four
    // items; n > 0 = n blocks walking the slots).  autotune() fills them
    // from measurements; tune_gen counts the changes.
    static constexpr int kSizeClasses = 8;
    int class_rows[kSizeClasses] = {};
    int grid_blocks[2] = {0, 0};
    int tune_gen = 0;
    bool use_graph = true;          // capture chunks of steps into hipGraphs

    struct Items {
        int hx = -1, rows = -1, gen = -1;
        WorkItem* d = nullptr;
        int n = 0;
        std::vector<int> box_begin;  // item range per box (per-box launch mode)
    };
    Items items_[3];

    // A captured chunk of steps names its state buffers by index, so it is
    // valid only while a replay starts from the buffer the capture started
    // from (graph_cur); replaying it leaves cur there and last_in as the
    // capture left it (graph_last_in).
    hipGraphExec_t graph_exec = nullptr;
    int graph_key = -1;
    int graph_cur = -1, graph_last_in = -1;

    bool profiling = false;
    std::vector<hipEvent_t> ev;
    double prof_ms[3] = {0, 0, 0};
    int prof_steps = 0;

    int plot_step = 0;
    double next_plot_time = 0.0;

    ~Solver() { release(); }

    // Host-to-device copy on the stream, through the pinned stage when set.
    void upload(void* dst, const void* src, size_t n) {
        if (stage) stage->upload(dst, src, n, stream);
        else ZAMR_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, stream));
    }

    // A device buffer of at least n elements for `role` in ptr: kept if it is
    // large enough, else taken from the cache or allocated (headroom times n;
    // a cached level's small buffers get twice n, so that the next level's
    // usually fit).
    template <class T>
    void dev_get(int role, T*& ptr, size_t n, double headroom = 1.0) {
        if (cache && role != BufferCache::POOL) headroom = std::max(headroom, 2.0);
        const size_t need = std::max<size_t>(n, 1) * sizeof(T);
        if (ptr && cap_bytes[role] >= need) return;
        if (ptr && !cache) ZAMR_HIP(hipStreamSynchronize(stream));  // queued launches may still use it
        dev_put(role, ptr);
        size_t got = 0;
        if (cache)
            if (void* q = cache->take(role, need, got)) {
                ptr = static_cast<T*>(q);
                cap_bytes[role] = got;
                return;
            }
        const size_t want = std::max(need, (size_t)((double)need * headroom) / sizeof(T) * sizeof(T));
        ZAMR_HIP(hipMalloc(&ptr, want));
        cap_bytes[role] = want;
    }
    template <class T>
    void dev_put(int role, T*& ptr) {
        if (!ptr) return;
        if (cache) cache->give(role, ptr, cap_bytes[role]);
        else hipFree(ptr);
        ptr = nullptr;
        cap_bytes[role] = 0;
    }

    void release() {
        jfnk.reset();
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        for (int k = 0; k < 3; ++k) dev_put(BufferCache::ITEMS0 + k, items_[k].d);
        dev_put(BufferCache::POOL, d_pool);
        dev_put(BufferCache::BOXES, d_boxes);
        dev_put(BufferCache::GROUPS, d_groups);
        dev_put(BufferCache::RECTS, d_rects);
        dev_put(BufferCache::HIST, d_hist);
        dev_put(BufferCache::PARTIAL, d_partial);
        for (auto e : ev) hipEventDestroy(e);
        ev.clear();
        if (own_stream && stream) hipStreamDestroy(stream);
        stream = nullptr;
    }

    // ---------------- configuration ----------------
    static int tag_index(const std::string& name) {  // ZoomyAmr.cpp:361-372
        if (name.empty()) return 0;
        auto tags = M::get_boundary_tags();
        for (int t = 0; t < (int)tags.size(); ++t)
            if (tags[t] == name) return t;
        std::string known;
        for (auto const& t : tags) known += (known.empty() ? "" : ", ") + t;
        throw Error(2, "boundary tag '" + name + "' is not declared by this model. Known tags: [" + known +
                           "]. Check the bc.* inputs keys against the model's boundary_conditions tags.");
    }

    void read_solver_keys(const Deck& pp) {
        pp.query("output.identifier", identifier);
        pp.query("output.plot_dt_interval", plot_dt_interval);
        pp.query("solver.time_end", time_end);
        pp.query("solver.max_step", max_step);
        pp.query("solver.cfl", cfl);
        pp.query("solver.dtmin", dtmin);
        pp.query("solver.dtmax", dtmax);
        pp.query("solver.spatial_order", spatial_order);
        pp.query("solver.implicit_source", implicit_source);
        pp.query("solver.implicit_global", implicit_global);
        pp.query("solver.well_balanced", well_balanced);
        pp.query("solver.clamp_positivity", clamp_positivity);
        pp.query("solver.positivity", positivity);
        pp.query("solver.face_based_rhs", face_based_rhs);
        pp.query("solver.profile_split", face_based_rhs);
        if (spatial_order != 1 && spatial_order != 2) throw Error(1, "solver.spatial_order must be 1 or 2");
        if (positivity != "none" && positivity != "mood") throw Error(1, "solver.positivity must be none or mood");
        auto defs = M::default_parameters();
        for (int i = 0; i < np; ++i) prm.v[i] = defs[i];
        auto names = M::parameter_names();
        for (int i = 0; i < np; ++i) {
            double v;
            if (pp.query("params." + names[i], v)) prm.v[i] = v;
        }
    }

    void configure_deck(const std::string& text) {
        Deck pp(text);
        read_solver_keys(pp);
        pp.query("amr.max_level", max_level);
        if (max_level < 0 || max_level > 3)
            throw Error(5, "amr.max_level > 3: this build's hierarchy has at most four levels (0 to 3)");
        if (max_level >= 1) {
            int rr = 2;
            pp.query("amr.ref_ratio", rr);
            if (rr != 2 && rr != 4) throw Error(5, "amr.ref_ratio must be 2 or 4 in this build");
        }
        std::vector<double> nc;
        if (!pp.queryarr("amr.n_cell", nc) || nc.size() < 2) throw Error(1, "amr.n_cell needs at least 2 values");
        nx = (int)nc[0];
        ny = (int)nc[1];
        pp.query("amr.max_grid_size", max_grid_size);
        std::vector<double> lo, hi;
        if (pp.queryarr("geometry.prob_lo", lo) && lo.size() >= 2) plo[0] = lo[0], plo[1] = lo[1];
        if (pp.queryarr("geometry.prob_hi", hi) && hi.size() >= 2) phi[0] = hi[0], phi[1] = hi[1];
        pp.query("bc.x_lo", bc_name[0]);
        pp.query("bc.x_hi", bc_name[1]);
        pp.query("bc.y_lo", bc_name[2]);
        pp.query("bc.y_hi", bc_name[3]);
        pp.query("init.dem_file", dem_file);
        pp.query("init.release_file", release_file);
        pp.query("init.friction_file", friction_file);
        pp.queryarr("init.state_rasters", state_rasters);
        if (nx <= 0 || ny <= 0 || max_grid_size <= 0) throw Error(1, "bad amr.n_cell / amr.max_grid_size");
        int bcidx[4];
        for (int k = 0; k < 4; ++k) bcidx[k] = tag_index(bc_name[k]);
        const double dx[2] = {(phi[0] - plo[0]) / nx, (phi[1] - plo[1]) / ny};
        const int ng = (spatial_order == 2) ? 2 : 1;
        for (auto& y : chop_axis(ny, max_grid_size))
            for (auto& x : chop_axis(nx, max_grid_size)) {
                HostBox b{};
                b.lo[0] = x.first;
                b.hi[0] = x.second;
                b.lo[1] = y.first;
                b.hi[1] = y.second;
                b.ng = 2;  // storage always carries 2 ghosts so either order can run
                b.dlo[0] = 0;
                b.dhi[0] = nx - 1;
                b.dlo[1] = 0;
                b.dhi[1] = ny - 1;
                b.dx[0] = dx[0];
                b.dx[1] = dx[1];
                b.plo[0] = plo[0];
                b.plo[1] = plo[1];
                b.group = 0;
                for (int k = 0; k < 4; ++k) b.bc[k] = bcidx[k];
                if (b.hi[0] - b.lo[0] + 1 < ng || (dim == 2 && b.hi[1] - b.lo[1] + 1 < ng))
                    throw Error(1, "boxes narrower than the ghost width are not supported");
                hb.push_back(b);
            }
        ngroups = 1;
        allocate();
    }

    // A refined level (ref_ratio `ratio` over level c): the given boxes of
    // valid cells in this level's index space, every solver key and the
    // stream taken from c.  It keeps its own clock; only level 0 stops at
    // time_end (ZoomyAmr.cpp:746-749).
    void configure_fine(const Solver& c, int ratio, const std::vector<std::array<int, 4>>& boxes) {
        spatial_order = c.spatial_order;
        implicit_source = c.implicit_source;
        implicit_global = c.implicit_global;
        well_balanced = c.well_balanced;
        clamp_positivity = c.clamp_positivity;
        positivity = c.positivity;
        face_based_rhs = c.face_based_rhs;
        cfl = c.cfl;
        dtmin = c.dtmin;
        dtmax = c.dtmax;
        prm = c.prm;
        c_cfl_dimension = c.c_cfl_dimension;
        c_cfl_degree_factor = c.c_cfl_degree_factor;
        identifier = c.identifier;
        for (int k = 0; k < 4; ++k) bc_name[k] = c.bc_name[k];
        time_end = 1.0e300;
        max_step = -1;
        nx = c.nx * ratio;
        ny = c.ny * ratio;
        plo[0] = c.plo[0];
        plo[1] = c.plo[1];
        phi[0] = c.phi[0];
        phi[1] = c.phi[1];
        rows_per_item = c.rows_per_item;
        xcd_remap = c.xcd_remap;
        per_box_launch = c.per_box_launch;
        stream = c.stream;
        own_stream = false;
        int bcidx[4];
        for (int k = 0; k < 4; ++k) bcidx[k] = tag_index(bc_name[k]);
        const double dx[2] = {(phi[0] - plo[0]) / nx, (phi[1] - plo[1]) / ny};
        for (auto& q : boxes) {
            HostBox b{};
            b.lo[0] = q[0];
            b.lo[1] = q[1];
            b.hi[0] = q[2];
            b.hi[1] = q[3];
            b.ng = 2;
            b.dlo[0] = 0;
            b.dhi[0] = nx - 1;
            b.dlo[1] = 0;
            b.dhi[1] = ny - 1;
            b.dx[0] = dx[0];
            b.dx[1] = dx[1];
            b.plo[0] = plo[0];
            b.plo[1] = plo[1];
            b.group = 0;
            for (int k = 0; k < 4; ++k) b.bc[k] = bcidx[k];
            if (b.hi[0] - b.lo[0] + 1 < 2 || b.hi[1] - b.lo[1] + 1 < 2)
                throw Error(1, "refined boxes narrower than the ghost width are not supported");
            hb.push_back(b);
        }
        ngroups = 1;
        allocate();
    }

    void configure_boxlist(const zamr_box* boxes, int nbox, const std::string& text) {
        Deck pp(text);
        read_solver_keys(pp);
        batched = true;
        if (nbox <= 0) throw Error(1, "empty box list");
        int gmax = 0;
        for (int k = 0; k < nbox; ++k) {
            const zamr_box& z = boxes[k];
            HostBox b{};
            for (int d = 0; d < 2; ++d) {
                b.lo[d] = z.lo[d];
                b.hi[d] = z.hi[d];
                b.dlo[d] = z.lo[d];
                b.dhi[d] = z.hi[d];
                b.dx[d] = z.dx[d];
                b.plo[d] = z.prob_lo[d] - z.lo[d] * z.dx[d];  // cell lo sits at prob_lo
            }
            b.ng = 2;
            b.group = z.group;
            if (z.group < 0) throw Error(1, "negative group id");
            gmax = std::max(gmax, z.group);
            for (int s = 0; s < 4; ++s) b.bc[s] = tag_index(z.bc[s] ? std::string(z.bc[s]) : std::string());
            const int ng = (spatial_order == 2) ? 2 : 1;
            if (b.hi[0] - b.lo[0] + 1 < ng || (dim == 2 && b.hi[1] - b.lo[1] + 1 < ng))
                throw Error(1, "boxes narrower than the ghost width are not supported");
            hb.push_back(b);
        }
        ngroups = gmax + 1;
        nx = hb[0].hi[0] - hb[0].lo[0] + 1;
        ny = hb[0].hi[1] - hb[0].lo[1] + 1;
        allocate();
    }

    void allocate() {
        if (!stream) {
            ZAMR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        size_t off = 0;
        auto align = [](size_t x) { return (x + 31) & ~size_t(31); };  // 256-byte alignment
        for (auto& b : hb) {
            b.gx = b.hi[0] - b.lo[0] + 1 + 2 * b.ng;
            b.gy = (dim == 2 || true) ? b.hi[1] - b.lo[1] + 1 + 2 * b.ng : 1;
            const size_t cells = (size_t)b.gx * b.gy;
            b.off_q[0] = off;
            off = align(off + cells * nq);
            b.off_q[1] = off;
            off = align(off + cells * nq);
            b.off_q[2] = b.off_q[1];
            if (nbuf() == 3) {  // Q^n survives the step (MOOD redo, stage-2 solve, AMR regrid overlap)
                b.off_q[2] = off;
                off = align(off + cells * nq);
            }
            b.off_aux = off;
            off = align(off + cells * std::max(na, 1));
        }
        pool_elems = off;
        dev_get(BufferCache::POOL, d_pool, pool_elems, pool_headroom);
        pool_cap = cap_bytes[BufferCache::POOL] / sizeof(double);
        ZAMR_HIP(hipMemsetAsync(d_pool, 0, pool_elems * sizeof(double), stream));
        std::vector<BoxDesc> bd(hb.size());
        for (size_t k = 0; k < hb.size(); ++k)
            bd[k] = desc_of(hb[k], d_pool + hb[k].off_q[0], d_pool + hb[k].off_q[1], d_pool + hb[k].off_q[2]);
        dev_get(BufferCache::BOXES, d_boxes, bd.size());
        upload(d_boxes, bd.data(), bd.size() * sizeof(BoxDesc));
        h_groups.assign(ngroups, GroupState{});
        for (auto& g : h_groups) {
            g.dt = 0.0;
            g.time = 0.0;
            g.time_end = time_end;
            g.cfl = cfl;
            g.dtmin = dtmin;
            g.dtmax = dtmax;
            g.divisor = double(c_cfl_dimension) * double(c_cfl_degree_factor);
            g.maxinv_bits = 0ull;
            g.step = 0;
            g.max_step = max_step;
            g.active = 1;
        }
        dev_get(BufferCache::GROUPS, d_groups, (size_t)ngroups);
        upload_groups();
        dev_get(BufferCache::HIST, d_hist, 2 * (size_t)hist_cap);
        build_rects();
        if (jfnk_mode()) {
            if (batched) throw Error(5, "solver.implicit_global = true is not part of the batched box list");
            std::vector<std::array<int, 4>> bx;
            for (auto& b : hb) bx.push_back({b.lo[0], b.lo[1], b.hi[0], b.hi[1]});
            jfnk = std::make_unique<Jfnk<M>>();
            jfnk->setup(stream, d_boxes, bx);
        }
        // the work items of the stage order exist from creation on, so that
        // entry points that may not allocate find them
        items(hx_of(spatial_order));
        if (!stage) ZAMR_HIP(hipStreamSynchronize(stream));
    }

    // Device descriptor of box b with the given state buffers.
    BoxDesc desc_of(const HostBox& b, double* q0, double* q1, double* q2 = nullptr) const {
        BoxDesc d;
        d.q[0] = q0;
        d.q[1] = q1;
        d.q[2] = q2 ? q2 : q1;
        d.aux = d_pool + b.off_aux;
        for (int i = 0; i < 2; ++i) {
            d.lo[i] = b.lo[i];
            d.hi[i] = b.hi[i];
            d.gl[i] = b.lo[i] - b.ng;
            d.dlo[i] = b.dlo[i];
            d.dhi[i] = b.dhi[i];
            d.dx[i] = b.dx[i];
            d.plo[i] = b.plo[i];
        }
        d.sy = b.gx;
        d.sc = (long long)b.gx * b.gy;
        d.group = b.group;
        for (int s = 0; s < 4; ++s) d.bc[s] = b.bc[s];
        return d;
    }

    void upload_groups() {
        if (stage) return upload(d_groups, h_groups.data(), ngroups * sizeof(GroupState));
        ZAMR_HIP(hipMemcpyAsync(d_groups, h_groups.data(), ngroups * sizeof(GroupState), hipMemcpyHostToDevice,
                                stream));
    }
    void download_groups() {
        ZAMR_HIP(hipMemcpyAsync(h_groups.data(), d_groups, ngroups * sizeof(GroupState), hipMemcpyDeviceToHost,
                                stream));
        ZAMR_HIP(hipStreamSynchronize(stream));
    }

    // Ghost cells of box b that lie inside the domain and inside another box
    // (same-level FillBoundary).  Batched boxes are independent domains.
    void build_rects() {
        rect_cells = 0;
        if (batched || hb.size() < 2) return;
        std::vector<CopyRect> rr;
        for (size_t b = 0; b < hb.size(); ++b)
            for (size_t o = 0; o < hb.size(); ++o) {
                if (o == b) continue;
                const HostBox& B = hb[b];
                const HostBox& O = hb[o];
                const int gy = (dim == 2) ? B.ng : 0;
                const int ilo = std::max(B.lo[0] - B.ng, O.lo[0]), ihi = std::min(B.hi[0] + B.ng, O.hi[0]);
                const int jlo = std::max(B.lo[1] - gy, O.lo[1]), jhi = std::min(B.hi[1] + gy, O.hi[1]);
                if (ilo > ihi || jlo > jhi) continue;
                rr.push_back(CopyRect{(int)o, (int)b, ilo, jlo, ihi - ilo + 1, jhi - jlo + 1, rect_cells});
                rect_cells += (long long)(ihi - ilo + 1) * (jhi - jlo + 1);
            }
        nrects = (int)rr.size();
        if (nrects) {
            dev_get(BufferCache::RECTS, d_rects, rr.size());
            upload(d_rects, rr.data(), rr.size() * sizeof(CopyRect));
        }
    }

    // Waves per SIMD the order's stage kernel runs at (its register use
    // decides: two for the 4-row systems, one for the 8-row ones), and the
    // device's CU count; both queried once.
    mutable int occ_cache[3] = {0, 0, 0};
    mutable int cu_count = 0;
    int stage_waves_per_simd(int hx) const {
        if (!occ_cache[hx]) {
            int blocks = 0;
            const void* k = hx == 2 ? (const void*)stage_kernel<M, N, 2, 1, 0, 1> : (const void*)stage_kernel<M, N, 1, 1, 0, 1>;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, NT, 0) != hipSuccess || blocks < 1) blocks = 2;
            occ_cache[hx] = blocks;  // a block is one wave on each of a CU's 4 SIMDs
        }
        if (!cu_count) {
            int dev = 0, n = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
                n = 256;
            cu_count = n;
        }
        return occ_cache[hx];
    }

    // Rows per work item.  Deeper items re-read fewer halo rows, shallower
    // ones put more waves on the SIMDs.  At two waves per SIMD: the deepest
    // of 32, 16, 8, 4 rows that still gives every SIMD two waves over the
    // whole box list (measured: 32 rows at 4096^2, 8 rows at 1024^2).  At one
    // wave per SIMD the launch runs in rounds of one item per SIMD, and a
    // last round that fills few SIMDs costs a whole item time: the depth
    // (8..64) that minimises rounds x (rows + 2), the 2 for the item's
    // prologue (measured at 2048^2 with 8 rows: 36 rows, 35 x 57 items in two
    // rounds, 1.25x the rate of 32 rows in three).
    int rows_for(int hx) const {
        if (rows_per_item > 0) return rows_per_item;
        const int txc = WV - 2 * hx;
        auto waves = [&](int rows) {
            long long w = 0;
            for (const HostBox& b : hb)
                w += (long long)((b.hi[0] - b.lo[0] + txc) / txc) * ((b.hi[1] - b.lo[1] + rows) / rows);
            return w;
        };
        if (stage_waves_per_simd(hx) == 1) {
            const long long slots = 4LL * cu_count;
            int best = 32;
            long long best_cost = -1;
            for (int rows = 8; rows <= 64; ++rows) {
                const long long cost = ((waves(rows) + slots - 1) / slots) * (rows + 2);
                if (best_cost < 0 || cost < best_cost) {
                    best_cost = cost;
                    best = rows;
                }
            }
            return best;
        }
        for (int rows : {32, 16, 8})
            if (waves(rows) >= 2 * 1024) return rows;
        return 4;
    }

    static int size_class(const HostBox& b) {
        const int side = std::max(b.hi[0] - b.lo[0] + 1, b.hi[1] - b.lo[1] + 1);
        int k = 0;
        while (k + 1 < kSizeClasses && side > (32 << k)) ++k;
        return k;
    }

    // Blocks of a stage launch over n items (EPI 0: a first stage; 1, 2: a
    // stage that ends the step).
    int grid_for(int n, int epi) const {
        const int s = slots_for(n), g = grid_blocks[epi == 0 ? 0 : 1];
        return (g > 0 && g < s) ? g : s;
    }

    // Strips of NT-2*hx columns by segments of rows, box by box: the rows of
    // the box's size class, or rows_for's one value.
    Items& items(int hx) {
        Items& it = items_[hx];
        const int rule = (dim == 2) ? rows_for(hx) : 1;
        if (it.d && it.hx == hx && it.rows == rule && it.gen == tune_gen) return it;
        std::vector<WorkItem> v;
        it.box_begin.clear();
        const int txc = WV - 2 * hx;  // one wave64 per item, HX halo columns each side
        for (size_t b = 0; b < hb.size(); ++b) {
            it.box_begin.push_back((int)v.size());
            const HostBox& B = hb[b];
            int rows = rule;
            if (dim == 2 && rows_per_item <= 0 && class_rows[size_class(B)] > 0) rows = class_rows[size_class(B)];
            for (int j0 = B.lo[1]; j0 <= B.hi[1]; j0 += rows)
                for (int i0 = B.lo[0]; i0 <= B.hi[0]; i0 += txc)
                    v.push_back(WorkItem{(int)b, i0, j0, std::min(rows, B.hi[1] - j0 + 1)});
        }
        it.box_begin.push_back((int)v.size());
        it.hx = hx;
        it.rows = rule;
        it.gen = tune_gen;
        it.n = (int)v.size();
        dev_get(BufferCache::ITEMS0 + hx, it.d, v.size());
        upload(it.d, v.data(), v.size() * sizeof(WorkItem));
        dev_get(BufferCache::PARTIAL, d_partial, (size_t)it.n);
        partial_cap = (int)(cap_bytes[BufferCache::PARTIAL] / sizeof(double));
        invalidate_graph();
        return it;
    }

    void invalidate_graph() {
        if (graph_exec) hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        graph_key = -1;
        graph_cur = graph_last_in = -1;
    }

    // ---------------- state I/O ----------------
    HostBox& level_box_check(int lev) {
        if (batched) {
            if (lev < 0 || lev >= (int)hb.size()) throw Error(1, "box index out of range");
            return hb[lev];
        }
        if (lev != 0) throw Error(1, "only level 0 exists in this build");
        return hb[0];
    }

    // Dense (ncomp, ny, nx) array of level 0 (or of box `lev` in batched mode).
    void copy_level(int lev, int field, double* host, size_t n, bool to_device) {
        level_box_check(lev);
        const int nc = (field == 0) ? nq : na;
        if (field != 0 && field != 1) throw Error(1, "field must be 0 (Q) or 1 (Qaux)");
        int gx0 = 0, gy0 = 0, NX = nx, NY = ny;
        std::vector<int> which;
        if (batched) {
            which.push_back(lev);
            gx0 = hb[lev].lo[0];
            gy0 = hb[lev].lo[1];
            NX = hb[lev].hi[0] - hb[lev].lo[0] + 1;
            NY = hb[lev].hi[1] - hb[lev].lo[1] + 1;
        } else {
            for (size_t b = 0; b < hb.size(); ++b) which.push_back((int)b);
        }
        if (n != (size_t)nc * NX * NY) throw Error(1, "array size mismatch: expected ncomp*ny*nx doubles");
        // Qaux is brought up to date before it is read, and before Q is
        // replaced (it then keeps what the last UpdateState left, as in the reference).
        if (field == 1 || to_device) ensure_aux();
        for (int b : which) {
            const HostBox& B = hb[b];
            const size_t cells = (size_t)B.gx * B.gy;
            std::vector<double> tmp(cells * nc, 0.0);
            double* dptr = d_pool + ((field == 0) ? B.off_q[cur] : B.off_aux);
            ZAMR_HIP(hipMemcpyAsync(tmp.data(), dptr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
            ZAMR_HIP(hipStreamSynchronize(stream));
            for (int c = 0; c < nc; ++c)
                for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                    for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                        const size_t li = (size_t)(i - B.lo[0] + B.ng) + (size_t)(j - B.lo[1] + B.ng) * B.gx +
                                          (size_t)c * cells;
                        const size_t gi = ((size_t)c * NY + (j - gy0)) * NX + (i - gx0);
                        if (to_device) tmp[li] = host[gi];
                        else host[gi] = tmp[li];
                    }
            if (to_device) {
                ZAMR_HIP(hipMemcpyAsync(dptr, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, stream));
                ZAMR_HIP(hipStreamSynchronize(stream));
            }
        }
        if (to_device) maxinv_fresh = false;
    }

    // init_solution (init_solution.cpp:78-115) on Q and Qaux: zero every
    // component, then b = 0 and h = 2 left of the domain's x midpoint, 1 right.
    void init_solution() {
        aux_stale = false;
        for (size_t b = 0; b < hb.size(); ++b) {
            const HostBox& B = hb[b];
            const double xmid = batched ? B.plo[0] + 0.5 * (B.dlo[0] + B.dhi[0] + 1) * B.dx[0]
                                        : 0.5 * (plo[0] + phi[0]);
            const size_t cells = (size_t)B.gx * B.gy;
            for (int field = 0; field < 2; ++field) {
                const int nc = field == 0 ? nq : std::max(na, 1);
                std::vector<double> tmp(cells * nc, 0.0);
                if (nc >= 2)
                    for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                        for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                            const double x = B.plo[0] + (i + 0.5) * B.dx[0];
                            const size_t li = (size_t)(i - B.lo[0] + B.ng) + (size_t)(j - B.lo[1] + B.ng) * B.gx;
                            tmp[li] = 0.0;
                            tmp[li + cells] = (x < xmid) ? 2.0 : 1.0;
                        }
                double* dptr = d_pool + (field == 0 ? B.off_q[cur] : B.off_aux);
                ZAMR_HIP(hipMemcpyAsync(dptr, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, stream));
                ZAMR_HIP(hipStreamSynchronize(stream));
            }
        }
        maxinv_fresh = false;
    }

    // readRasterIntoComponent (init_solution.cpp:7-75): raw float64, idx = i + nx*j.
    void read_raster(const std::string& file, int field, int comp) {
        std::ifstream ifs(file, std::ios::binary);
        if (!ifs.good()) throw Error(3, "### Cannot open raw file: " + file);
        ifs.seekg(0, std::ios::end);
        const size_t actual = (size_t)ifs.tellg();
        ifs.seekg(0);
        const size_t need = (size_t)nx * ny * sizeof(double);
        if (actual < need)
            throw Error(3, "Raw file size mismatch: " + file + " has " + std::to_string(actual) + " < expected " +
                               std::to_string(need) + " bytes");
        std::vector<double> r((size_t)nx * ny);
        ifs.read(reinterpret_cast<char*>(r.data()), need);
        const int nc = field == 0 ? nq : na;
        std::vector<double> all((size_t)nc * nx * ny);
        copy_level(0, field, all.data(), all.size(), false);
        std::copy(r.begin(), r.end(), all.begin() + (size_t)comp * nx * ny);
        copy_level(0, field, all.data(), all.size(), true);
    }

    static bool exists(const std::string& f) {
        std::ifstream i(f);
        return i.good();
    }

    void init_data() {  // ZoomyAmr::InitData (ZoomyAmr.cpp:112-156), single level
        init_solution();
        if (!batched) {
            for (int c = 0; c < (int)state_rasters.size() && c < nq; ++c)
                if (exists(state_rasters[c])) read_raster(state_rasters[c], 0, c);
            if (!dem_file.empty() && exists(dem_file)) read_raster(dem_file, 0, 0);
            if (!release_file.empty() && exists(release_file)) read_raster(release_file, 0, 1);
            if (!friction_file.empty() && exists(friction_file)) read_raster(friction_file, 1, 1);
        }
        reset_clock();
    }

    void reset_clock() {
        download_groups();
        for (auto& g : h_groups) {
            g.time = 0.0;
            g.step = 0;
            g.dt = 0.0;
            g.active = 1;
            g.maxinv_bits = 0ull;
        }
        upload_groups();
        maxinv_fresh = false;
        plot_step = 0;
        next_plot_time = 0.0;
    }

    bool mood() const { return positivity == "mood" && spatial_order == 2; }
    // The source is solved globally after each stage (ZoomyAmr.cpp:647-651).
    bool jfnk_mode() const { return implicit_source && implicit_global; }
    // An order-2 level 0 under AMR keeps Q^n in its own buffer through the
    // step too: the regrid that runs beside the level-0 step fills a new fine
    // level from it.
    bool amr_base() const { return max_level >= 1 && spatial_order == 2; }
    int nbuf() const { return (mood() || jfnk_mode() || amr_base()) ? 3 : 2; }

    // ---------------- kernels ----------------
    int hx_of(int order) const { return order == 2 ? 2 : 1; }
    static int aux_grid(int nitems) { return (nitems + NT / WV - 1) / (NT / WV); }

    StageArgs stage_args(int src, int dst, int old) const {
        StageArgs a;
        a.p = prm;
        a.src = src;
        a.dst = dst;
        a.old = old;
        a.wb = well_balanced ? 1 : 0;
        a.impl = implicit_source ? (implicit_global ? 2 : 1) : 0;
        a.clamp = (clamp_positivity && positivity != "mood") ? 1 : 0;
        a.mood = 0;
        a.pre = nullptr;
        a.dstp = nullptr;
        a.sw = nullptr;
        return a;
    }

    int clamp_flag() const { return (clamp_positivity && positivity != "mood") ? 1 : 0; }

    // FillPatch's same-level part before a stage (update: ghosts take the
    // owners' values after UpdateState) or a plain copy of the stored state.
    void fill_boundary(int buf, bool update = true) {
        if (nrects == 0) return;
        hipLaunchKernelGGL(fill_boundary_kernel<M>, dim3((unsigned)((rect_cells + 255) / 256)), dim3(256), 0, stream,
                           d_boxes, d_rects, nrects, rect_cells, buf, prm, clamp_flag(), update ? 1 : 0);
        ZAMR_HIP(hipGetLastError());
    }

    template <int ORDER, int EPI>
    void launch_stage(const StageArgs& a) {
        Items& it = items(hx_of(ORDER));
        auto kern = a.wb ? stage_kernel<M, N, ORDER, EPI, 0, 1> : stage_kernel<M, N, ORDER, EPI, 0, 0>;
        if (per_box_launch) {
            for (size_t b = 0; b + 1 < it.box_begin.size(); ++b) {
                const int n0 = it.box_begin[b], n = it.box_begin[b + 1] - n0;
                if (n <= 0) continue;
                hipLaunchKernelGGL(kern, dim3(slots_for(n)), dim3(NT), 0, stream, d_boxes, it.d + n0, n, a, d_groups, 0);
            }
        } else {
            hipLaunchKernelGGL(kern, dim3(grid_for(it.n, EPI)), dim3(NT), 0, stream, d_boxes, it.d, it.n, a, d_groups,
                               xcd_remap);
        }
        ZAMR_HIP(hipGetLastError());
    }

    void launch_stage_dyn(int order, int epi, const StageArgs& a) {
        if (order == 2) {
            if (epi == 0) launch_stage<2, 0>(a);
            else if (epi == 1) launch_stage<2, 1>(a);
            else launch_stage<2, 2>(a);
        } else {
            if (epi == 0) launch_stage<1, 0>(a);
            else if (epi == 1) launch_stage<1, 1>(a);
            else launch_stage<1, 2>(a);
        }
    }

    void begin_step() {
        hipLaunchKernelGGL(begin_step_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, d_groups, ngroups);
        ZAMR_HIP(hipGetLastError());
    }
    void end_step() {
        hipLaunchKernelGGL(end_step_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, d_groups, ngroups, d_hist,
                           hist_cap);
        ZAMR_HIP(hipGetLastError());
    }

    // Group maxima from the state and aux as stored (ComputeDt's reduction).
    void reduce_stored() {
        ensure_aux();
        download_groups();
        for (auto& g : h_groups) g.maxinv_bits = 0ull;
        upload_groups();
        Items& it = items(hx_of(spatial_order));
        hipLaunchKernelGGL((max_inv_kernel<M, N>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d, it.n,
                           cur, prm, d_groups, WV - 2 * it.hx);
        ZAMR_HIP(hipGetLastError());
        maxinv_fresh = true;
    }

    void ensure_aux() {
        if (aux_stale) materialize_aux();
    }

    void materialize_aux() {
        Items& it = items(hx_of(spatial_order));
        download_groups();
        hipLaunchKernelGGL((materialize_aux_kernel<M>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d,
                           it.n, cur, prm, h_groups[0].time, WV - 2 * it.hx, (const GroupState*)nullptr);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipStreamSynchronize(stream));
        aux_stale = false;
    }

    // The stage kernels of Advance for an order-`order` step on the device
    // clock (dt and time read from the group state).
    // End of a step whose stages solved the source globally (Heun with
    // buffer old, or none at order 1; UpdateState; CFL max) in place on dst.
    void finish_step(int dst, int old, int heun_only = 0) {
        Items& it = items(hx_of(spatial_order));
        hipLaunchKernelGGL((finish_step_kernel<M, N>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d, it.n,
                           dst, old, prm, clamp_flag(), d_groups, WV - 2 * it.hx, heun_only);
        ZAMR_HIP(hipGetLastError());
    }

    // Advance with solver.implicit_global: each stage is the flux update
    // followed by the global source solve on its output; the host drives
    // Newton/GMRES, so dt and time come down once per step.
    void advance_global_source() {
        aux_stale = true;
        download_groups();
        const GroupState& g = h_groups[0];
        if (!g.active) return;
        const double dt = g.dt, time = g.time;
        const int t1 = (cur + 1) % 3, t2 = (cur + 2) % 3;
        if (spatial_order == 1) {
            fill_boundary(cur);
            launch_stage_dyn(1, 0, stage_args(cur, t1, cur));
            jfnk->solve(t1, dt, time, prm);
            finish_step(t1, -1);
            last_in = cur;
            cur = t1;
        } else {
            fill_boundary(cur);
            launch_stage_dyn(2, 0, stage_args(cur, t1, cur));
            jfnk->solve(t1, dt, time, prm);
            fill_boundary(t1);
            launch_stage_dyn(2, 0, stage_args(t1, t2, cur));
            jfnk->solve(t2, dt, time, prm);
            if (mood()) {
                // the order-2 candidate (Heun) and its troubled cells; with any,
                // the whole step again at order 1 from Q^n (ZoomyAmr.cpp:699-720),
                // each stage with its global solve
                finish_step(t2, cur, 1);
                hipLaunchKernelGGL(mood_reset_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, d_groups,
                                   ngroups);
                ZAMR_HIP(hipGetLastError());
                download_groups();
                if (h_groups[0].troubled > 0) {
                    launch_stage_dyn(1, 0, stage_args(cur, t1, cur));
                    jfnk->solve(t1, dt, time, prm);
                    fill_boundary(t1);
                    launch_stage_dyn(1, 0, stage_args(t1, t2, cur));
                    jfnk->solve(t2, dt, time, prm);
                    finish_step(t2, cur);
                } else {
                    finish_step(t2, -1);
                }
            } else {
                finish_step(t2, cur);
            }
            last_in = t1;
            cur = t2;
        }
    }

    void advance_kernels() {
        if (jfnk_mode()) {
            advance_global_source();
            return;
        }
        aux_stale = true;
        if (spatial_order == 1) {
            fill_boundary(cur);
            launch_stage_dyn(1, 2, stage_args(cur, 1 - cur, cur));
            last_in = cur;
            cur = 1 - cur;
        } else if (mood()) {
            // the order-2 candidate goes to a third buffer so that Q^n stays for
            // an order-1 redo of the groups whose candidate has troubled cells
            // (ZoomyAmr.cpp:699-720); the buffers rotate by one per step
            const int t1 = (cur + 1) % 3, t2 = (cur + 2) % 3;
            fill_boundary(cur);
            record(0);
            launch_stage_dyn(2, 0, stage_args(cur, t1, cur));
            record(1);
            fill_boundary(t1);
            StageArgs a2 = stage_args(t1, t2, cur);
            a2.mood = 1;
            launch_stage_dyn(2, 1, a2);
            record(2);
            hipLaunchKernelGGL(mood_reset_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, d_groups, ngroups);
            ZAMR_HIP(hipGetLastError());
            StageArgs r1 = stage_args(cur, t1, cur);
            r1.mood = 2;
            launch_stage_dyn(1, 0, r1);
            fill_boundary(t1);
            StageArgs r2 = stage_args(t1, t2, cur);
            r2.mood = 2;
            launch_stage_dyn(1, 1, r2);
            last_in = t1;
            cur = t2;
        } else if (nbuf() == 3) {  // the same stages, Q^{n+1} into a third buffer
            const int t1 = (cur + 1) % 3, t2 = (cur + 2) % 3;
            fill_boundary(cur);
            record(0);
            launch_stage_dyn(2, 0, stage_args(cur, t1, cur));
            record(1);
            fill_boundary(t1);
            launch_stage_dyn(2, 1, stage_args(t1, t2, cur));
            record(2);
            last_in = t1;
            cur = t2;
        } else {
            fill_boundary(cur);
            record(0);
            launch_stage_dyn(2, 0, stage_args(cur, 1 - cur, cur));
            record(1);
            fill_boundary(1 - cur);
            launch_stage_dyn(2, 1, stage_args(1 - cur, cur, cur));
            record(2);
            last_in = 1 - cur;
        }
    }

    // Profiling: four events per step (before begin_step, stage 1, stage 2,
    // after stage 2) recorded on the stream without any synchronisation; the
    // elapsed times are read after the run.
    int prof_slot = -1;
    void record(int k) {
        if (!profiling || prof_slot < 0) return;
        const size_t idx = (size_t)prof_slot * 4 + k;
        while (ev.size() <= idx) {
            hipEvent_t e;
            ZAMR_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        ZAMR_HIP(hipEventRecord(ev[idx], stream));
    }

    // Graph replay for the device-clock loop; the per-box launch mode stays
    // eager, as the baseline it stands for (one launch per box per pass).
    bool graph_ok() const {
        // (the legacy null stream, which a caller may hand over, cannot be captured)
        return use_graph && stream != nullptr && spatial_order == 2 && !profiling && !per_box_launch && !jfnk_mode();
    }

    void run_steps(int nsteps, int* done, double* time) {
        download_groups();
        const int step0 = h_groups[0].step;
        if (!maxinv_fresh) reduce_stored();
        prof_ms[0] = prof_ms[1] = prof_ms[2] = 0.0;
        prof_steps = 0;
        int k = 0;
        // with three state buffers the rotation repeats every 3 steps, so a
        // captured chunk is a multiple of 3 steps
        const int chunk = profiling ? std::max(nsteps, 1) : (nbuf() == 3 ? 15 : 16);
        if (mood()) items(1);  // the redo's work items exist before any capture
        while (k < nsteps) {
            const int n = std::min(chunk, nsteps - k);
            if (graph_ok() && n == chunk) {
                // eager steps of a three-buffer rotation, a single stage or an
                // Advance move cur; a graph captured from another buffer is stale
                if (!graph_exec || graph_key != chunk || graph_cur != cur) {
                    invalidate_graph();
                    items(hx_of(spatial_order));
                    const int cur0 = cur;
                    hipGraph_t g;
                    ZAMR_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                    for (int s = 0; s < chunk; ++s) {
                        begin_step();
                        advance_kernels();
                        end_step();
                    }
                    ZAMR_HIP(hipStreamEndCapture(stream, &g));
                    ZAMR_HIP(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0));
                    hipGraphDestroy(g);
                    if (cur != cur0) throw Error(4, "internal: a captured chunk must return to its first buffer");
                    graph_key = chunk;
                    graph_cur = cur0;
                    graph_last_in = last_in;
                }
                ZAMR_HIP(hipGraphLaunch(graph_exec, stream));
                cur = graph_cur;
                last_in = graph_last_in;
            } else {
                for (int s = 0; s < n; ++s) {
                    if (profiling) prof_slot = k + s;
                    record(3);
                    begin_step();
                    advance_kernels();
                    end_step();
                }
                prof_slot = -1;
            }
            k += n;
            download_groups();
            bool any = false;
            for (auto& g : h_groups) any = any || ((g.time < g.time_end) && (g.max_step < 0 || g.step < g.max_step));
            if (!any) break;
        }
        maxinv_fresh = true;
        download_groups();
        if (profiling && spatial_order == 2) {
            const int n = std::min(k, nsteps);
            for (int i = 0; i < n; ++i) {
                float a = 0, b = 0, c = 0;
                ZAMR_HIP(hipEventElapsedTime(&a, ev[4 * i + 0], ev[4 * i + 1]));
                ZAMR_HIP(hipEventElapsedTime(&b, ev[4 * i + 1], ev[4 * i + 2]));
                ZAMR_HIP(hipEventElapsedTime(&c, ev[4 * i + 3], ev[4 * i + 0]));
                prof_ms[0] += a;
                prof_ms[1] += b;
                prof_ms[2] += c;
            }
            prof_steps = n;
        }
        if (done) *done = h_groups[0].step - step0;
        if (time) *time = h_groups[0].time;
    }

    // Measured launch-shape tuner.  The stage kernels of one step are timed
    // on the current state with HIP events (one untimed launch, then the
    // shortest of `reps` launches per candidate): first under the analytic
    // rule, then with rows per work item in {4, 8, 16, 32, 64} for each box
    // size class present, one class at a time from the one with the most
    // cells (cost: both kernels), then the grid of each kernel (one block
    // per slot, or 1x, 2x, 4x the blocks the device holds at once, walking
    // the slots); each keeps its fastest candidate, which has to beat the
    // one kept by 1%.  The first stage writes the buffer
    // the next step's first stage overwrites, the step-ending stage a
    // buffer no state lives in (a temporary one when a box has two), so the
    // state is left as it was; the group CFL maxima are recomputed before
    // the next step.  The cut into items changes no result (every face is
    // evaluated by the same code whichever item holds it).  Returns the
    // per-step stage-kernel time (ms) of the analytic shape and of the
    // chosen one.
    std::pair<double, double> autotune(int reps) {
        if (dim != 2) throw Error(1, "the tuner cuts 2-D boxes into row segments; 1-D items are single rows");
        if (jfnk_mode()) throw Error(5, "the tuner times the point-implicit stages; not with solver.implicit_global");
        reps = std::max(reps, 1);
        const int hx = hx_of(spatial_order);
        const bool two = spatial_order == 2;
        const int d1 = (nbuf() == 2) ? 1 - cur : (cur + 1) % 3;
        int d2 = (cur + 2) % 3;
        double* tmp = nullptr;
        if (two && nbuf() == 2) {
            std::vector<size_t> off(hb.size());
            size_t total = 0;
            for (size_t k = 0; k < hb.size(); ++k) {
                off[k] = total;
                total = (total + (size_t)hb[k].gx * hb[k].gy * nq + 31) & ~size_t(31);  // 256-byte alignment
            }
            ZAMR_HIP(hipMalloc(&tmp, total * sizeof(double)));
            std::vector<BoxDesc> bd(hb.size());
            for (size_t k = 0; k < hb.size(); ++k)
                bd[k] = desc_of(hb[k], d_pool + hb[k].off_q[0], d_pool + hb[k].off_q[1], tmp + off[k]);
            upload(d_boxes, bd.data(), bd.size() * sizeof(BoxDesc));
            ZAMR_HIP(hipStreamSynchronize(stream));
            d2 = 2;
        }
        const int epi_first = two ? 0 : 2, epi_last = two ? 1 : 2;
        const StageArgs a1 = stage_args(cur, d1, cur);
        const StageArgs a2 = stage_args(d1, d2, cur);
        rows_per_item = 0;
        for (int& r : class_rows) r = 0;
        grid_blocks[0] = grid_blocks[1] = 0;
        ++tune_gen;
        invalidate_graph();
        // a candidate's time: the shortest of `reps` launches, each between
        // two events (the minimum is robust to clock ramps and to other work
        // on the device; a candidate must beat the kept one by 1%)
        std::vector<hipEvent_t> e(reps + 1);
        for (auto& x : e) ZAMR_HIP(hipEventCreate(&x));
        auto time_kernel = [&](int epi, const StageArgs& a) {
            items(hx);
            launch_stage_dyn(spatial_order, epi, a);
            for (int r = 0; r < reps; ++r) {
                ZAMR_HIP(hipEventRecord(e[r], stream));
                launch_stage_dyn(spatial_order, epi, a);
            }
            ZAMR_HIP(hipEventRecord(e[reps], stream));
            ZAMR_HIP(hipEventSynchronize(e[reps]));
            double best_ms = 1e30;
            for (int r = 0; r < reps; ++r) {
                float ms = 0.0f;
                ZAMR_HIP(hipEventElapsedTime(&ms, e[r], e[r + 1]));
                best_ms = std::min(best_ms, (double)ms);
            }
            return best_ms;
        };
        auto cost = [&]() { return time_kernel(epi_first, a1) + (two ? time_kernel(epi_last, a2) : 0.0); };
        fill_boundary(cur);
        if (two) {  // the step-ending stage reads the first stage's output with its ghosts filled
            items(hx);
            launch_stage_dyn(spatial_order, epi_first, a1);
            fill_boundary(d1);
        }
        for (int w = 0; w < 4; ++w) cost();  // clocks up before the first candidate
        const double before = cost();
        double best = before;
        long long cells[kSizeClasses] = {};
        int tallest[kSizeClasses] = {};
        for (const HostBox& b : hb) {
            const int k = size_class(b), h = b.hi[1] - b.lo[1] + 1;
            cells[k] += (long long)(b.hi[0] - b.lo[0] + 1) * h;
            tallest[k] = std::max(tallest[k], h);
        }
        int order[kSizeClasses];
        for (int k = 0; k < kSizeClasses; ++k) order[k] = k;
        std::stable_sort(order, order + kSizeClasses, [&](int x, int y) { return cells[x] > cells[y]; });
        for (int k : order) {
            if (!cells[k]) continue;
            int keep = 0;
            for (int rows : {4, 8, 16, 32, 64}) {
                class_rows[k] = rows;
                ++tune_gen;
                const double t = cost();
                if (t < 0.99 * best) {
                    best = t;
                    keep = rows;
                }
                if (rows >= tallest[k]) break;  // deeper items cut this class no differently
            }
            class_rows[k] = keep;
            ++tune_gen;
        }
        const int resident = stage_waves_per_simd(hx) * cu_count;
        for (int kind = 0; kind < (two ? 2 : 1); ++kind) {
            const int epi = kind == 0 ? epi_first : epi_last;
            const StageArgs& a = kind == 0 ? a1 : a2;
            const int slot = two ? kind : 1;
            double tb = time_kernel(epi, a);
            int keep = 0;
            for (int g : {resident, 2 * resident, 4 * resident}) {
                grid_blocks[slot] = g;
                const double t = time_kernel(epi, a);
                if (t < 0.99 * tb) {
                    tb = t;
                    keep = g;
                }
            }
            grid_blocks[slot] = keep;
        }
        if (!two) grid_blocks[0] = grid_blocks[1];
        const double after = cost();
        for (auto& x : e) hipEventDestroy(x);
        if (tmp) {
            std::vector<BoxDesc> bd(hb.size());
            for (size_t k = 0; k < hb.size(); ++k)
                bd[k] = desc_of(hb[k], d_pool + hb[k].off_q[0], d_pool + hb[k].off_q[1], d_pool + hb[k].off_q[2]);
            upload(d_boxes, bd.data(), bd.size() * sizeof(BoxDesc));
            ZAMR_HIP(hipStreamSynchronize(stream));
            ZAMR_HIP(hipFree(tmp));
        }
        maxinv_fresh = false;
        invalidate_graph();
        return {before, after};
    }

    // ---------------- reference-shaped entry points ----------------
    double compute_dt_host() {  // ZoomyAmr::ComputeDt
        reduce_stored();
        download_groups();
        const double m = [&] {
            double mm = 0.0;
            for (auto& g : h_groups) {
                double v;
                std::memcpy(&v, &g.maxinv_bits, sizeof(v));
                mm = std::max(mm, v);
            }
            return mm;
        }();
        maxinv_fresh = false;
        if (m < 1e-14) return dtmax;
        const double div = double(c_cfl_dimension) * double(c_cfl_degree_factor);
        const double dt = cfl / (div * m);
        return std::max(std::min(dt, dtmax), dtmin);
    }

    // ComputeDt as compute_dt_host computes it, queued on the stream with no
    // allocation and no host synchronisation: the stored aux brought up to
    // date at each group's time, the CFL maxima reduced into the ngroups group
    // states at `scratch` (device memory of at least ngroups GroupStates,
    // cleared here), the step size written to the device double `out`.
    void compute_dt_device(GroupState* scratch, double* out) {
        const int hx = hx_of(spatial_order);
        Items& it = items_[hx];
        if (!it.d || it.hx != hx) throw Error(1, "the context's work items are not set up");
        if (aux_stale) {
            hipLaunchKernelGGL((materialize_aux_kernel<M>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d,
                               it.n, cur, prm, 0.0, WV - 2 * it.hx, (const GroupState*)d_groups);
            ZAMR_HIP(hipGetLastError());
            aux_stale = false;
        }
        hipLaunchKernelGGL(clear_maxinv_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, scratch, ngroups);
        ZAMR_HIP(hipGetLastError());
        hipLaunchKernelGGL((max_inv_kernel<M, N>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d, it.n,
                           cur, prm, scratch, WV - 2 * it.hx);
        ZAMR_HIP(hipGetLastError());
        hipLaunchKernelGGL(dt_from_maxinv_kernel, dim3(1), dim3(64), 0, stream, (const GroupState*)scratch, ngroups,
                           cfl, double(c_cfl_dimension) * double(c_cfl_degree_factor), dtmin, dtmax, out);
        ZAMR_HIP(hipGetLastError());
    }

    void set_clock(double dt, double time) {
        hipLaunchKernelGGL(set_dt_kernel, dim3((ngroups + 63) / 64), dim3(64), 0, stream, d_groups, ngroups, dt, time);
        ZAMR_HIP(hipGetLastError());
    }

    void advance_host(double time, double dt) {  // ZoomyAmr::Advance with the caller's dt
        set_clock(dt, time);
        advance_kernels();
        materialize_aux();
        maxinv_fresh = false;
        download_groups();
    }

    void stage_host(int order, double time, double dt) {  // do_stage(order) in place
        aux_stale = true;
        materialize_aux();  // do_stage's UpdateState leaves Qaux = update_aux(stage-start state)
        set_clock(dt, time);
        fill_boundary(cur);
        const int nxt = (cur + 1) % nbuf();
        launch_stage_dyn(order, 0, stage_args(cur, nxt, cur));
        if (jfnk_mode()) jfnk->solve(nxt, dt, time, prm);
        last_in = cur;
        cur = nxt;
        ZAMR_HIP(hipStreamSynchronize(stream));
        maxinv_fresh = false;
    }

    void update_state_host(double time) {
        Items& it = items(hx_of(spatial_order));
        hipLaunchKernelGGL((update_state_kernel<M>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d, it.n,
                           cur, prm, time, (clamp_positivity && positivity != "mood") ? 1 : 0, WV - 2 * it.hx);
        ZAMR_HIP(hipGetLastError());
        ZAMR_HIP(hipStreamSynchronize(stream));
        maxinv_fresh = false;
    }

    // ComputeTotalMass of this level (deterministic item-order sum), without
    // the cells flagged in `cover` (device, one byte per cell, row length nx).
    double total_mass(const unsigned char* cover = nullptr) {
        Items& it = items(hx_of(spatial_order));
        hipLaunchKernelGGL((mass_partial_kernel<M>), dim3(aux_grid(it.n)), dim3(NT), 0, stream, d_boxes, it.d, it.n,
                           cur, d_partial, WV - 2 * it.hx, cover, nx);
        ZAMR_HIP(hipGetLastError());
        std::vector<double> part(it.n);
        ZAMR_HIP(hipMemcpyAsync(part.data(), d_partial, it.n * sizeof(double), hipMemcpyDeviceToHost, stream));
        ZAMR_HIP(hipStreamSynchronize(stream));
        double s = 0.0;
        for (double v : part) s += v;
        return s;
    }

    double time_step_host(double time) {  // ZoomyAmr::TimeStep, level 0
        double dt = compute_dt_host();
        const double remaining = time_end - time;
        if (remaining > 0.0 && dt > remaining) dt = remaining;
        advance_host(time, dt);
        return dt;
    }

    void write_plotfile(const std::string& dir, int step, double time);

    void evolve() {  // ZoomyAmr::Evolve
        reset_clock();
        double time = 0.0;
        int step = 0;
        next_plot_time = 0.0;
        const auto t0 = std::chrono::steady_clock::now();
        const double mass0 = total_mass();
        std::cout << "Initial total mass = " << std::setprecision(12) << mass0 << "\n";
        while (time < time_end && (max_step < 0 || step < max_step)) {
            if (time >= next_plot_time) {
                write_plotfile("", plot_step, time);
                next_plot_time += plot_dt_interval;
                plot_step++;
            }
            int done = 0;
            double tnew = time;
            run_steps(1, &done, &tnew);
            if (done == 0) break;
            if (mood() && h_groups[0].troubled > 0)  // ZoomyAmr.cpp:712-713
                std::cout << "  [MOOD] lev 0 troubled cells: " << (long)h_groups[0].troubled
                          << " -> whole-step order-1 redo\n";
            const double dt0 = h_groups[0].dt;
            time = tnew;
            step++;
            const double mass = total_mass();
            const double drift = (mass0 != 0.0) ? (mass - mass0) / mass0 : (mass - mass0);
            std::cout << "Step " << step << " dt: " << std::setprecision(12) << dt0 << " time: "
                      << std::setprecision(12) << time << "s"
                      << " levels: " << 1 << " mass: " << std::setprecision(14) << mass
                      << " drift: " << std::setprecision(4) << drift << "\n";
        }
        materialize_aux();
        write_plotfile("", plot_step, time);
        const double evo = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Total evolution time = " << evo << " seconds\n";
        std::cout.flush();
    }
};

}  // namespace zamr
