"""Benchmark: cell-updates per second per RK stage (FP64) of the single-level
2-D shallow-water advance on one MI355X, with the HBM-roofline fraction of
the dominant kernel and the CPU oracle timed beside it.

Workload: 4096^2 cells on [0,1]^2, bathymetry = 32 seeded Gaussian bumps,
eta = 1 + 1e-2 random-phase perturbation, dry where b > eta, Manning
n = 0.03 through the point-implicit source, well-balanced order-2 MUSCL /
Rusanov, SSP-RK2.  A "step" is one SSP-RK2 step (two RK stages).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "zoomy-amrex_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK = 8.0e12  # B/s, MI355X
N_Q = 4            # state rows b, h, hu, hv
GRAPH_CHUNK = 16   # steps per hipGraph the library captures in run_steps (15 with three state buffers)
BYTES_STAGE1 = 16 * N_Q  # read U^n, write U^(1)
BYTES_STAGE2 = 24 * N_Q  # read U^(1) and U^n, write U^(n+1)


def parse_traffic(n_cells):
    """HBM bytes per stage-2 launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE in KiB, corrected by the copy-kernel factor the
    summary states), or None when no summary is present."""
    f = ROOT / "profiles" / "stage2_swe4096_pmc.txt"
    if not f.exists():
        return None
    txt = f.read_text()
    m = re.search(r"stage2_bytes_per_launch\s*=\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def pattern_ceiling(n, t1_ms, t2_ms):
    """The HBM rate the stage kernels' memory access pattern reaches with no
    arithmetic (zamr_probe_pattern: the same loads and stores in the same
    decomposition over an n^2 box of 4 rows) and a plain 1 GiB device copy,
    from the calibration library; None when it is not built."""
    import ctypes
    from zamr import compile as zc
    path = zc.LIBDIR / "libzamr_probe.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    lib.zamr_probe_pattern.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    lib.zamr_probe_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    out = {}
    cells = n * n
    for key, heun, algo in (("stage1", 0, BYTES_STAGE1), ("stage2", 1, BYTES_STAGE2)):
        if lib.zamr_probe_pattern(n, N_Q, heun, 20, ctypes.byref(ms)) != 0:
            return None
        out[key + "_pattern_ms"] = ms.value
        out[key + "_pattern_GBps"] = algo * cells / (ms.value * 1e-3) / 1e9
    if lib.zamr_probe_copy(1 << 30, 20, ctypes.byref(ms)) != 0:
        return None
    out["copy_GBps"] = 2 * (1 << 30) / (ms.value * 1e-3) / 1e9
    out["stage2_of_pattern"] = out["stage2_pattern_ms"] / t2_ms if t2_ms > 0 else 0.0
    out["stage1_of_pattern"] = out["stage1_pattern_ms"] / t1_ms if t1_ms > 0 else 0.0
    out["note"] = ("the stage kernels' loads and stores alone (no arithmetic) and a plain copy: "
                   "stage2_of_pattern = pattern time / stage-2 launch time")
    return out


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, Q0, threads, steps_all=2):
    """The oracle (C++ restatement of the reference, OpenMP over boxes) on the
    same 4096^2 workload cut into 64^2 boxes, as the reference's MPI-over-boxes
    decomposition cuts it (max_grid_size 64): `steps_all` SSP-RK2 steps on
    `threads` threads (every core this process may use, at most 16, the CPU
    share one GPU's job gets), then one step on a single core."""
    from oracle.pyoracle import Oracle
    cells = case["nx"] * case["ny"]

    def leg(th, steps):
        c = dict(case, max_grid_size=64, max_step=steps)
        o = Oracle("swe2d", c, threads=th)
        o.init_solution()
        o.set_state(Q0)
        t0 = time.perf_counter()
        n, _, _, _ = o.evolve()
        el = time.perf_counter() - t0
        del o
        return 2 * n * cells / el, n, el

    v_all, n_all, el_all = leg(threads, steps_all)
    v_one, n_one, el_one = leg(1, 1)
    nb = (case["nx"] // 64) * (case["ny"] // 64)
    return dict(value=v_all, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"{n_all} SSP-RK2 steps (2 stages each) of the same {case['nx']}^2 workload in {nb} boxes "
                       f"of 64^2 on {threads} threads, {el_all:.2f} s",
                one_core=dict(value=v_one, unit="cell-updates/s", cores=1,
                              sample=f"{n_one} SSP-RK2 step of the same workload and boxes on 1 thread, "
                                     f"{el_one:.2f} s"),
                nproc=os.cpu_count(), usable_cores=len(os.sched_getaffinity(0)), cpu_model=cpu_model(),
                build="g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp, OpenMP over boxes")


def bench_boxlist(args):
    """Batched box list: 600 patches of a 3-level hierarchy (sides 16..128) and
    64 dam-break ensemble members (256..1024 cells a side), per-group dt on the
    device.  Four launch modes of the same kernels from the same state:
    one launch per box per pass (eager, the baseline), one launch per pass
    for all boxes eager, the same captured in hipGraphs, and that with the
    measured launch-shape tuner run first."""
    import numpy as np
    import torch
    from zamr import cases
    from zamr.lib import from_boxes

    steps = args.steps
    # the library captures full chunks of GRAPH_CHUNK steps into a hipGraph:
    # the warm-up runs at least one, so no capture falls in the timed region
    warm = max(args.warmup, GRAPH_CHUNK)
    boxes, states, keys = cases.box_list(seed=0, steps=warm + 2 * steps)
    cells = sum((b["hi"][0] + 1) * (b["hi"][1] + 1) for b in boxes)
    res, tuned = {}, None
    for mode in ("per_box", "batched_eager", "batched_graph", "batched_tuned"):
        z = from_boxes(boxes, states, keys)
        if mode == "per_box":
            z.set_tuning(per_box_launch=1)
        if mode == "batched_eager":
            z.set_graph(False)
        if mode == "batched_tuned":
            tuned = z.autotune(reps=10)
        z.run_steps(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done, _ = z.run_steps(steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[mode] = dict(seconds=el, steps=done, rate=2 * done * cells / el)
        res[mode + "_state"] = [z.get_state(lev=k) for k in (0, len(boxes) - 1)]
        del z
    same = all(np.array_equal(a, b) for m in ("batched_eager", "batched_graph", "batched_tuned")
               for a, b in zip(res[m + "_state"], res["per_box_state"]))
    out = {"metric": "batched box-list speed-up over one launch per box per pass",
           "value": res["per_box"]["seconds"] / res["batched_tuned"]["seconds"], "unit": "x",
           "higher_is_better": True, "n_gpus": 1, "steps": steps, "warmup": args.warmup, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "600 AMR patches (3 levels, sides 16/32/64/128) + 64 dam-break members "
                                  "(256..1024)^2, order 2, SSP-RK2, per-group dt", "boxes": len(boxes),
                      "cells": cells},
           "cell_updates_per_s": {m: res[m]["rate"] for m in ("per_box", "batched_eager", "batched_graph",
                                                              "batched_tuned")},
           "speedup_batching_alone": res["per_box"]["seconds"] / res["batched_eager"]["seconds"],
           "speedup_graph_over_eager": res["batched_eager"]["seconds"] / res["batched_graph"]["seconds"],
           "tuner": {"rows_per_size_class": tuned[0], "grid_blocks": tuned[1],
                     "stage1_ms_analytic": tuned[2][0], "stage1_ms_tuned": tuned[2][1]},
           "results_identical": bool(same)}
    print(json.dumps(out), flush=True)


def bench_adjoint(reps=3):
    """Adjoint cost on the 1024^2, 200-step gauge-misfit inversion (per-cell
    Manning system, checkpoint every 20 steps): wall time of the backward
    (segment recompute + reverse stages) and of the checkpointed forward,
    over the wall time of the plain graph-captured forward of the same 200
    steps from the same state.  Host copies of the gauge arrays are inside
    the timed calls; building the input arrays is not.  The PyTorch leg times
    the whole zero-copy GaugeSeries round on CUDA tensors: the initial state
    built from (b, n) in torch, the forward, the misfit and loss.backward()
    down to the gradients of b and n."""
    import numpy as np
    import torch
    from zamr import cases
    from zamr.adjoint import GaugeRun, GaugeSeries, frozen_dts
    from zamr.deck import deck_from_case
    from zamr.lib import ZoomyAmr

    c = cases.swe_adjoint(n=1024, seed=0, steps=200, ngauges=256, nfd=32, seg=20)
    Q0 = cases.initial_state(c["b0"], c["eta0"])
    dts = frozen_dts(c["case"], Q0, c["n0"], 200)
    run = GaugeRun(c["case"], c["gauges"], dts, seg=20)
    truth = run.forward(cases.initial_state(c["b_truth"], c["eta0"]), c["n_truth"])

    plain = ZoomyAmr(deck_from_case(dict(c["case"], max_step=10 ** 9)), system="swe2d_nf", dim=2, order=2)
    plain.init_data()
    aux = plain.get_state(field=1)
    aux[1] = c["n0"]
    dev = torch.device("cuda")
    b_t, n_t, eta_t = (torch.tensor(a, device=dev) for a in (c["b0"], c["n0"], c["eta0"]))
    truth_t = torch.tensor(truth, device=dev)

    def torch_round():
        b = b_t.clone().requires_grad_(True)
        nf = n_t.clone().requires_grad_(True)
        h = torch.clamp(eta_t - b, min=0.0)
        z = torch.zeros_like(h)
        G = GaugeSeries.apply(torch.stack([b, h, z, z]), nf, run)
        loss = 0.5 * ((G - truth_t) ** 2).sum()
        loss.backward()
        return b.grad, nf.grad

    tf, tc, tb, tt = [], [], [], []
    for r in range(reps + 1):
        plain.set_state(Q0)
        plain.set_state(aux, field=1)
        plain.reset_clock()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done, _ = plain.run_steps(200)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        assert done == 200
        run.set_inputs(Q0, c["n0"])
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        G = run.z.adj_forward()
        t3 = time.perf_counter()
        run.z.adj_backward(G - truth)
        t4 = time.perf_counter()
        qb, nb = run.z.adj_gradient()
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        gb_t, gn_t = torch_round()
        torch.cuda.synchronize()
        t6 = time.perf_counter()
        if r:  # the first round loads kernels and builds the graph
            tf.append(t1 - t0)
            tc.append(t3 - t2)
            tb.append(t4 - t3)
            tt.append(t6 - t5)
    assert np.isfinite(qb).all() and np.isfinite(nb).all()
    assert bool(torch.isfinite(gb_t).all()) and bool(torch.isfinite(gn_t).all())
    f, ck, bw, tr = min(tf), min(tc), min(tb), min(tt)
    return {"adjoint_over_forward": bw / f, "gradient_over_forward": (ck + bw) / f,
            "torch_gradient_over_forward": tr / f, "torch_gradient_ms": 1e3 * tr,
            "forward_ms": 1e3 * f, "checkpointed_forward_ms": 1e3 * ck, "backward_ms": 1e3 * bw,
            "workload": "2-D SWE 1024x1024, per-cell Manning, 200 SSP-RK2 steps (frozen dt), 256 gauges, "
                        "checkpoint every 20 steps; gradients w.r.t. the initial state and the Manning field",
            "timing": f"best of {reps} after one warm-up round"}


def bench_sme(steps=50, warmup=3):
    """The level-2 shallow-moment workload at 2048^2 (8 state rows: b, h, three
    moments of each velocity component), slip + Newtonian friction through the
    8x8 point-implicit solve, order 2, no well-balancing: cell-updates/s per RK
    stage and the stage-2 kernel's HBM fraction (algorithmic bytes 16 n_q and
    24 n_q per cell for stages 1 and 2, n_q = 8)."""
    import numpy as np
    import torch
    from zamr import cases
    from zamr.deck import deck_from_case
    from zamr.lib import ZoomyAmr

    n = 2048
    case, Q0 = cases.sme_bumps(n=n, seed=0, steps=warmup + steps)
    z = ZoomyAmr(deck_from_case(case), system="sme2d", dim=2, order=2)
    z.init_data()
    z.set_state(Q0)
    z.run_steps(warmup)
    torch.cuda.synchronize()
    z.set_profiling(True)
    t0 = time.perf_counter()
    done, _ = z.run_steps(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kt, nk = z.kernel_times()
    assert done == steps and np.isfinite(z.get_state()).all()
    cells = n * n
    t1, t2 = kt[0] * 1e-3, kt[1] * 1e-3
    return {"workload": f"level-2 shallow moments {n}x{n}, 16 seeded bumps, moments U[-0.05,0.05] h, slip + "
                        "Newtonian friction point-implicit (8x8), order-2 path-conservative Rusanov, SSP-RK2",
            "value": 2 * steps * cells / el, "unit": "cell-updates/s", "ms_per_step": 1e3 * el / steps,
            "steps": steps, "stage1_avg_launch_ms": kt[0], "stage2_avg_launch_ms": kt[1],
            "stage2_frac": (24 * 8 * cells / t2) / HBM_PEAK if t2 > 0 else 0.0,
            "stage1_frac": (16 * 8 * cells / t1) / HBM_PEAK if t1 > 0 else 0.0}


def bench_amr(steps=10, warmup=2, n=2048):
    """The two-level AMR advance (amr.max_level = 1) on seeded water columns
    over a dry bed, n^2 at level 0 (level 1 in (2n)^2 index space where the
    water is): ms per coarse step (regrid, level-0 step, two level-1
    substeps, average_down), composite cell-updates/s per RK stage (level-0
    cells + 2 x level-1 cells per coarse step), the host share of regrid, and
    the same two level-1 steps on the uniform (2n)^2 grid for comparison."""
    import numpy as np
    import torch
    from zamr import cases
    from zamr.deck import deck_from_case
    from zamr.lib import ZoomyAmr

    amr = dict(max_level=1, ref_ratio=2, blocking_factor=2, n_error_buf=1, grid_eff=0.7, threshold=0.02,
               b_max=1.0e30)
    case, Q0 = cases.dry_columns(n=n, seed=0, steps=warmup + steps + 1)
    z = ZoomyAmr(deck_from_case(case, amr=amr), system="swe2d", dim=2, order=2)
    z.init_data()
    z.set_state(Q0)
    z.init_levels()
    z.run_steps(warmup)
    torch.cuda.synchronize()
    r0, m0, ms0 = z.amr_stats(with_time=True)
    fine_cells, nboxes, lanes = [], [], []
    t_total = 0.0
    for _ in range(steps):
        bx = z.level_boxes(1)
        fine_cells.append(sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in bx))
        nboxes.append(len(bx))
        lanes.append(sum(60 * -(-(b[2] - b[0] + 1) // 60) * (b[3] - b[1] + 1) for b in bx))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done, _ = z.run_steps(1)
        torch.cuda.synchronize()
        t_total += time.perf_counter() - t0
        assert done == 1
    r1, m1, ms1 = z.amr_stats(with_time=True)
    assert np.isfinite(z.get_state()).all()
    upd = 2 * sum(n * n + 2 * f for f in fine_cells)
    frac = float(np.mean(fine_cells)) / (4 * n * n)
    del z
    # the same two fine steps on the uniform grid (its warm-up runs a full
    # graph chunk, so the capture is not timed)
    warm = max(2 * warmup, GRAPH_CHUNK)
    cu, Qu = cases.dry_columns(n=2 * n, seed=0, steps=warm + 2 * steps + 1)
    u = ZoomyAmr(deck_from_case(cu), system="swe2d", dim=2, order=2)
    u.init_data()
    u.set_state(Qu)
    u.run_steps(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u.run_steps(2 * steps)
    torch.cuda.synchronize()
    t_uni = time.perf_counter() - t0
    del u
    return {"workload": f"2-D SWE two-level AMR, {n}x{n} level 0, ref_ratio 2, 24 seeded water columns over a dry "
                        "bed, walls, Manning 0.03 point-implicit, positivity clamp, order 2, regrid every coarse "
                        "step",
            "value": upd / t_total, "unit": "cell-updates/s", "ms_per_coarse_step": 1e3 * t_total / steps,
            "steps": steps, "level1_fraction_of_domain": frac, "regrids": r1 - r0, "remakes": m1 - m0,
            "regrid_host_ms_per_step": (ms1 - ms0) / steps,
            "level1_boxes_mean": float(np.mean(nboxes)),
            "level1_strip_lane_use": float(sum(fine_cells) / max(sum(lanes), 1)),
            "uniform_fine_ms_per_coarse_step": 1e3 * t_uni / steps,
            "speedup_vs_uniform_fine": t_uni / t_total}


def bench_jfnk(steps=10, warmup=2, n=1024):
    """solver.implicit_global on the seeded-bump workload at n^2 with moving
    water (Manning 0.03): ms per SSP-RK2 step with the global Newton + GMRES
    source solve after each stage, against the point-implicit step of the
    same deck, and the iteration counts."""
    import numpy as np
    import torch
    from zamr import cases
    from zamr.deck import deck_from_case
    from zamr.lib import ZoomyAmr

    case, Q0 = cases.swe_bumps(n=n, seed=0, steps=2 * (steps + warmup) + 1)
    Q0 = Q0.copy()
    x = np.linspace(0.0, 1.0, n)
    Q0[2] = 0.3 * Q0[1] * np.sin(2 * np.pi * x)[None, :]
    Q0[3] = -0.2 * Q0[1] * np.cos(2 * np.pi * x)[:, None]
    out = {}
    for label, glob in (("global", True), ("point", False)):
        z = ZoomyAmr(deck_from_case(dict(case, implicit_global=glob)), system="swe2d", dim=2, order=2)
        z.init_data()
        z.set_state(Q0)
        z.run_steps(warmup)
        torch.cuda.synchronize()
        s0 = z.jfnk_stats()
        t0 = time.perf_counter()
        done, _ = z.run_steps(steps)
        torch.cuda.synchronize()
        out[label] = (time.perf_counter() - t0) / done
        if glob:
            s1 = z.jfnk_stats()
            out["solves"], out["newton"], out["gmres"] = (s1[0] - s0[0], s1[1] - s0[1], s1[2] - s0[2])
        assert np.isfinite(z.get_state()).all()
        del z
    return {"workload": f"2-D SWE {n}x{n}, 32 seeded bumps, moving water, Manning 0.03, well-balanced order 2, "
                        "solver.implicit_global = true (Newton + GMRES(30) after each stage)",
            "value": 1e3 * out["global"], "unit": "ms/step", "point_implicit_ms_per_step": 1e3 * out["point"],
            "steps": steps, "newton_per_solve": out["newton"] / max(out["solves"], 1),
            "gmres_per_solve": out["gmres"] / max(out["solves"], 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true", help="run the measured launch-shape tuner before the warm-up")
    ap.add_argument("--workload", default="swe4096", choices=["swe4096", "boxlist", "adjoint", "sme", "amr", "jfnk"],
                    help="swe4096: the single-level 4096^2 workload (default); boxlist: 600 AMR patches + "
                         "64 ensemble members, batched launch vs one launch per box; adjoint: the 1024^2 "
                         "gauge-inversion gradient alone")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg of the default workload")
    ap.add_argument("--no-sme", action="store_true", help="skip the shallow-moment leg of the default workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the run, copy a 1 GiB buffer with a plain device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("this benchmark runs on one GPU only: the box decomposition over several GPUs (one halo "
              "exchange per stage, one max-reduction of the CFL bound per step) is designed but not built",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from zamr import cases
    from zamr.deck import deck_from_case
    from zamr.lib import ZoomyAmr

    if args.workload == "boxlist":
        return bench_boxlist(args)
    if args.workload == "sme":
        print(json.dumps({"metric": "cell-updates/s per RK stage (FP64) + % HBM roofline", "n_gpus": 1,
                          "dtype": "f64", "data": "synthetic", "higher_is_better": True, **bench_sme()}), flush=True)
        return
    if args.workload == "jfnk":
        print(json.dumps({"metric": "ms per SSP-RK2 step with the global implicit-source solve", "n_gpus": 1,
                          "dtype": "f64", "data": "synthetic", "higher_is_better": False, **bench_jfnk()}), flush=True)
        return
    if args.workload == "amr":
        print(json.dumps({"metric": "cell-updates/s per RK stage (FP64), two-level AMR", "n_gpus": 1,
                          "dtype": "f64", "data": "synthetic", "higher_is_better": True, **bench_amr()}), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps({"metric": "adjoint/forward time ratio", "unit": "x", "higher_is_better": False,
                          "n_gpus": 1, "dtype": "f64", "data": "synthetic", **bench_adjoint()}), flush=True)
        return

    case, Q0 = cases.swe_bumps(n=args.n, seed=0, steps=args.warmup + args.steps)
    z = ZoomyAmr(deck_from_case(case), system="swe2d", dim=2, order=2)
    z.init_data()
    z.set_state(Q0)
    if args.rows:
        z.set_tuning(rows=args.rows)
    tuned = z.autotune(reps=10) if args.tune else None
    if args.warmup:
        z.run_steps(args.warmup)
    torch.cuda.synchronize()

    z.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, _ = z.run_steps(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ktimes, nk = z.kernel_times()
    z.set_profiling(False)
    assert done == args.steps, (done, args.steps)
    Q = z.get_state()
    assert np.isfinite(Q).all()

    cells = args.n * args.n
    cu = 2 * args.steps * cells
    value = cu / el
    t2 = ktimes[1] * 1e-3
    t1 = ktimes[0] * 1e-3
    achieved = BYTES_STAGE2 * cells / t2 if t2 > 0 else 0.0
    traffic = parse_traffic(cells)
    out = {
        "metric": "cell-updates/s per RK stage (FP64) + % HBM roofline; adjoint/forward time ratio",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2-D SWE {args.n}x{args.n}, 32 seeded Gaussian bumps, wet/dry, Manning 0.03 "
                               "point-implicit, well-balanced order-2 MUSCL/Rusanov, SSP-RK2",
                   "nx": args.n, "ny": args.n, "steps": args.steps, "seed": 0,
                   "launch_shape": ({"rows_per_size_class": tuned[0], "grid_blocks": tuned[1],
                                     "stage1_ms_analytic": tuned[2][0], "stage1_ms_tuned": tuned[2][1]}
                                    if tuned else "analytic rule")},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": "not measured in this run: stage2_bytes_per_launch of "
                                       "profiles/stage2_swe4096_pmc.txt (rocprofv3 --pmc FETCH_SIZE and "
                                       "WRITE_SIZE of this workload, scaled by the copy-kernel factor "
                                       "stated there)",
                     "kernel": "stage_kernel<swe2d, order 2, stage 2>",
                     "algorithmic_bytes_per_launch": BYTES_STAGE2 * cells,
                     "avg_launch_ms": ktimes[1], "stage1_avg_launch_ms": ktimes[0],
                     "stage1_frac": (BYTES_STAGE1 * cells / t1) / HBM_PEAK if t1 > 0 else 0.0,
                     "step_frac": ((BYTES_STAGE1 + BYTES_STAGE2) * cells / (t1 + t2)) / HBM_PEAK if t1 + t2 > 0
                     else 0.0,
                     "timed_launches": nk},
    }
    pc = pattern_ceiling(args.n, ktimes[0], ktimes[1])
    if pc is not None:
        out["roofline"]["pattern_ceiling"] = pc
    if args.calibrate_copy:
        # a plain fill (writes 1 GiB) and a plain sum (reads 1 GiB)
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        a.fill_(1.0)
        torch.cuda.synchronize()
        s = a.sum()
        torch.cuda.synchronize()
        out["calibration_bytes"] = a.numel() * 8
        out["calibration_sum"] = float(s)
    del z
    if not args.no_adjoint:
        out["adjoint"] = bench_adjoint()
    if not args.no_sme:
        out["sme"] = bench_sme()
    if not args.no_cpu:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = cpu_baseline(case, Q0, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
