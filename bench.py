#!/usr/bin/env python3
"""Benchmark: wall seconds per timestep at 8192^2, relative residual 1e-10.

Flagship: FP32 inner CG with FP64 residual replacement, warm-started from
the learned U-Net when trained weights are present (falling back per step
to the previous-step guess when the prediction's residual is worse).
Baseline: plain FP64 CG from the previous-step guess (the reference's
u = energy*density start).  Both run on the same seeded random-rectangle
problem (background rho=100, e=1e-4 plus 32 rectangles; dt 0.004) and must
reach ||u0 - A u|| <= 1e-10 ||u0|| in FP64 every step.

A step is one timestep of diffuse!: halo(energy, density), the solve,
solvefinished! and the field summary, all on the device, with the fields
resident in HBM.  W untimed warmup steps, then K timed steps per method.

Prints one JSON line (rank 0).  `roofline` is for the flagship's dominant
kernel (largest total time among the FP32 CG kernels), timed per launch with
HIP events on the library's stream during the timed steps; `cpu_baseline`
times the CPU oracle's FP64 CG on the host for a bounded number of
iterations and extrapolates to the GPU's iteration count.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tealeaf.jl_amd"))

import numpy as np  # noqa: E402

import tealeaf_hip as T  # noqa: E402
from tealeaf_hip import _lib, decks  # noqa: E402

PEAK_HBM_GBS = 8000.0
PROFILE_PERIOD = 15  # odd: both parities of the alternating FP32 UR kernel are sampled
# algorithmic bytes per interior cell per launch (arrays read once + written once).
# A CG iteration is two kernels.
#   FP64:  PW = [p = r + beta p; w = A p] reads p, r, Kx, Ky, writes p', w  (48 B)
#          UR = [u += alpha p; r -= alpha w] reads u, p, r, w, writes u, r   (48 B)
#   FP32 inner CG (operator from one 16-bit cell-coefficient array gc: half with a
#   power-of-two scale, or bfloat16 when the range does not fit; w not stored):
#          PW reads p, r (4 B each), gc (2 B), writes p'                    (14 B)
#          UR = [r -= alpha A p] reads p, r, gc, writes r                   (14 B)
#          and at even iterations also x += alpha(t-1) p(t-1) + alpha(t) p(t):
#          reads x, p(t-1), writes x                                     (+12 B)
#          i.e. 20 B per launch on average
#   FP32 inner CG as one kernel per iteration (the default when nx % 4 == 0):
#          reads r(t-1), p(t-1) (4 B each), gc (2 B), writes r(t), p(t)       (18 B)
#          and at even iterations reads and writes x                      (+8 B)
#          i.e. 22 B per launch (= per iteration) on average
#   FP32 inner CG as pairs of iterations (two sweeps per two iterations):
#          first sweep reads r(t-1), p(t-1), gc                           (10 B)
#          second sweep reads r(t-1), p(t-1), gc, x, writes r(t+1), p(t+1), x (26 B)
#          i.e. 18 B per iteration
# P64 is only the final p update of an FP64 solve.
KERNEL_BYTES = {
    _lib.TL_K_CG_W64: 48, _lib.TL_K_CG_UR64: 48, _lib.TL_K_CG_P64: 24,
    _lib.TL_K_CG_W32: 14, _lib.TL_K_CG_UR32: 20, _lib.TL_K_CG_P32: 0, _lib.TL_K_CGF32: 22,
    _lib.TL_K_CGP_G32: 10, _lib.TL_K_CGP_U32: 26,
}
# kernels whose odd and even iterations move different bytes: (odd, even) B/cell
ALTERNATING = {_lib.TL_K_CG_UR32: (14, 26), _lib.TL_K_CGF32: (18, 26)}
WEIGHTS = os.path.join(ROOT, "tealeaf.jl_amd", "weights", "unet_warmstart.npy")
# HBM bytes per launch of the CG kernels at 8192^2 from rocprofv3 FETCH_SIZE /
# WRITE_SIZE counter runs, scaled by a copy kernel of known size
PMC_TRAFFIC = os.path.join(ROOT, "profiles", "pmc_traffic_8192.json")


def measured_traffic(kernel, n):
    if n != 8192 or not os.path.exists(PMC_TRAFFIC):
        return None
    with open(PMC_TRAFFIC) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_settings(n, seed, precision, warm):
    text = decks.random_rect_deck(seed, n=n, steps=1000, precision=precision, warm_start=warm)
    return T.Settings.from_text(text)


FP32_KERNELS = {"auto": _lib.TL_FP32_KERNELS_AUTO, "two": _lib.TL_FP32_KERNELS_TWO,
                "fused": _lib.TL_FP32_KERNELS_FUSED, "pairs": _lib.TL_FP32_KERNELS_PAIRS}


def run_method(name, n, seed, precision, warm, steps, warmup, profile_mask=0, launch="stream", ir_delta=None,
               fp32_kernels="auto"):
    s = make_settings(n, seed, precision, warm)
    s.set(launch={"stream": T.TL_LAUNCH_STREAM, "graph": T.TL_LAUNCH_GRAPH}[launch],
          fp32_kernels=FP32_KERNELS[fp32_kernels])
    if ir_delta is not None:
        s.set(ir_delta=ir_delta)
    chunk = T.initialiseapp(s)
    if warm == "learned":
        from tealeaf_hip import unet
        unet.load_weights(chunk, WEIGHTS)
    stats = []
    for w in range(warmup):
        st = T.step(chunk, s)
        T.fieldsummary(chunk)
        log(f"[{name}] warmup step {w + 1}: {st['iterations']} it, {st['solve_ms']:.0f} ms, "
            f"rel_resid {st['rel_resid']:.2e}")
    L = _lib.load()
    if profile_mask:
        # one launch in PROFILE_PERIOD of each profiled kernel is bracketed by
        # HIP events (an event pair costs a few microseconds of GPU time)
        _lib.check(L.tl_profile_begin(chunk._ctx, profile_mask, PROFILE_PERIOD), chunk._ctx, "profile")
    times = []
    summaries = []
    t_all0 = time.perf_counter()
    for k in range(steps):
        t0 = time.perf_counter()
        st = T.step(chunk, s)
        summ = T.fieldsummary(chunk)
        t1 = time.perf_counter()
        times.append(t1 - t0)
        stats.append(st)
        summaries.append(summ)
        log(f"[{name}] step {warmup + k + 1}: {t1 - t0:.3f} s, {st['iterations']} it, "
            f"refreshes {st['refreshes']}, fallback {st['fallback']}, rel_resid {st['rel_resid']:.2e}")
    t_all1 = time.perf_counter()
    prof = {}
    if profile_mask:
        for kid in range(_lib.TL_K_COUNT):
            if profile_mask & (1 << kid):
                cnt = C.c_int()
                _lib.check(L.tl_profile_read(chunk._ctx, kid, None, 0, C.byref(cnt)), chunk._ctx, "profile")
                buf = np.zeros(cnt.value)
                _lib.check(L.tl_profile_read(chunk._ctx, kid, buf.ctypes.data_as(C.POINTER(C.c_double)),
                                             cnt.value, C.byref(cnt)), chunk._ctx, "profile")
                prof[kid] = buf
        L.tl_profile_end(chunk._ctx)
    chunk.close()
    return {"times": times, "total": t_all1 - t_all0, "stats": stats, "summaries": summaries, "prof": prof,
            "ir_delta": s.ir_delta}


def roofline_from(prof, n, coef_format=None):
    """Dominant kernel by total time; average over launches that ran (launches
    made after a solve's stop flag exit at once and are excluded)."""
    best = None
    for kid, d in prof.items():
        if len(d) == 0:
            continue
        med = float(np.median(d))
        live = d[d > 0.25 * med]
        tot = float(live.sum())
        if best is None or tot > best[1]:
            best = (kid, tot, live, len(d) - len(live))
    if best is None:
        return None
    kid, tot, live, skipped = best
    avg_ms = float(live.mean())
    algo_bytes = KERNEL_BYTES[kid] * n * n
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(_lib.kernel_name(kid, coef_format), n),
        "kernel": _lib.kernel_name(kid, coef_format), "avg_ms": round(avg_ms, 5), "timed_launches": int(len(live)),
        "sampling": f"1 launch in {PROFILE_PERIOD}",
        "stopped_launches_excluded": int(skipped), "algorithmic_bytes_per_launch": algo_bytes,
        "bytes_per_cell": KERNEL_BYTES[kid],
    }


def pair_summary(prof, n):
    """The FP32 CG pair as a whole (both sweeps per two iterations): 36 B/cell
    of algorithmic bytes over the sum of the two sweeps' average live launch
    durations; None when the pair form did not run."""
    d = [prof.get(_lib.TL_K_CGP_G32), prof.get(_lib.TL_K_CGP_U32)]
    if any(x is None or len(x) == 0 for x in d):
        return None
    ms = []
    for x in d:
        live = x[x > 0.25 * float(np.median(x))]
        ms.append(float(live.mean()))
    b = (KERNEL_BYTES[_lib.TL_K_CGP_G32] + KERNEL_BYTES[_lib.TL_K_CGP_U32]) * n * n
    gbs = b / ((ms[0] + ms[1]) * 1e-3) / 1e9
    return {"first_sweep_ms": round(ms[0], 5), "second_sweep_ms": round(ms[1], 5),
            "ms_per_iteration": round((ms[0] + ms[1]) / 2, 5), "bytes_per_cell_per_iteration": b // (2 * n * n),
            "achieved_GBps": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4)}


def cpu_model():
    """CPU model name of this host (first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(n, seed, gpu_iters, budget_s, c1_n=512):
    """The CPU oracle's FP64 CG (the reference's algorithm restated in C) on
    this host.  The reference is single-threaded, so `value` is the 1-thread
    figure: a bounded number of CG iterations of the benchmark's first step at
    n^2 is timed and s/timestep extrapolated with the GPU's FP64 iteration
    count.  The same with OpenMP on every core this process may use is
    reported beside it, and the 512^2 reference deck (10 CG steps) is timed in
    full, on all cores and (when it fits the budget) on 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    d = O.parse_deck(decks.random_rect_deck(seed, n=n, steps=1))
    d.stop_rule, d.rel_tol, d.maxiters = O.STOP_RELATIVE, 1e-10, 2000
    t0 = time.perf_counter()
    o = O.Oracle(d, threads=cores)
    o.initialise()
    log(f"cpu baseline: {n}^2 problem set up on {cores} threads in {time.perf_counter() - t0:.1f} s")

    def per_iteration(threads, budget):
        O.lib().or_set_threads(threads)

        def timed(iters):
            o.set_maxiters(iters)
            t0 = time.perf_counter()
            o.step()
            return time.perf_counter() - t0

        timed(2)  # first touch of the arrays on this thread count
        t2, t6 = timed(2), timed(6)
        per = max((t6 - t2) / 4, 1e-9)
        fixed = max(t2 - 2 * per, 0.0)
        iters = int(max(4, min(2000, (budget - fixed) / per)))
        tn = timed(iters)
        per = (tn - fixed) / iters
        r = {"s_per_timestep": float(f"{fixed + per * gpu_iters:.5g}"), "threads": threads,
             "iterations_timed": iters, "timed_s": round(tn, 2), "ms_per_iteration": float(f"{per * 1e3:.5g}"),
             "step_overhead_s": float(f"{fixed:.4g}")}
        log(f"cpu baseline: {threads} thread(s): {r['ms_per_iteration']} ms/iteration, {iters} iterations timed")
        return r

    allc = per_iteration(cores, budget_s * 0.3)
    one = per_iteration(1, budget_s * 0.4)
    del o

    def c1(threads):
        t0 = time.perf_counter()
        _, (summ, st) = O.run_deck(decks.reference_deck(c1_n, 10), threads=threads)
        dt = time.perf_counter() - t0
        log(f"cpu baseline: {c1_n}^2 deck, 10 CG steps on {threads} thread(s): {dt:.1f} s")
        return {"threads": threads, "s_total": round(dt, 3), "s_per_timestep": round(dt / 10, 4),
                "iterations": [x["iterations"] for x in st], "temp_step10": float(summ[-1][3])}

    c1_runs = [c1(cores)]
    est_one = c1_runs[0]["s_total"] * one["ms_per_iteration"] / max(allc["ms_per_iteration"], 1e-9)
    if est_one <= budget_s:
        c1_runs.append(c1(1))
    return {
        "value": one["s_per_timestep"], "unit": "s/timestep", "cores": 1, "kind": "port",
        "cpu_model": cpu_model(), "host_cores_available": cores,
        "sample": (f"CPU oracle (the reference's FP64 CG restated in C, -O3) at {n}^2 (seed {seed}) on 1 thread: "
                   f"{one['iterations_timed']} CG iterations timed ({one['timed_s']} s, {one['ms_per_iteration']} "
                   f"ms/iteration), extrapolated to the GPU FP64 CG's {gpu_iters:.0f} iterations/step"),
        "all_cores": dict(allc, note=f"same sample with OpenMP on {cores} threads, extrapolated likewise"),
        f"c1_{c1_n}_10_steps": c1_runs,
    }


def solvers_run(n, steps):
    """The 3-state reference deck at n^2 (presteps 8, 10 PPCG inner steps, eps
    1e-15, each solver's own stop rule) for CG, Chebyshev and PPCG: one JSON
    line per solver with s/timestep, iterations and field summaries, then the
    largest relative summary difference from CG.  For kernel traces of the
    FP64 solver kernels."""
    res = {}
    for solver in ("cg", "cheby", "ppcg"):
        s = T.Settings.from_text(decks.reference_deck(n, steps, solver=solver, presteps=8, ppcg_inner_steps=10))
        chunk = T.initialiseapp(s)
        times, its, summ = [], [], []
        for k in range(steps):
            t0 = time.perf_counter()
            st = T.step(chunk, s)
            sm = T.fieldsummary(chunk)
            times.append(time.perf_counter() - t0)
            its.append([st["iterations"], st["cg_iters"], st["cheby_iters"]])
            summ.append(sm)
        chunk.close()
        res[solver] = np.array(summ)
        print(json.dumps({"solver": solver, "n": n, "s_per_step": [round(x, 4) for x in times],
                          "iterations_cg_cheby": its, "summary": [list(map(float, x)) for x in summ]}), flush=True)
    for solver in ("cheby", "ppcg"):
        d = float(np.max(np.abs(res[solver] - res["cg"]) / np.abs(res["cg"])))
        print(json.dumps({"solver": solver, "max_rel_summary_diff_vs_cg": d}), flush=True)


def compare_run(n, seed, steps, fp32_kernels="auto"):
    """The benchmark problem for `steps` timesteps with the FP32-refined CG
    and with FP64 CG side by side: one JSON line per step with both field
    summaries, their largest relative difference, iterations, true residuals
    and times."""
    runs = {}
    for prec in ("fp32_ir", "fp64"):
        s = make_settings(n, seed, prec, "none").set(fp32_kernels=FP32_KERNELS[fp32_kernels])
        runs[prec] = (s, T.initialiseapp(s))
    worst = 0.0
    for k in range(steps):
        row = {"step": k + 1}
        summ = {}
        for prec, (s, chunk) in runs.items():
            t0 = time.perf_counter()
            st = T.step(chunk, s)
            summ[prec] = T.fieldsummary(chunk)
            row[prec] = {"s": round(time.perf_counter() - t0, 3), "iterations": st["iterations"],
                         "refreshes": st["refreshes"], "rel_resid": st["rel_resid"]}
        d = float(np.max(np.abs(summ["fp32_ir"] - summ["fp64"]) / np.abs(summ["fp64"])))
        worst = max(worst, d)
        row["summary_fp64"] = [float(x) for x in summ["fp64"]]
        row["max_rel_diff"] = d
        print(json.dumps(row), flush=True)
    print(json.dumps({"steps": steps, "max_rel_diff_all_steps": worst}), flush=True)


def kernel_repeat(n, seed, kid, reps):
    """For counter runs on the 8192^2 problem: `reps` iterations of the CG
    that kernel kid belongs to (tl_kernel_repeat), or for kid -1 / -2 that
    many launches of a same-size copy kernel (16 B/cell)."""
    s = make_settings(n, seed, "fp64", "none")
    chunk = T.initialiseapp(s)
    _lib.check(_lib.load().tl_kernel_repeat(chunk._ctx, kid, reps), chunk._ctx, "kernel_repeat")
    chunk.close()
    print(json.dumps({"kernel": _lib.KERNEL_NAMES[kid], "reps": reps, "n": n}))


def unet_time(n, seed, reps):
    """The learned start's inference alone on the n^2 benchmark problem: after
    one FP32-refined timestep, `reps` calls of tl_unet_predict (face
    coefficients, features, the U-Net forward pass, the guess), each timed on
    the host; one JSON line with the mean and the fastest call."""
    s = make_settings(n, seed, "fp32_ir", "learned")
    chunk = T.initialiseapp(s)
    from tealeaf_hip import unet
    unet.load_weights(chunk, WEIGHTS)
    T.step(chunk, s)
    L = _lib.load()
    rx, ry = s.dtinit / (s.dx * s.dx), s.dtinit / (s.dy * s.dy)
    times = []
    for i in range(reps + 1):
        t0 = time.perf_counter()
        # returns after a synchronisation of the context's stream
        _lib.check(L.tl_unet_predict(chunk._ctx, rx, ry, s.coefficient), chunk._ctx, "tl_unet_predict")
        if i:  # the first call allocates the activations
            times.append((time.perf_counter() - t0) * 1e3)
        log(f"[unet] call {i + 1}: {(time.perf_counter() - t0) * 1e3:.2f} ms")
    chunk.close()
    print(json.dumps({"unet_predict_ms_mean": round(float(np.mean(times)), 3),
                      "unet_predict_ms_min": round(float(np.min(times)), 3), "calls": reps, "n": n}), flush=True)


# algorithmic bytes per interior cell of the kernels kernel_table measures
# (FP64 solver kernels: arrays read once, written once; k_residual in the form
# tl_kernel_repeat runs it, writing the FP64 r)
TABLE_BYTES = dict(KERNEL_BYTES)
TABLE_BYTES.update({_lib.TL_K_CHEBY64: 56, _lib.TL_K_PPCG64: 64, _lib.TL_K_JACOBI64: 40, _lib.TL_K_RESID64: 40,
                    _lib.TL_K_PPCG2_64: 64,   # two PPCG inner steps: 64 B/cell per launch
                    _lib.TL_K_CHEBY2_64: 56})  # two Chebyshev iterations: 56 B/cell per launch


def kernel_table(n, seed, reps):
    """Per-kernel timing on the n^2 benchmark problem: `reps` launches (CG:
    iterations) of each hot kernel bracketed by HIP events on the library's
    stream (tl_kernel_repeat + tl_profile_*), one JSON line per kernel with the
    mean duration and the algorithmic GB/s; the 16-byte copy kernel is timed
    with the host clock around `reps` launches as the bandwidth reference."""
    L = _lib.load()
    groups = [(_lib.TL_K_CG_W64, (_lib.TL_K_CG_W64, _lib.TL_K_CG_UR64)),
              (_lib.TL_K_CG_W32, (_lib.TL_K_CG_W32, _lib.TL_K_CG_UR32)),
              (_lib.TL_K_CGF32, (_lib.TL_K_CGF32,)),
              (_lib.TL_K_CGP_G32, (_lib.TL_K_CGP_G32, _lib.TL_K_CGP_U32)),
              (_lib.TL_K_CHEBY64, (_lib.TL_K_CHEBY64,)), (_lib.TL_K_PPCG64, (_lib.TL_K_PPCG64,)),
              (_lib.TL_K_PPCG2_64, (_lib.TL_K_PPCG2_64,)), (_lib.TL_K_CHEBY2_64, (_lib.TL_K_CHEBY2_64,)),
              (_lib.TL_K_JACOBI64, (_lib.TL_K_JACOBI64,)), (_lib.TL_K_RESID64, (_lib.TL_K_RESID64,))]
    s = make_settings(n, seed, "fp64", "none")
    chunk = T.initialiseapp(s)
    ctx = chunk._ctx
    for run_id, kids in groups:
        mask = 0
        for k in kids:
            mask |= 1 << k
        _lib.check(L.tl_profile_begin(ctx, mask, 1), ctx, "profile")
        _lib.check(L.tl_kernel_repeat(ctx, run_id, reps), ctx, "kernel_repeat")
        for k in kids:
            cnt = C.c_int()
            _lib.check(L.tl_profile_read(ctx, k, None, 0, C.byref(cnt)), ctx, "profile")
            buf = np.zeros(cnt.value)
            _lib.check(L.tl_profile_read(ctx, k, buf.ctypes.data_as(C.POINTER(C.c_double)), cnt.value,
                                         C.byref(cnt)), ctx, "profile")
            # launch i is iteration i + 1; the first quarter counts as warm-up
            idx = np.arange(len(buf))
            keep = idx >= len(buf) // 4
            if k in ALTERNATING:
                forms = {"odd": buf[keep & (idx % 2 == 0)], "even": buf[keep & (idx % 2 == 1)]}
            else:
                forms = {"": buf[keep]}
            for form, v in forms.items():
                b = TABLE_BYTES[k] if form == "" else ALTERNATING[k][0 if form == "odd" else 1]
                ms = float(np.mean(v))
                print(json.dumps({"kernel": _lib.KERNEL_NAMES[k] + (f" ({form} t)" if form else ""), "n": n,
                                  "launches": int(len(v)), "avg_ms": round(ms, 5), "median_ms": round(
                                      float(np.median(v)), 5), "bytes_per_cell": b,
                                  "GBps": round(b * n * n / (ms * 1e-3) / 1e9, 1)}), flush=True)
        L.tl_profile_end(ctx)
    for kid, b in ((-2, 16),):
        L.tl_kernel_repeat(ctx, kid, 5)
        t0 = time.perf_counter()
        _lib.check(L.tl_kernel_repeat(ctx, kid, reps), ctx, "kernel_repeat")
        ms = (time.perf_counter() - t0) / reps * 1e3
        pitch = (n + chunk.hd + 32 + 63) // 64 * 64  # the library's row pitch (elements)
        nbytes = 2 * 8 * (pitch * chunk.Y + 64)       # one field allocation read and written
        print(json.dumps({"kernel": _lib.KERNEL_NAMES[kid], "n": n, "launches": reps, "avg_ms": round(ms, 5),
                          "bytes": nbytes, "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
    chunk.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--kernel-repeat", type=int, nargs=2, metavar=("KERNEL_ID", "REPS"))
    ap.add_argument("--compare-steps", type=int, default=0,
                    help="FP32-refined vs FP64 CG side by side for this many timesteps (per-step summaries)")
    ap.add_argument("--only", choices=("flagship", "fp64"), default=None,
                    help="run one method only (long runs); per-step field summaries are in the JSON")
    ap.add_argument("--fp32-kernels", choices=tuple(FP32_KERNELS), default="auto",
                    help="FP32-refined CG: one fused kernel per iteration or the two-kernel form")
    ap.add_argument("--launch", choices=("stream", "graph"), default="stream",
                    help="CG iterations as stream launches or hipGraph replays (graph: no per-kernel timing)")
    ap.add_argument("--ir-delta", type=float, default=None,
                    help="FP32-refined CG: refresh the FP64 residual when the FP32 one falls by this factor")
    ap.add_argument("--solvers", action="store_true",
                    help="CG, Chebyshev and PPCG on the 3-state reference deck at --n (default 4096 here)")
    ap.add_argument("--unet-time", type=int, default=0, metavar="REPS",
                    help="time REPS calls of the learned start's inference alone at --n")
    ap.add_argument("--kernel-table", type=int, default=0, metavar="REPS",
                    help="per-kernel HIP-event timing of every hot kernel, REPS launches each, at --n")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the solve would shard by 2-D domain decomposition "
              "(halo exchange of p and u, two scalar all-reduces per CG iteration) but no multi-GPU path "
              "is built, so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if a.kernel_table:
        kernel_table(a.n, a.seed, a.kernel_table)
        return
    if a.unet_time:
        unet_time(a.n, a.seed, a.unet_time)
        return
    if a.kernel_repeat:
        kernel_repeat(a.n, a.seed, a.kernel_repeat[0], a.kernel_repeat[1])
        return
    if a.solvers:
        solvers_run(a.n if a.n != 8192 else 4096, a.steps)
        return
    if a.compare_steps:
        compare_run(a.n, a.seed, a.compare_steps, a.fp32_kernels)
        return
    learned = os.path.exists(WEIGHTS)
    warm = "learned" if learned else "none"
    if a.only == "fp64":
        base = run_method("fp64-cg", a.n, a.seed, "fp64", "none", a.steps, a.warmup, launch=a.launch)
        print(json.dumps({"method": "fp64-cg", "grid": a.n, "seed": a.seed, "steps": a.steps,
                          "s_per_step": [round(t, 4) for t in base["times"]],
                          "iterations": [st["iterations"] for st in base["stats"]],
                          "rel_resid": [st["rel_resid"] for st in base["stats"]],
                          "summaries": [list(map(float, x)) for x in base["summaries"]]}))
        return
    mask = ((1 << _lib.TL_K_CG_W32) | (1 << _lib.TL_K_CG_UR32) | (1 << _lib.TL_K_CGF32) | (1 << _lib.TL_K_CGP_G32)
            | (1 << _lib.TL_K_CGP_U32))
    flag = run_method("fp32-ir" + ("+learned" if learned else ""), a.n, a.seed, "fp32_ir", warm, a.steps,
                      a.warmup, mask if a.launch == "stream" else 0, a.launch, a.ir_delta, a.fp32_kernels)
    base = None
    if not a.no_baseline and a.only is None:
        base = run_method("fp64-cg", a.n, a.seed, "fp64", "none", a.steps, a.warmup, launch=a.launch)
    value = float(np.mean(flag["times"]))
    out = {
        "metric": "s/timestep at 8192², rel. resid 1e-10: learned+FP32-refined vs plain FP64 CG",
        "value": round(value, 4), "unit": "s/timestep", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(value * 1e3, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 inner / fp64 residual+solution+reductions", "data": "synthetic",
        "config": {
            "workload": f"TeaLeaf {a.n}x{a.n} random rectangles (32, seed {a.seed}), dt 0.004, rel. resid 1e-10",
            "grid": a.n, "seed": a.seed, "warm_start": warm, "parallelism": "single GPU", "launch": a.launch,
            "fp32_kernels": a.fp32_kernels,
        },
        "flagship": {
            "s_per_step": [round(t, 4) for t in flag["times"]],
            "median": round(float(np.median(flag["times"])), 4),
            "p90": round(float(np.percentile(flag["times"], 90)), 4),
            "iterations": [st["iterations"] for st in flag["stats"]],
            "refreshes": [st["refreshes"] for st in flag["stats"]],
            "ir_delta": flag["ir_delta"],
            "warm_attempted": [st["warm_attempted"] for st in flag["stats"]],
            "warm_ms": [round(st["warm_ms"], 2) for st in flag["stats"]],
            "fallback": [st["fallback"] for st in flag["stats"]],
            "start": [["previous", "learned"][st["start"]] for st in flag["stats"]],
            "init_rel_resid": [st["init_rel_resid"] for st in flag["stats"]],
            "rel_resid": [st["rel_resid"] for st in flag["stats"]],
        },
    }
    if base is not None:
        bval = float(np.mean(base["times"]))
        out["baseline_fp64_cg"] = {
            "s_per_step": [round(t, 4) for t in base["times"]], "mean": round(bval, 4),
            "median": round(float(np.median(base["times"])), 4),
            "p90": round(float(np.percentile(base["times"], 90)), 4),
            "iterations": [st["iterations"] for st in base["stats"]],
            "rel_resid": [st["rel_resid"] for st in base["stats"]],
        }
        out["speedup_vs_fp64_cg"] = round(bval / value, 3)
        diffs = [float(np.max(np.abs(f - b) / np.abs(b))) for f, b in zip(flag["summaries"], base["summaries"])]
        out["field_summary_max_rel_diff_vs_fp64"] = max(diffs)
    if a.only == "flagship":
        out["flagship"]["summaries"] = [list(map(float, x)) for x in flag["summaries"]]
    fmt = flag["stats"][0].get("coef_format") if flag["stats"] else None
    out["roofline"] = roofline_from(flag["prof"], a.n, fmt)
    pair = pair_summary(flag["prof"], a.n)
    if pair is not None:
        out["fp32_pair"] = pair
    # every profiled kernel of the flagship's timed steps, and the share of the
    # timed wall time they cover (launches past a solve's stop excluded; the
    # totals scale the sampled launches by the sampling period)
    kern = {}
    for kid, d in flag["prof"].items():
        if len(d) == 0:
            continue
        live = d[d > 0.25 * float(np.median(d))]
        kern[_lib.kernel_name(kid, fmt)] = {"timed_launches": int(len(live)), "avg_ms": round(float(live.mean()), 5),
                                         "total_s_est": round(float(live.sum()) * 1e-3 * PROFILE_PERIOD, 4),
                                         "GBps": round(KERNEL_BYTES[kid] * a.n * a.n / (float(live.mean()) * 1e-3) / 1e9, 1)}
    out["flagship_kernels"] = kern
    out["flagship_inner_cg_time_frac"] = round(sum(k["total_s_est"] for k in kern.values()) / flag["total"], 4)
    if not a.no_cpu:
        it = (np.mean([st["iterations"] for st in base["stats"]]) if base is not None
              else np.mean([st["iterations"] for st in flag["stats"]]))
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(a.n, a.seed, it, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
