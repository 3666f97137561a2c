"""Benchmark of the fastEDM prediction core on one MI355X.

Workload (default): the convergent-cross-mapping sweep on coupled logistic maps
  x_{t+1} = x_t (3.8 - 3.8 x_t - 0.02 y_t),  y_{t+1} = y_t (3.5 - 3.5 y_t - 0.1 x_t),  x0 = 0.4, y0 = 0.2,
T = 20000 points, edm(t, y, x, E=3, library = 40 log-spaced sizes 10..19000, numReps = 100): cross-map x
from the embedding of y, simplex with k = E+1 = 4, 4000 tasks of 19998 predictions each.

One step = every prediction of the sweep (fedm_plan_run), inputs resident on the GPU.  Reported:
  value          predictions/s of the batched mixed-precision route (long libraries: FP32-screened rank
                 tables with FP64 re-rank; short libraries: the exact FP64 direct kernel)
  plain          the plain FP64 drop-in route, one edm_predict per task (the reference's per-task seam)
  plain_batched  FP64 in the batched launch, every library through the direct kernel
  roofline       the stage that dominates the device time, with its algorithmic rate from HIP-event
                 kernel times against the matching MI355X peak
  cpu_baseline   the CPU oracle (a restatement of the reference CPU path, one thread per host core)
                 timed on a bounded sample of the same sweep
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simplex/S-map predictions/sec, batched-mixed vs plain GPU, at max|d yhat|<=1e-6"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def load_product():
    import importlib.util
    spec = importlib.util.spec_from_file_location("fastedm_r_amd", os.path.join(ROOT, "fastedm-r_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(ROOT, "fastedm-r_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fastedm_r_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def coupled_logistic(T):
    x = np.empty(T)
    y = np.empty(T)
    x[0], y[0] = 0.4, 0.2
    for i in range(T - 1):
        x[i + 1] = x[i] * (3.8 - 3.8 * x[i] - 0.02 * y[i])
        y[i + 1] = y[i] * (3.5 - 3.5 * y[i] - 0.1 * x[i])
    return x, y


def ccm_libraries():
    return np.round(np.exp(np.linspace(np.log(10), np.log(19000), 40))).astype(int)


def ccm_workload(reps=100):
    x, y = coupled_logistic(20000)
    t = np.arange(1, 20001, dtype=float)
    libs = ccm_libraries()
    kw = dict(E=3, library=libs, numReps=reps, savePredictions=False)
    # cross-map x from the embedding of y: edm(t, y, x)
    return t, y, x, kw


def pair_count(libs, reps, n_pred=19998):
    return int(reps) * int(np.sum(libs)) * n_pred


def timed_steps(plan, fe, steps, warmup):
    """Wall time per step and the mean HIP-event stage times [batch, rank build, rank scan, direct] (ms)."""
    for _ in range(warmup):
        plan.run()
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
        stages.append(fe.timings())
    wall = time.perf_counter() - t0
    return wall / steps, np.mean(np.asarray(stages), axis=0)


HBM_PEAK_GBS = 8000.0
N_TARGETS = 19998
K = 4


def rank_threshold(k, npool):
    import ctypes
    fe = sys.modules["fastedm_r_amd"]
    f = fe.load_library().fedm_rank_threshold
    f.argtypes = [ctypes.c_int, ctypes.c_int]
    return f(k, npool)


def rank_route_rates(st, libs, reps):
    """The rank-table route in its own units: the build ranks every (pool row, target) pair once per
    step (two FP32 passes over the pool per target), the scan walks ranked-prefix entries per (task,
    target): about k * pool / L of them for a library of L rows drawn from the pool."""
    pool = N_TARGETS
    lthr = rank_threshold(K, pool)
    rank_libs = libs[libs >= lthr]
    pairs = float(pool) * N_TARGETS
    entries = reps * float(np.sum(K * pool / rank_libs)) * N_TARGETS
    return {"rank_tasks": int(reps * rank_libs.size), "library_threshold": int(lthr),
            "build_pool_target_pairs": pairs, "build_pairs_per_s": pairs / (st[1] * 1e-3),
            "scan_prefix_entries": entries, "scan_entries_per_s": entries / (st[2] * 1e-3),
            "note": "one ranked prefix per target serves every long-library task, so the brute-force "
                    "pair count of the step (215.8e9) overstates the work this route does"}


def roofline(st, libs, reps):
    """Roofline of the stage that dominates the device time, with that stage's algorithmic work.

    direct     (batch_simplex): 3 E_a = 9 flops per library-target pair scanned (FP64 for short libraries);
    rank build (rank_build):    9 flops per pool-target pair and histogram/filter pass (2 passes);
    rank scan  (rank_scan_multi + rank_scatter): HBM bytes of the ranked-prefix entries the tasks' k-NN
                                selections need (k pool / L entries of 4 B plus k entries of 8 B per (task,
                                target)) and of their outputs (16 B per prediction)."""
    pool = N_TARGETS
    lthr = rank_threshold(K, pool)
    direct_libs = libs[libs < lthr]
    rank_libs = libs[libs >= lthr]
    work = {
        # libraries below the rank threshold are short: the direct kernel runs its FP64 variant; both
        # kernels run on the vector units (no matrix-core instruction), hence "valu"
        "direct": ("valu", 9.0 * reps * float(np.sum(direct_libs)) * N_TARGETS, "TFLOP/s", FP64_PEAK_TFLOPS,
                   "batch_simplex<4,3>"),
        "rank_build": ("valu", 2 * 9.0 * pool * N_TARGETS, "TFLOP/s", FP32_PEAK_TFLOPS, "rank_build<4>"),
        "rank_scan": ("hbm", reps * float(np.sum(K * pool / rank_libs * 4.0 + K * 8.0 + 16.0)) * N_TARGETS, "GB/s",
                      HBM_PEAK_GBS, "rank_scan_multi<4,1> + rank_scatter"),
    }
    # HBM bytes per launch from rocprofv3 FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) passes over a steady-state
    # step, scaled by the same counters on a 1 GiB copy (profiles/ccm_sweep_E3_k4.md); the direct kernel's
    # writes include its register spills (0.56 GB) beside the 0.64 GB of outputs
    traffic = {"direct": 1.40e9, "rank_build": 0.16e9, "rank_scan": 1.03e9 + 0.98e9}
    times = {"rank_build": st[1], "rank_scan": st[2], "direct": st[3]}
    name = max(times, key=lambda n: times[n])
    bound, amount, unit, peak, kernel = work[name]
    ms = float(times[name])
    achieved = amount / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "traffic": traffic[name],
            "traffic_source": "not measured by this run: rocprofv3 FETCH_SIZE/WRITE_SIZE passes over the same "
                              "command, scaled by a 1 GiB copy (profiles/ccm_sweep_E3_k4.md)",
            "kernel": kernel, "kernel_ms": ms,
            "algorithmic_work_per_launch": amount,
            "engine": {"direct": "fp64 VALU", "rank_build": "fp32 VALU", "rank_scan": "HBM3E"}[name]}


def panel_smap_leg(fe):
    """Secondary line: S-map on 512 logistic-map panels (E = 8 with dt, E_a = 16, k = 18, theta 0, 1, 3;
    265 665 library rows x 265 665 targets), mixed against plain, device stage times of one call each."""
    from tests import configs
    t, x, panel = configs.cfg4_panels()
    kw = dict(panel=panel, E=8, dt=True, theta=[0, 1, 3], algorithm="smap", savePredictions=True)
    out = {"workload": "panel_smap_512_logistic_panels_E8_dt_k18_3thetas", "predictions": None}
    res = {}
    for name, prec in (("mixed", fe.MIXED), ("plain", fe.PLAIN)):
        t0 = time.perf_counter()
        r = fe.edm(t, x, precision=prec, **kw)
        wall = time.perf_counter() - t0
        st = fe.timings()
        res[name] = r
        out[name] = {"device_ms": float(st[3]), "select_ms": float(st[0]), "merge_ms": float(st[1]),
                     "smap_ms": float(st[2]), "wall_s": wall}
    out["predictions"] = int(np.sum(~np.isnan(res["plain"]["predictions"]))) * 3
    out["device_ratio_mixed_vs_plain"] = out["plain"]["device_ms"] / out["mixed"]["device_ms"]
    a, b = res["mixed"]["predictions"], res["plain"]["predictions"]
    ok = ~np.isnan(b)
    out["max_abs_d_yhat_mixed_vs_plain"] = float(np.max(np.abs(a[ok] - b[ok]))) if ok.any() else 0.0
    return out


def baseline_oracle():
    """The oracle built -O3 -march=native -ffp-contract=off on this host (4 s); the prebuilt -O2 one if
    the compiler is missing.  Returns (module, build flags)."""
    import subprocess
    import tempfile
    from oracle import oracle
    out = tempfile.mkdtemp(prefix="fedm_baseline_")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"NATIVE_OUT={out}"],
                       check=True, capture_output=True, timeout=120)
        oracle.use_library(os.path.join(out, "liboracle_native.so"))
        return oracle, "-O3 -march=native -ffp-contract=off"
    except Exception:
        return oracle, "-O2 -ffp-contract=off (prebuilt; the native build failed)"


def cpu_baseline(t, x, y, kw, seconds):
    """The oracle's edm() (tasks serial, targets over a thread per host core, as the reference CPU path
    runs them) on the first replicates of the sweep: a sample sized to about `seconds`, timed three
    times, median reported.  The timed call includes the oracle's manifold builds and rho/MAE."""
    oracle, flags = baseline_oracle()
    cores = os.cpu_count() or 1
    reps = 1
    libs = kw["library"]
    while True:
        sample = dict(kw, numReps=reps, numThreads=cores)
        t0 = time.perf_counter()
        oracle.edm(t, x, y, **sample)
        dt = time.perf_counter() - t0
        if dt >= seconds / 2 or reps >= 100:
            break
        reps = min(100, max(reps + 1, int(reps * seconds / max(dt, 1e-3)) + 1))
    times = []
    res = None
    for _ in range(3):
        sample = dict(kw, numReps=reps, numThreads=cores, savePredictions=True)
        t0 = time.perf_counter()
        res = oracle.edm(t, x, y, **sample)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    preds = reps * len(libs) * N_TARGETS
    return {"value": preds / dt, "unit": "predictions/s", "cores": cores, "kind": "port",
            "build": flags, "runs_s": [round(v, 3) for v in times],
            "sample": f"the first {reps} of the 100 replicates x 40 libraries ({preds} predictions), median of 3 "
                      f"runs ({dt:.1f} s, {cores} threads), manifold builds and rho/MAE included; rc={res['rc']}"}, \
        res, reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reps", type=int, default=100, help="replicates of the CCM sweep (100 = the full workload)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-panel", action="store_true", help="skip the panel S-map line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    fe = load_product()
    t, x, y, kw = ccm_workload(args.reps)
    libs = kw["library"]
    pairs = pair_count(libs, args.reps)

    # batched mixed-precision route (the headline)
    print("building the batched mixed-precision plan", file=sys.stderr, flush=True)
    mixed = fe.Plan(t, x, y, precision=fe.MIXED, **kw)
    if mixed.device_batches != 1:
        raise SystemExit("the CCM sweep did not map to one persistent device launch")
    n_pred = mixed.num_predictions
    sec, st = timed_steps(mixed, fe, args.steps, args.warmup)
    kms = float(st[0])
    print(f"mixed: {sec * 1e3:.1f} ms/step, device {st[0]:.1f} ms (rank build {st[1]:.1f}, rank scan {st[2]:.1f}, "
          f"direct {st[3]:.1f})", file=sys.stderr, flush=True)
    redos = fe.exact_redos()
    res_mixed = mixed.result()
    out_mixed = mixed.fetch()  # every prediction of the last timed step

    out = {"metric": METRIC, "value": n_pred / sec, "unit": "predictions/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "value_excludes": "fedm_plan_create (host normalisation, splits, point tables and their upload, once "
                             "per command) and fedm_plan_result (copy-back and rho/MAE); the step is fedm_plan_run "
                             "with every input resident in HBM",
           "config": {"workload": "ccm_sweep_coupled_logistic_T20000_E3_40lib_10to19000",
                      "replicates": args.reps, "tasks": int(args.reps * len(libs)), "predictions_per_step": n_pred,
                      "library_target_pairs_per_step": pairs, "k": 4, "algorithm": "simplex",
                      "form": "batched: rank-table route (FP32-screened exact ranking, one pass per target "
                              "over every long-library task) for libraries >= 1.5 sqrt(k n), direct exact FP64 "
                              "kernel for shorter ones"}}

    out["stages_ms"] = {"device_total": float(st[0]), "rank_build": float(st[1]), "rank_scan": float(st[2]),
                        "direct": float(st[3])}
    out["exact_reselections"] = redos  # (task, target) pairs of the last step handed to the exact selection
    out["roofline"] = roofline(st, libs, args.reps)
    out["rank_route"] = rank_route_rates(st, libs, args.reps)

    # plain forms and parity
    print("building the plain plans", file=sys.stderr, flush=True)
    pb = fe.Plan(t, x, y, precision=fe.PLAIN, **kw)
    sec_pb, st_pb = timed_steps(pb, fe, 3, 1)
    kms_pb = float(st_pb[0])
    res_pb = pb.result()
    out_pb = pb.fetch()
    out["plain_batched"] = {"value": n_pred / sec_pb, "ms_per_step": sec_pb * 1e3, "kernel_ms": kms_pb, "steps": 3}
    d_rho = np.nanmax(np.abs(res_mixed["stats"]["rho"] - res_pb["stats"]["rho"]))
    d_mae = np.nanmax(np.abs(res_mixed["stats"]["mae"] - res_pb["stats"]["mae"]))
    pm, pp = out_mixed[0], out_pb[0]
    out["parity"] = {"max_abs_d_rho_mixed_vs_plain": float(d_rho), "max_abs_d_mae_mixed_vs_plain": float(d_mae),
                     "k_range": [res_mixed["kMin"], res_mixed["kMax"]],
                     "predictions_compared": int(pm.size),
                     "max_abs_d_yhat_mixed_vs_plain_batched": float(np.max(np.abs(pm - pp))) if pm.size else 0.0,
                     "predictions_differing_bits_mixed_vs_plain_batched": int(np.sum(pm.view(np.int64) !=
                                                                                     pp.view(np.int64))),
                     "rc_and_k_used_equal_mixed_vs_plain_batched": bool(np.array_equal(out_mixed[1], out_pb[1]) and
                                                                       np.array_equal(out_mixed[2], out_pb[2]))}
    del out_pb
    if not args.no_plain:
        pt = fe.Plan(t, x, y, precision=2, **kw)
        t0 = time.perf_counter()
        pt.run()
        sec_pt = time.perf_counter() - t0
        res_pt = pt.result()
        out["plain"] = {"value": n_pred / sec_pt, "ms_per_step": sec_pt * 1e3,
                        "route": "one edm_predict per task (host gather + upload per task, as the drop-in seam)"}
        out["ratio_mixed_vs_plain"] = out["value"] / out["plain"]["value"]
        out["parity"]["max_abs_d_rho_plain_taskwise_vs_plain_batched"] = float(
            np.nanmax(np.abs(res_pt["stats"]["rho"] - res_pb["stats"]["rho"])))
        pt.close()
    pb.close()
    mixed.close()

    if not args.no_panel:
        print("panel S-map line (mixed and plain)", file=sys.stderr, flush=True)
        out["panel_smap"] = panel_smap_leg(fe)

    if not args.no_cpu:
        print("timing the CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"], res_cpu, reps = cpu_baseline(t, x, y, kw, args.cpu_seconds)
        n = len(res_cpu["stats"]["rho"])
        out["parity"]["oracle_sample_tasks"] = n
        out["parity"]["max_abs_d_rho_mixed_vs_oracle"] = float(
            np.nanmax(np.abs(res_mixed["stats"]["rho"][:n] - res_cpu["stats"]["rho"])))
        out["parity"]["max_abs_d_mae_mixed_vs_oracle"] = float(
            np.nanmax(np.abs(res_mixed["stats"]["mae"][:n] - res_cpu["stats"]["mae"])))
        # the oracle saves the last task's predictions (replicate `reps`, library 19000): all of them
        # against the same task of the mixed step
        # (R's layout: one row per table row, missing where a row is not a target; here every usable row
        # is one, in row order, as in the task's outputs)
        ref = res_cpu["predictions"][:, 0]
        ref = ref[~np.isnan(ref)]
        task = reps * len(libs) - 1
        got = pm[task * N_TARGETS:(task + 1) * N_TARGETS]
        if ref.size == got.size:
            out["parity"]["oracle_saved_task"] = int(task)
            out["parity"]["max_abs_d_yhat_mixed_vs_oracle_saved_task"] = float(np.max(np.abs(got - ref)))
            out["parity"]["max_rel_d_yhat_mixed_vs_oracle_saved_task"] = float(
                np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
