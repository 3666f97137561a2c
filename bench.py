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
  cpu_baseline   the CPU oracle (a restatement of the reference CPU path, one thread per host core, built
                 -O3 -march=native on the host it runs on) timed on a bounded sample of the same sweep
  graph_replay   the same step replayed as one captured hipGraph
  configs        the other workloads (S-map theta sweep, panel S-map, Lorenz-96 S-map): mixed and plain
                 timed runs with inputs resident, their rooflines, CPU baselines and parity figures
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
    """Wall time per step and the mean HIP-event stage times [batch, rank build, rank scan, direct stage,
    direct kernel, pruned direct kernel, exact re-selection, -] (ms)."""
    for _ in range(warmup):
        plan.run()
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
        stages.append(fe.timings_ex())
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


PR_MIN, PR_BCH = 64, 512  # library sizes the pruned direct kernel takes (kernels' constants)


def roofline(st, libs, reps):
    """Roofline of the kernel with the largest device time in the step, with that kernel's algorithmic work.

    batch_simplex          3 E_a = 9 FP64 flops per (target, library row) pair of its tasks (libraries below
                           64 rows);
    batch_simplex_pruned   the same unit over every pair of its tasks (libraries of 64..423 rows), although it
                           computes only the pairs the triangle bound does not exclude;
    rank_build             9 FP32 flops per pool-target pair and pass (2 passes);
    rank_scan              HBM bytes of the ranked-prefix entries the tasks' selections need (k pool / L entries
                           of 4 B plus k entries of 8 B per (task, target)) and of their outputs (16 B per
                           prediction)."""
    pool = N_TARGETS
    lthr = rank_threshold(K, pool)
    direct_libs = libs[libs < lthr]
    pruned = direct_libs[(direct_libs >= PR_MIN) & (direct_libs <= PR_BCH)]
    plain = direct_libs[(direct_libs < PR_MIN) | (direct_libs > PR_BCH)]
    rank_libs = libs[libs >= lthr]
    work = {
        "direct": ("valu", 9.0 * reps * float(np.sum(plain)) * N_TARGETS, "TFLOP/s", FP64_PEAK_TFLOPS,
                   "batch_simplex<4,3,6>", "fp64 VALU"),
        "direct_pruned": ("valu", 9.0 * reps * float(np.sum(pruned)) * N_TARGETS, "TFLOP/s", FP64_PEAK_TFLOPS,
                          "batch_simplex_pruned<4,3>", "fp64 VALU"),
        "rank_build": ("valu", 2 * 9.0 * pool * N_TARGETS, "TFLOP/s", FP32_PEAK_TFLOPS, "rank_build<4>", "fp32 VALU"),
        "rank_scan": ("hbm", reps * float(np.sum(K * pool / rank_libs * 4.0 + K * 8.0 + 16.0)) * N_TARGETS, "GB/s",
                      HBM_PEAK_GBS, "rank_scan_multi<4,1> + rank_scatter", "HBM3E"),
    }
    # HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over a steady-state step,
    # scaled by the same counters on a 1 GiB copy (profiles/ccm_sweep_E3_k4.md)
    traffic = {"direct": 0.67e9, "direct_pruned": 0.54e9, "rank_build": 0.14e9, "rank_scan": 0.98e9 + 0.98e9}
    times = {"rank_build": st[1], "rank_scan": st[2], "direct": st[4], "direct_pruned": st[5]}
    name = max(times, key=lambda n: times[n])
    bound, amount, unit, peak, kernel, engine = work[name]
    ms = float(times[name])
    achieved = amount / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "traffic": traffic[name],
            "traffic_source": "not measured by this run: rocprofv3 FETCH_SIZE/WRITE_SIZE passes over the same "
                              "command, scaled by a 1 GiB copy (profiles/ccm_sweep_E3_k4.md)",
            "kernel": kernel, "kernel_ms": ms, "algorithmic_work_per_launch": amount, "engine": engine,
            "kernel_ms_all": {k: float(v) for k, v in times.items()},
            "note": NOTES.get(name, "")}


# What the profiles show beside each kernel's roofline (profiles/ccm_sweep_E3_k4.md)
NOTES = {
    "rank_scan": "the measured HBM traffic (FETCH_SIZE + WRITE_SIZE, 1.96 GB) is below the algorithmic bytes: the "
                 "prefix entries of one target are read once from HBM and reused from L2 by its 2 000 tasks; "
                 "the scan is bound by VALU issue (83 % active), not by HBM bandwidth",
    "direct_pruned": "VALU 66 % active; the brute-force pair unit counts pairs the triangle bound skips",
    "direct": "VALU 82 % active; with libraries of 10..57 rows the list updates, not the distances, are the work",
    "rank_build": "VALU 57 % active; the pair unit counts the pairs of the pool chunks its box test skips",
}


SMAP_UNIT = ("k(n+1)n + 2kn flops per S-map prediction, n = E_a + 1: one multiply and one add per neighbour for each "
             "entry of the upper triangle of X^T X and of X^T y (the weights' exps and the SVD are not counted)")


def smap_flops(k, Ea):
    n = Ea + 1
    return 2.0 * k * (n * (n + 1) / 2 + n)


def stage_roofline(stages, work):
    """Roofline of the stage with the largest device time of a single-task run: stages = [select, merge,
    S-map] ms (HIP events around the kernels on the context stream); work[name] = (kernel, algorithmic
    flops per launch, engine, peak TFLOP/s, bound)."""
    times = {"select": stages[0], "smap": stages[2]}
    name = max(times, key=lambda n: times[n])
    kernel, flops, engine, peak, bound = work[name]
    ms = float(times[name])
    achieved = flops / (ms * 1e-3) / 1e12
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": None, "kernel": kernel, "kernel_ms": ms, "algorithmic_work_per_launch": flops,
            "engine": engine, "stage": name}


def timed_runs(obj, fe, steps, warmup):
    for _ in range(warmup):
        obj.run()
    st = []
    t0 = time.perf_counter()
    for _ in range(steps):
        obj.run()
        st.append(fe.timings())
    return (time.perf_counter() - t0) / steps, np.mean(np.asarray(st), axis=0)


def median_time(fn, runs=3):
    times = []
    out = None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times, out


def form_line(n_pred, sec, st):
    return {"predictions_per_s": n_pred / sec, "ms_per_step": sec * 1e3, "device_ms": float(st[3]),
            "select_ms": float(st[0]), "merge_ms": float(st[1]), "smap_ms": float(st[2])}


def bits_equal(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return int(np.sum(a.view(np.int64) != b.view(np.int64)))


def cfg2_leg(fe, oracle, cores, steps):
    """S-map theta sweep on the coupled maps, T = 5000, E = 4, k = all library points (2498), six thetas:
    one task of 2499 targets (14 994 predictions) through the plan's resident task route."""
    from tests import configs
    x, _ = configs.coupled_maps(5000)
    t = np.arange(1, 5001, dtype=float)
    kw = dict(E=4, k=np.inf, theta=[0, 0.5, 1, 2, 4, 8], algorithm="smap", savePredictions=True)
    out = {"workload": "smap_theta_sweep_coupled_logistic_T5000_E4_kall_6thetas"}
    raw, res = {}, {}
    for name, prec in (("mixed", fe.MIXED), ("plain", fe.PLAIN)):
        plan = fe.Plan(t, x, precision=prec, **kw)
        sec, st = timed_runs(plan, fe, steps, 2)
        n_pred = plan.num_predictions
        out[name] = form_line(n_pred, sec, st)
        raw[name] = plan.fetch()
        res[name] = plan.result()
        plan.close()
    out["predictions"] = n_pred
    out["ratio_mixed_vs_plain"] = out["mixed"]["predictions_per_s"] / out["plain"]["predictions_per_s"]
    out["note"] = ("k = all takes the dense distance route in both forms: there is no candidate list to "
                   "screen, so the forms run the same FP64 kernels")
    out["predictions_differing_bits_mixed_vs_plain"] = bits_equal(raw["mixed"][0], raw["plain"][0])
    k, Ea = 2498, 4
    st = [out["mixed"]["select_ms"], out["mixed"]["merge_ms"], out["mixed"]["smap_ms"]]
    out["roofline"] = stage_roofline(st, {
        "select": ("dense_dist_f64<8>", 3.0 * Ea * 2498 * 2499, "fp64 VALU", FP64_PEAK_TFLOPS, "valu"),
        "smap": ("dense_means + gram_dense<1,3> (systems) + smap_rows<8,0> (solve)", smap_flops(k, Ea) * n_pred,
                 "fp64 VALU", FP64_PEAK_TFLOPS, "valu")})
    out["roofline"]["unit_of_work"] = SMAP_UNIT
    if oracle is not None:
        dt, times, b = median_time(lambda: oracle.edm(t, x, numThreads=cores, **kw))
        out["cpu_baseline"] = {"value": n_pred / dt, "unit": "predictions/s", "cores": cores, "kind": "port",
                               "runs_s": [round(v, 3) for v in times],
                               "sample": f"the whole command ({n_pred} predictions), median of 3"}
        a_, b_ = res["mixed"]["predictions"], b["predictions"]
        ok = ~np.isnan(b_)
        out["max_abs_d_yhat_mixed_vs_oracle"] = float(np.max(np.abs(a_[ok] - b_[ok])))
        out["max_rel_d_yhat_mixed_vs_oracle"] = float(np.max(np.abs(a_[ok] - b_[ok]) / np.abs(b_[ok])))
        out["max_abs_d_rho_mixed_vs_oracle"] = float(np.nanmax(np.abs(res["mixed"]["stats"]["rho"] -
                                                                      b["stats"]["rho"])))
    return out


def cfg4_leg(fe, oracle, cores, steps):
    """S-map on 512 logistic-map panels of lengths 200..2000 with 5 % missing values, dt on, E = 8 (E_a = 16),
    k = 18, thetas 0, 1, 3: one task of 265 665 library rows and 265 665 targets (796 995 predictions),
    through the plan's resident task route."""
    from tests import configs
    t, x, panel = configs.cfg4_panels()
    kw = dict(panel=panel, E=8, dt=True, theta=[0, 1, 3], algorithm="smap")
    out = {"workload": "panel_smap_512_logistic_panels_E8_dt_k18_3thetas"}
    raw = {}
    for name, prec, st_ in (("mixed", fe.MIXED, steps), ("plain", fe.PLAIN, max(1, steps // 2))):
        plan = fe.Plan(t, x, precision=prec, **kw)
        sec, st = timed_runs(plan, fe, st_, 1)
        n_pred = plan.num_predictions
        out[name] = form_line(n_pred, sec, st)
        out[name]["steps"] = st_
        raw[name] = plan.fetch()
        plan.close()
    out["predictions"] = n_pred
    out["ratio_mixed_vs_plain"] = out["mixed"]["predictions_per_s"] / out["plain"]["predictions_per_s"]
    out["device_ratio_mixed_vs_plain"] = out["plain"]["device_ms"] / out["mixed"]["device_ms"]
    out["predictions_differing_bits_mixed_vs_plain"] = bits_equal(raw["mixed"][0], raw["plain"][0])
    out["rc_and_k_used_equal_mixed_vs_plain"] = bool(np.array_equal(raw["mixed"][1], raw["plain"][1]) and
                                                     np.array_equal(raw["mixed"][2], raw["plain"][2]))
    Ea, k, NL = 16, 18, n_pred // 3
    st = [out["mixed"]["select_ms"], out["mixed"]["merge_ms"], out["mixed"]["smap_ms"]]
    out["roofline"] = stage_roofline(st, {
        "select": ("knn_seed + knn_screen<16> (FP32 matrix cores, seeded floor, k-d ordered library chunks "
                   "skipped by box bounds) + knn_rerank<32,16> (FP64) + knn_mfma<32,16> on the targets that "
                   "overflow; the work unit counts every pair, most of which the boxes skip",
                   3.0 * Ea * float(NL) * NL,
                   "fp32 MFMA", FP32_PEAK_TFLOPS, "mfma"),
        "smap": ("smap_pack<16,1> (Gram + Eigen-order Jacobi SVD, two rows per lane)", smap_flops(k, Ea) * n_pred,
                 "fp64 VALU",
                 FP64_PEAK_TFLOPS, "valu")})
    out["roofline"]["unit_of_work"] = SMAP_UNIT
    if oracle is not None:
        # the oracle's own manifold of the command; a prefix of the targets against the full library
        c, keep, _ = oracle.make_command(t, x, **kw)
        n = len(t)
        usable = np.zeros(n, np.int32)
        nu = oracle.lib().orc_usable(oracle.C.byref(c), 8, 0, oracle.iptr(usable))
        rows = np.nonzero(usable)[0]

        def manifold(sel, ps):
            f = np.zeros(n, np.int32)
            f[sel] = 1
            o_ = np.zeros(n * Ea)
            tg = np.zeros(n)
            m = oracle.lib().orc_manifold(oracle.C.byref(c), 8, oracle.iptr(f), int(ps), 0, oracle.dptr(o_),
                                          oracle.dptr(tg), n)
            return o_[:m * Ea].reshape(m, Ea), tg[:m]

        M, Mt = manifold(rows[:nu // 2], False)
        Mp, _ = manifold(rows[nu // 2:], True)
        m = 20000
        pi = panel.astype(np.int32)
        Md = {"data": M, "targets": Mt, "panel": pi[rows[:nu // 2]]}
        Mpd = {"data": Mp[:m], "panel": pi[rows[nu // 2:]][:m]}
        opts = {"algorithm": 1, "k": k, "thetas": [0.0, 1.0, 3.0], "panel_mode": 1, "nthreads": cores}
        dt, times, r = median_time(lambda: oracle.predict(Md, Mpd, opts))
        out["cpu_baseline"] = {"value": 3 * m / dt, "unit": "predictions/s", "cores": cores, "kind": "port",
                               "runs_s": [round(v, 3) for v in times],
                               "sample": f"the first {m} targets x 3 thetas against the full library (manifold "
                                         f"build not timed), median of 3"}
        got = raw["mixed"][0].reshape(3, -1)[:, :m]
        ref = r[0]
        ok = ref != 1e100
        out["max_abs_d_yhat_mixed_vs_oracle_sample"] = float(np.max(np.abs(got[ok] - ref[ok])))
        out["max_rel_d_yhat_mixed_vs_oracle_sample"] = float(np.max(np.abs(got[ok] - ref[ok]) / np.abs(ref[ok])))
    return out


def cfg5_leg(fe, oracle, cores, steps):
    """S-map on Lorenz-96 (N = 40, F = 8, RK4 dt = 0.01, 100 000 steps): E = 10 lags of x_0 plus 30
    unlagged coordinates (E_a = 40), theta 2, k = 42, 80/20 rows: one task of 79 992 library rows and
    19 998 targets through a resident batch."""
    from tests import configs
    series = configs.lorenz96()
    lib, pred = configs.cfg5_manifold(series)
    nl, npd = lib["data"].shape[0], pred["data"].shape[0]
    table = {key: np.concatenate([lib[key], pred[key]]) for key in ("data", "targets")}
    table.update(E_x=10, E_extras=30)
    opts = {"algorithm": 1, "k": 42, "thetas": [2.0]}
    out = {"workload": "lorenz96_N40_smap_E10_plus30extras_k42_theta2"}
    raw = {}
    for name, prec in (("mixed", fe.MIXED), ("plain", fe.PLAIN)):
        b = fe.Batch(table, table, dict(opts, precision=prec), [np.arange(nl)], [np.arange(nl, nl + npd)], [42])
        assert b.route == 1
        sec, st = timed_runs(b, fe, steps, 1)
        out[name] = form_line(npd, sec, st)
        raw[name] = b.fetch()[0]
        b.close()
    out["predictions"] = npd
    out["ratio_mixed_vs_plain"] = out["mixed"]["predictions_per_s"] / out["plain"]["predictions_per_s"]
    out["predictions_differing_bits_mixed_vs_plain"] = bits_equal(raw["mixed"][0], raw["plain"][0])
    Ea, k = 40, 42
    st = [out["mixed"]["select_ms"], out["mixed"]["merge_ms"], out["mixed"]["smap_ms"]]
    out["roofline"] = stage_roofline(st, {
        "select": ("knn_seed + knn_mfma<64,64> (FP32 matrix cores from a seeded floor, FP64 re-rank)",
                   3.0 * Ea * float(nl) * npd, "fp32 MFMA",
                   FP32_PEAK_TFLOPS, "mfma"),
        "smap": ("smap_hyb (Gram + Eigen-order Jacobi SVD, n = 41)", smap_flops(k, Ea) * npd, "fp64 VALU",
                 FP64_PEAK_TFLOPS, "valu")})
    out["roofline"]["unit_of_work"] = SMAP_UNIT
    if oracle is not None:
        m = npd
        dt, times, r = median_time(lambda: oracle.predict(lib, {"data": pred["data"][:m]}, dict(opts, nthreads=cores)))
        out["cpu_baseline"] = {"value": m / dt, "unit": "predictions/s", "cores": cores, "kind": "port",
                               "runs_s": [round(v, 3) for v in times],
                               "sample": f"every target ({m}) against the full library (the whole task, manifold "
                                         f"build not timed), median of 3"}
        got = raw["mixed"][0][0, :m]
        ref = r[0][0]
        out["max_abs_d_yhat_mixed_vs_oracle_sample"] = float(np.max(np.abs(got - ref)))
        out["max_rel_d_yhat_mixed_vs_oracle_sample"] = float(np.max(np.abs(got - ref) / np.abs(ref)))
    return out


def graph_leg(fe, t, x, y, kw, steps):
    """The headline sweep with graph replay on: the second and later runs replay the captured launches."""
    plan = fe.Plan(t, x, y, precision=fe.MIXED, **kw)
    plan.run()
    ref = plan.fetch()[0].copy()
    plan.set_graph(True)
    for _ in range(2):
        plan.run()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    sec = (time.perf_counter() - t0) / steps
    got = plan.fetch()[0]
    plan.close()
    return {"ms_per_step": sec * 1e3, "predictions_per_s": got.size / sec,
            "predictions_differing_bits_vs_eager": bits_equal(got, ref)}


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


def host_cores():
    """Threads for the CPU baseline: the host's cores, capped by the box's thread allowance."""
    return max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "1024"))))


def cpu_baseline(oracle, flags, t, x, y, kw, seconds):
    """The oracle's edm() (tasks serial, targets over a thread per host core, as the reference CPU path
    runs them) on the first replicates of the sweep: a sample sized to about `seconds`, timed three
    times, median reported.  The timed call includes the oracle's manifold builds and rho/MAE."""
    cores = host_cores()
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
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--configs", default="2,4,5",
                    help="other workloads to measure beside the CCM sweep (comma list of 2, 4, 5; empty: none)")
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
                              "over every long-library task) for libraries >= 1.5 sqrt(k n), exact FP64 direct "
                              "kernels for shorter ones (spatially pruned from 64 rows)"}}

    out["stages_ms"] = {"device_total": float(st[0]), "rank_build": float(st[1]), "rank_scan": float(st[2]),
                        "direct": float(st[3]), "direct_kernel": float(st[4]), "direct_pruned_kernel": float(st[5]),
                        "exact_reselection": float(st[6])}
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
        out["ratio_mixed_vs_plain_batched"] = out["value"] / out["plain_batched"]["value"]
        out["ratio_note"] = ("plain is the drop-in (one edm_predict per task, host gather and upload per task); "
                             "plain_batched runs every library through the FP64 direct kernel in the persistent "
                             "launch.  Most of the mixed route's lead comes from its algorithm: one exact ranked "
                             "prefix per target serves all 2 000 long-library tasks, and an FP64 ranking would "
                             "give the plain form the same reuse; the FP32 screen saves only the ranking's "
                             "distance work")
        out["parity"]["max_abs_d_rho_plain_taskwise_vs_plain_batched"] = float(
            np.nanmax(np.abs(res_pt["stats"]["rho"] - res_pb["stats"]["rho"])))
        pt.close()
    pb.close()
    mixed.close()

    print("graph replay of the sweep", file=sys.stderr, flush=True)
    out["graph_replay"] = graph_leg(fe, t, x, y, kw, args.steps)

    oracle_mod, flags = (None, None) if args.no_cpu else baseline_oracle()
    cores = host_cores()
    legs = {"2": ("cfg2_smap_theta_sweep", cfg2_leg), "4": ("cfg4_panel_smap", cfg4_leg),
            "5": ("cfg5_lorenz96_smap", cfg5_leg)}
    out["configs"] = {}
    for key in [v for v in args.configs.split(",") if v]:
        name, fn = legs[key]
        print(f"{name}: mixed and plain", file=sys.stderr, flush=True)
        out["configs"][name] = fn(fe, oracle_mod, cores, 5)
        if oracle_mod is not None:
            out["configs"][name]["cpu_baseline"]["build"] = flags

    if not args.no_cpu:
        print("timing the CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"], res_cpu, reps = cpu_baseline(oracle_mod, flags, t, x, y, kw, args.cpu_seconds)
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
