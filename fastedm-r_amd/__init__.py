"""fastEDM prediction core on AMD MI355X (gfx950).

Python mirror of the reference's R entry point ``edm()`` (R/edm.R:232-393) over the C ABI of
``libfastedm_hip.so`` (include/fastedm_hip.h).  All predictions are made by the HIP kernels; if the
library or a GPU is missing, every call raises — there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastedm_hip.so")
MISSING_D = 1.0e100
MISSING_I = -(2**31)
PLAIN, MIXED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
KEEP_GOING = C.CFUNCTYPE(C.c_int)


class ManifoldView(C.Structure):
    """edm_manifold_view"""
    _fields_ = [("data", _dp), ("targets", _dp), ("point_times", _dp), ("target_times", _dp), ("panel", _ip),
                ("num_points", C.c_int), ("E_x", C.c_int), ("E_dt", C.c_int), ("E_extras", C.c_int),
                ("E_lagged_extras", C.c_int), ("E_actual", C.c_int)]


class Knobs(C.Structure):
    """fedm_knobs: launch and route choices that leave every result unchanged."""
    _fields_ = [("rank_scale", C.c_double), ("direct_blocks_per_cu", C.c_int), ("direct_alt_waves", C.c_int),
                ("list_blocks", C.c_int), ("mfma_blocks", C.c_int), ("use_mfma", C.c_int), ("smap_lds", C.c_int),
                ("direct_prune", C.c_int), ("concurrent_routes", C.c_int), ("smap_block", C.c_int),
                ("mfma_seed", C.c_int), ("mfma_prune", C.c_int),
                ("mfma_split", C.c_int)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class Options(C.Structure):
    """edm_options"""
    _fields_ = [("algorithm", C.c_int), ("distance", C.c_int), ("k", C.c_int), ("num_thetas", C.c_int),
                ("thetas", _dp), ("metrics", _ip), ("missing_distance", C.c_double), ("panel_mode", C.c_int),
                ("idw", C.c_double), ("num_panels", C.c_int), ("id_weights", _fp), ("has_missing", C.c_int),
                ("has_categorical", C.c_int), ("use_only_past", C.c_int), ("force_compute", C.c_int),
                ("save_smap_coeffs", C.c_int), ("low_memory", C.c_int), ("aspect_ratio", C.c_double),
                ("precision", C.c_int), ("nthreads", C.c_int)]


class Command(C.Structure):
    """fedm_command"""
    _fields_ = [("n", C.c_int), ("t", _dp), ("x", _dp), ("y", _dp), ("co_x", _dp), ("panel", _ip),
                ("num_extras", C.c_int), ("extras", _dp), ("num_E", C.c_int), ("E", _ip), ("tau", C.c_int),
                ("num_thetas", C.c_int), ("thetas", _dp), ("num_libraries", C.c_int), ("libraries", _ip),
                ("k", C.c_int), ("algorithm", C.c_int), ("num_reps", C.c_int), ("p", C.c_int),
                ("crossfold", C.c_int), ("full", C.c_int), ("shuffle", C.c_int), ("seed", C.c_int),
                ("save_final_predictions", C.c_int), ("save_final_copredictions", C.c_int),
                ("save_manifolds", C.c_int), ("save_smap_coeffs", C.c_int), ("dt", C.c_int), ("reldt", C.c_int),
                ("dt_weight", C.c_double), ("allow_missing", C.c_int), ("missing_distance", C.c_double),
                ("panel_weight", C.c_double), ("panel_weights", _dp), ("panel_weights_dim", C.c_int),
                ("num_threads", C.c_int), ("low_memory", C.c_int), ("predict_with_past", C.c_int),
                ("distance", C.c_int), ("aspect_ratio", C.c_double), ("precision", C.c_int),
                ("keep_going", C.c_void_p)]


EXPORTS = ["edm_ctx_create", "edm_ctx_destroy", "edm_predict", "edm_predict_batch", "edm_ctx_timings",
           "edm_ctx_exact_redos",
           "edm_batch_create", "edm_batch_run", "edm_batch_fetch", "edm_batch_destroy", "fedm_plan_create",
           "fedm_plan_run", "fedm_plan_result", "fedm_plan_num_predictions", "fedm_plan_device_batches", "fedm_plan_destroy",
           "fedm_plan_num_tasks", "fedm_plan_task", "fedm_plan_fetch", "edm_batch_set_keep_going",
           "edm_batch_route", "edm_batch_set_graph", "edm_batch_fetch_coeffs", "fedm_plan_set_graph",
           "fedm_plan_resident_batches", "edm_predict_noalloc", "edm_selftest_exp", "fedm_default_knobs",
           "edm_batch_create_knobs", "edm_batch_knobs", "fedm_plan_tune", "fedm_plan_group_knobs",
           "edm_ctx_timings_ex", "edm_batch_on_device", "edm_selftest_mfma", "fedm_rank_threshold", "fedm_run_command", "fedm_run_task_group", "fedm_result_rc", "fedm_result_num_stats", "fedm_result_stats", "fedm_result_k",
           "fedm_result_scalar", "fedm_result_matrix_shape", "fedm_result_matrix", "fedm_result_num_manifolds",
           "fedm_result_manifold_shape", "fedm_result_manifold", "fedm_result_free"]

_lib = None
_ctx = None


class FastEDMError(RuntimeError):
    pass


def load_library():
    """Loads libfastedm_hip.so (built in-tree by build.py); raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FastEDMError(f"{LIB_PATH} is not built; run fastedm-r_amd/build.py")
        L = C.CDLL(LIB_PATH)
        L.fedm_run_command.argtypes = [C.c_void_p, C.POINTER(Command), C.POINTER(C.c_void_p)]
        L.edm_ctx_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        L.edm_ctx_destroy.argtypes = [C.c_void_p]
        L.edm_ctx_timings.argtypes = [C.c_void_p, _dp]
        L.edm_predict.argtypes = [C.c_void_p, C.POINTER(ManifoldView), C.POINTER(ManifoldView), C.POINTER(Options),
                                  _dp, _ip, _dp, _ip, _ip, C.c_int, C.c_void_p]
        L.edm_predict_noalloc.argtypes = [C.c_void_p, C.POINTER(ManifoldView), C.POINTER(ManifoldView),
                                          C.POINTER(Options), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.edm_predict_batch.argtypes = [C.c_void_p, C.POINTER(ManifoldView), C.POINTER(ManifoldView),
                                        C.POINTER(Options), C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _i64p, _dp,
                                        _ip, _ip, C.c_void_p]
        for name in ["fedm_result_rc", "fedm_result_num_stats", "fedm_result_matrix_shape",
                     "fedm_result_num_manifolds"]:
            getattr(L, name).restype = C.c_int
        L.fedm_result_scalar.restype = C.c_double
        L.fedm_plan_create.argtypes = [C.c_void_p, C.POINTER(Command), C.POINTER(C.c_void_p)]
        L.fedm_plan_run.argtypes = [C.c_void_p]
        L.fedm_plan_result.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.fedm_plan_num_predictions.argtypes = [C.c_void_p]
        L.fedm_plan_num_predictions.restype = C.c_int64
        L.fedm_plan_destroy.argtypes = [C.c_void_p]
        L.fedm_plan_device_batches.argtypes = [C.c_void_p]
        L.fedm_plan_num_tasks.argtypes = [C.c_void_p]
        L.fedm_plan_task.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip]
        L.fedm_plan_fetch.argtypes = [C.c_void_p, _dp, _ip, _ip]
        L.fedm_plan_set_graph.argtypes = [C.c_void_p, C.c_int]
        L.fedm_plan_resident_batches.argtypes = [C.c_void_p]
        for name in ["fedm_result_rc", "fedm_result_num_stats", "fedm_result_stats", "fedm_result_k",
                     "fedm_result_scalar", "fedm_result_matrix_shape", "fedm_result_matrix",
                     "fedm_result_num_manifolds", "fedm_result_manifold_shape", "fedm_result_manifold",
                     "fedm_result_free"]:
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def context():
    """The process-wide device context on cuda:0 (one HIP stream, pooled device buffers)."""
    global _ctx
    if _ctx is None:
        L = load_library()
        h = C.c_void_p()
        rc = L.edm_ctx_create(C.byref(h), 0)
        if rc != 0:
            raise FastEDMError(f"edm_ctx_create failed ({rc}): no usable GPU")
        _ctx = h
    return _ctx


def _d(a, keep):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data_as(_dp)


def _i(a, keep):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.int32)
    keep.append(a)
    return a.ctypes.data_as(_ip)


def _f(a, keep):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    keep.append(a)
    return a.ctypes.data_as(_fp)


def _isnothing(v):
    return v is None or (isinstance(v, float) and np.isnan(v))


def _command(t, x, y, panel, E, tau, theta, library, k, algorithm, p, crossfold, full, shuffle, seed, copredict,
             savePredictions, saveCoPredictions, saveManifolds, saveSMAPCoeffs, extras, allowMissing,
             missingDistance, dt, reldt, dtWeight, numReps, panelWeight, panelWeights, numThreads, lowMemory,
             predictWithPast, distance, aspectRatio, precision, keep, keep_going=None):
    n = len(t)
    if len(x) != n:
        raise ValueError("The time and x variables should be the same length")
    if y is not None and len(y) and len(y) != n:
        raise ValueError("The y variable is the wrong length")
    if panel is not None and len(panel) and len(panel) != n:
        raise ValueError("The panel id variable is the wrong length")
    if copredict is not None and len(copredict) and len(copredict) != n:
        raise ValueError("Coprediction vector is the wrong length")
    for e in extras or []:
        if len(e) != n:
            raise ValueError("An extra variable is not the right size")
    y = None if y is None or len(y) == 0 else y
    panel = None if panel is None or len(panel) == 0 else panel
    copredict = None if copredict is None or len(copredict) == 0 else copredict
    # R/edm.R:277-322 defaults
    if numReps > 1:
        shuffle = True
    explore = y is None
    p = (1 if explore else 0) if p is None else int(p)
    k = 0 if _isnothing(k) else k
    Es = np.atleast_1d(np.asarray(E, dtype=np.float64))
    Es = (np.array([2.0]) if np.isnan(Es[0]) else Es) if Es.size == 1 else Es[~np.isnan(Es)]
    k = -1 if k == np.inf else int(k)
    panelWeight = -1.0 if panelWeight == np.inf else float(panelWeight)
    thetas = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    c = Command()
    c.n = n
    c.t = _d(t, keep)
    c.x = _d(x, keep)
    c.y = _d(y, keep)
    c.co_x = _d(copredict, keep)
    if panel is not None:
        pa = np.asarray(panel, dtype=np.float64)
        c.panel = _i(np.where(np.isnan(pa), MISSING_I, pa).astype(np.int64), keep)
    ex = list(extras or [])
    c.num_extras = len(ex)
    c.extras = _d(np.concatenate([np.asarray(e, np.float64) for e in ex]) if ex else None, keep)
    c.num_E = Es.size
    c.E = _i(Es.astype(np.int64), keep)
    c.tau = int(tau)
    c.num_thetas = thetas.size
    c.thetas = _d(thetas, keep)
    if library is not None:
        libs = np.atleast_1d(np.asarray(library)).astype(np.int64)
        c.num_libraries = libs.size
        c.libraries = _i(libs, keep)
    c.k = k
    c.algorithm = {"simplex": 0, "smap": 1}.get(algorithm, 9)
    c.num_reps = int(numReps)
    c.p = p
    c.crossfold = int(crossfold)
    c.full = int(bool(full))
    c.shuffle = int(bool(shuffle))
    c.seed = int(seed)
    c.save_final_predictions = int(bool(savePredictions))
    c.save_final_copredictions = int(bool(saveCoPredictions))
    c.save_manifolds = int(bool(saveManifolds))
    c.save_smap_coeffs = int(bool(saveSMAPCoeffs))
    c.dt = int(bool(dt))
    c.reldt = int(bool(reldt))
    c.dt_weight = float(dtWeight)
    c.allow_missing = int(bool(allowMissing))
    c.missing_distance = float(missingDistance)
    c.panel_weight = panelWeight
    if panelWeights is not None:
        pw = np.asarray(panelWeights, np.float64)
        c.panel_weights = _d(pw.ravel(), keep)
        c.panel_weights_dim = pw.shape[0]
    c.num_threads = int(numThreads)
    c.low_memory = int(bool(lowMemory))
    c.predict_with_past = int(bool(predictWithPast))
    c.distance = {"mae": 0, "euclidean": 1, "wasserstein": 2}[distance]
    c.aspect_ratio = float(aspectRatio)
    c.precision = int(precision)
    if keep_going is not None:
        cb = KEEP_GOING(keep_going)
        keep.append(cb)
        c.keep_going = C.cast(cb, C.c_void_p)
    return c


def _na(a):
    a = np.asarray(a, dtype=np.float64)
    return np.where(a == MISSING_D, np.nan, a)


def _aggregate(st):
    """stats::aggregate(cbind(rho, mae) ~ E + library + theta, na.omit(stats), mean)."""
    ok = ~(np.isnan(st["rho"]) | np.isnan(st["mae"]))
    groups = {}
    for i in np.nonzero(ok)[0]:
        groups.setdefault((float(st["theta"][i]), int(st["library"][i]), int(st["E"][i])), []).append(i)
    keys = sorted(groups)
    return {"E": np.array([k[2] for k in keys], dtype=np.int64),
            "library": np.array([k[1] for k in keys], dtype=np.int64),
            "theta": np.array([k[0] for k in keys]),
            "rho": np.array([np.mean(st["rho"][groups[k]]) for k in keys]),
            "mae": np.array([np.mean(st["mae"][groups[k]]) for k in keys])}


def edm(t, x, y=None, panel=None, E=2, tau=1, theta=1, library=None, k=0, algorithm="simplex", p=None, crossfold=0,
        full=False, shuffle=False, seed=0, copredict=None, savePredictions=False, saveCoPredictions=False,
        saveManifolds=False, saveSMAPCoeffs=False, extras=None, allowMissing=False, missingDistance=0.0,
        dt=False, reldt=False, dtWeight=0.0, numReps=1, panelWeight=0, panelWeights=None, verbosity=0,
        showProgressBar=None, numThreads=1, lowMemory=False, predictWithPast=False, distance="euclidean",
        aspectRatio=1.0, precision=PLAIN, keep_going=None):
    """The reference's edm() (R/edm.R:232), predictions made on the GPU.

    Returns a dict shaped like R's result list: rc, stats (E, library, theta, rho, mae; NaN = NA),
    summary, kMin, kMax and, when requested, copredStats, predictions, copredictions, coeffs, Ms, Mps,
    missingdistance, dtWeight.  ``precision`` selects the plain FP64 or the mixed form.  ``keep_going``
    (a function returning a truth value, or None) plays the part of R's interrupt check: once it says
    stop, the call returns with rc 1 (BREAK_HIT) and no statistics.
    """
    L = load_library()
    keep = []
    c = _command(t, x, y, panel, E, tau, theta, library, k, algorithm, p, crossfold, full, shuffle, seed,
                 copredict, savePredictions, saveCoPredictions, saveManifolds, saveSMAPCoeffs, extras, allowMissing,
                 missingDistance, dt, reldt, dtWeight, numReps, panelWeight, panelWeights, numThreads, lowMemory,
                 predictWithPast, distance, aspectRatio, precision, keep, keep_going)
    h = C.c_void_p()
    status = L.fedm_run_command(context(), C.byref(c), C.byref(h))
    if status < 0:
        raise FastEDMError(f"fedm_run_command failed ({status})")
    try:
        return _collect(L, h, copredict is not None and len(copredict) > 0, allowMissing, dt or reldt)
    finally:
        L.fedm_result_free(h)


def _collect(L, h, copred, allow_missing, dt):
    res = {"rc": L.fedm_result_rc(h)}

    def stats(co):
        m = L.fedm_result_num_stats(h, co)
        E = np.zeros(m, np.int32)
        lib_ = np.zeros(m, np.int32)
        th, rho, mae = np.zeros(m), np.zeros(m), np.zeros(m)
        L.fedm_result_stats(h, co, E.ctypes.data_as(_ip), lib_.ctypes.data_as(_ip), th.ctypes.data_as(_dp),
                            rho.ctypes.data_as(_dp), mae.ctypes.data_as(_dp))
        return {"E": E, "library": lib_, "theta": th, "rho": _na(rho), "mae": _na(mae)}

    res["stats"] = stats(0)
    if copred:
        res["copredStats"] = stats(1)
    kmin, kmax = C.c_int(), C.c_int()
    L.fedm_result_k(h, C.byref(kmin), C.byref(kmax))
    res["kMin"] = None if kmin.value == MISSING_I else kmin.value
    res["kMax"] = None if kmax.value == MISSING_I else kmax.value
    for which, name in [(0, "predictions"), (1, "copredictions"), (2, "coeffs")]:
        rows, cols = C.c_int(), C.c_int()
        if L.fedm_result_matrix_shape(h, which, C.byref(rows), C.byref(cols)):
            out = np.zeros(rows.value * cols.value)
            L.fedm_result_matrix(h, which, out.ctypes.data_as(_dp))
            res[name] = _na(out.reshape(cols.value, rows.value).T)
    nm = L.fedm_result_num_manifolds(h)
    if nm:
        res["Ms"], res["Mps"] = [], []
        for i in range(nm):
            for ps, key in [(0, "Ms"), (1, "Mps")]:
                rows, cols = C.c_int(), C.c_int()
                L.fedm_result_manifold_shape(h, i, ps, C.byref(rows), C.byref(cols))
                out = np.zeros(rows.value * cols.value)
                L.fedm_result_manifold(h, i, ps, out.ctypes.data_as(_dp))
                res[key].append(_na(out.reshape(rows.value, cols.value)))
    if allow_missing:
        res["missingdistance"] = L.fedm_result_scalar(h, 0)
    if dt:
        res["dtWeight"] = L.fedm_result_scalar(h, 1)
    res["summary"] = _aggregate(res["stats"])
    return res


def manifold_view(m, keep):
    """dict(data=[N x Ea], targets, point_times, target_times, panel, E_x, E_dt, E_extras) -> view."""
    v = ManifoldView()
    data = np.ascontiguousarray(m["data"], np.float64)
    n, Ea = data.shape
    keep.append(data)
    v.data = data.ctypes.data_as(_dp)
    v.targets = _d(m.get("targets", np.zeros(n)), keep)
    v.point_times = _d(m.get("point_times", np.zeros(n)), keep)
    v.target_times = _d(m.get("target_times", np.zeros(n)), keep)
    v.panel = _i(m.get("panel"), keep)
    v.num_points = n
    v.E_x = m.get("E_x", Ea)
    v.E_dt = m.get("E_dt", 0)
    v.E_extras = m.get("E_extras", 0)
    v.E_lagged_extras = m.get("E_lagged_extras", 0)
    v.E_actual = Ea
    return v


def options(d, keep):
    o = Options()
    th = np.atleast_1d(np.asarray(d["thetas"], np.float64))
    o.algorithm = d.get("algorithm", 0)
    o.distance = d.get("distance", 1)
    o.k = d["k"]
    o.num_thetas = th.size
    o.thetas = _d(th, keep)
    o.metrics = _i(d.get("metrics"), keep)
    o.missing_distance = d.get("missing_distance", 0.0)
    o.panel_mode = d.get("panel_mode", 0)
    o.idw = d.get("idw", 0.0)
    if d.get("id_weights") is not None:
        w = np.asarray(d["id_weights"], np.float32)
        o.id_weights = _f(w, keep)
        o.num_panels = w.shape[0]
    o.has_missing = d.get("has_missing", 0)
    o.has_categorical = d.get("has_categorical", 0)
    o.use_only_past = d.get("use_only_past", 0)
    o.force_compute = d.get("force_compute", 1)
    o.save_smap_coeffs = d.get("save_smap_coeffs", 0)
    o.low_memory = d.get("low_memory", 0)
    o.aspect_ratio = d.get("aspect_ratio", 1.0)
    o.precision = d.get("precision", PLAIN)
    o.nthreads = 1
    return o


def predict(M, Mp, opts, nn_cap=0, keep_going=None):
    """Task-level call of edm_predict: returns (pred [T x Np], rc [T x Np], coeffs [Np x (Ea+1)], k_used, nn_idx)."""
    L = load_library()
    keep = []
    mv, mpv, o = manifold_view(M, keep), manifold_view(Mp, keep), options(opts, keep)
    Np, Ea = np.shape(Mp["data"])
    T = int(np.atleast_1d(opts["thetas"]).size)
    pred = np.zeros(T * Np)
    rc = np.zeros(T * Np, np.int32)
    coeffs = np.zeros(Np * (Ea + 1))
    ku = np.zeros(Np, np.int32)
    nn = np.zeros(max(1, Np * nn_cap), np.int32)
    cb = KEEP_GOING(keep_going) if keep_going is not None else None
    status = L.edm_predict(context(), C.byref(mv), C.byref(mpv), C.byref(o), pred.ctypes.data_as(_dp),
                           rc.ctypes.data_as(_ip), coeffs.ctypes.data_as(_dp), ku.ctypes.data_as(_ip),
                           nn.ctypes.data_as(_ip) if nn_cap else None, nn_cap,
                           C.cast(cb, C.c_void_p) if cb is not None else None)
    if status < 0:
        raise FastEDMError(f"edm_predict failed ({status})")
    return (pred.reshape(T, Np), rc.reshape(T, Np), coeffs.reshape(Np, Ea + 1), ku,
            nn[:Np * nn_cap].reshape(Np, nn_cap) if nn_cap else None, status)


def timings():
    """Device time of the last call's stages in ms: [distance/select, merge/simplex, S-map, total]."""
    out = np.zeros(4)
    load_library().edm_ctx_timings(context(), out.ctypes.data_as(_dp))
    return out


def timings_ex():
    """timings() followed by the persistent route's direct stage split: [4] direct kernel, [5] pruned
    direct kernel, [6] exact re-selection, [7] 1 when the direct kernels ran beside the rank-table
    kernels on the second stream."""
    out = np.zeros(8)
    f = load_library().edm_ctx_timings_ex
    f.argtypes = [C.c_void_p, _dp]
    f(context(), out.ctypes.data_as(_dp))
    return out


def exact_redos():
    """(task, target) pairs the last batched run re-selected with the sqrt-ordered list."""
    f = load_library().edm_ctx_exact_redos
    f.argtypes = [C.c_void_p]
    return int(f(context()))


class Plan:
    """A command prepared for repeated device runs (fedm_plan_*): create once, run() any number of
    times with inputs resident on the GPU, then result() for R's result list."""

    def __init__(self, t, x, y=None, **kw):
        L = load_library()
        self._keep = []
        args = dict(panel=None, E=2, tau=1, theta=1, library=None, k=0, algorithm="simplex", p=None, crossfold=0,
                    full=False, shuffle=False, seed=0, copredict=None, savePredictions=False,
                    saveCoPredictions=False, saveManifolds=False, saveSMAPCoeffs=False, extras=None,
                    allowMissing=False, missingDistance=0.0, dt=False, reldt=False, dtWeight=0.0, numReps=1,
                    panelWeight=0, panelWeights=None, numThreads=1, lowMemory=False, predictWithPast=False,
                    distance="euclidean", aspectRatio=1.0, precision=PLAIN, keep_going=None)
        args.update(kw)
        self._flags = (args["copredict"] is not None and len(args["copredict"]) > 0, args["allowMissing"],
                       args["dt"] or args["reldt"])
        self._c = _command(t, x, y, keep=self._keep, **args)
        self._h = C.c_void_p()
        status = L.fedm_plan_create(context(), C.byref(self._c), C.byref(self._h))
        if status != 0:
            raise FastEDMError(f"fedm_plan_create failed ({status})")

    @property
    def num_predictions(self):
        return int(load_library().fedm_plan_num_predictions(self._h))

    def tune(self, trials=3, report_len=1 << 16):
        """The autotuner over every group's knobs (fedm_plan_tune): returns (groups tuned, report)."""
        L = load_library()
        L.fedm_plan_tune.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        buf = C.create_string_buffer(report_len)
        n = L.fedm_plan_tune(self._h, int(trials), buf, report_len)
        if n < 0:
            raise FastEDMError(f"fedm_plan_tune failed ({n})")
        return n, buf.value.decode()

    def knobs(self, group=0):
        L = load_library()
        L.fedm_plan_group_knobs.argtypes = [C.c_void_p, C.c_int, C.POINTER(Knobs)]
        k = Knobs()
        if L.fedm_plan_group_knobs(self._h, int(group), C.byref(k)) != 0:
            return None
        return k.as_dict()

    def set_graph(self, on=True):
        """Replays each group's launches as one captured hipGraph from the second run on; returns the
        number of groups that do."""
        return int(load_library().fedm_plan_set_graph(self._h, int(bool(on))))

    @property
    def resident_batches(self):
        """Number of groups on the resident task route (kernels per task, inputs uploaded once)."""
        return int(load_library().fedm_plan_resident_batches(self._h))

    @property
    def device_batches(self):
        """Number of (E, copredict) groups that run as one persistent device launch."""
        return int(load_library().fedm_plan_device_batches(self._h))

    def run(self):
        """One device run of every task; returns 0, or 1 (BREAK_HIT) when keep_going stopped it."""
        status = load_library().fedm_plan_run(self._h)
        if status < 0 or status > 1:
            raise FastEDMError(f"fedm_plan_run failed ({status})")
        return status

    def tasks(self, rows=False, indices=None):
        """The plan's tasks in the reference's order (or those at ``indices``): dicts with E_actual,
        library, k, copredict, E and the numbers of library and prediction rows (and, with rows=True,
        the table rows themselves)."""
        L = load_library()
        out = []
        info = np.zeros(7, np.int32)
        if indices is None:
            indices = range(int(L.fedm_plan_num_tasks(self._h)))
        for i in indices:
            L.fedm_plan_task(self._h, i, info.ctypes.data_as(_ip), None, None)
            d = {"E_actual": int(info[0]), "library": int(info[1]), "k": int(info[2]), "num_lib": int(info[3]),
                 "num_pred": int(info[4]), "copredict": bool(info[5]), "E": int(info[6])}
            if rows:
                lr = np.zeros(max(d["num_lib"], 1), np.int32)
                pr = np.zeros(max(d["num_pred"], 1), np.int32)
                L.fedm_plan_task(self._h, i, info.ctypes.data_as(_ip), lr.ctypes.data_as(_ip), pr.ctypes.data_as(_ip))
                d["lib_rows"], d["pred_rows"] = lr[:d["num_lib"]], pr[:d["num_pred"]]
            out.append(d)
        return out

    def fetch(self):
        """Every task's outputs of the last run, concatenated in task order: pred and rc ([T x Np] per
        task, flattened) and k_used ([Np] per task)."""
        L = load_library()
        T = max(int(self._c.num_thetas), 1)
        n = self.num_predictions
        pred = np.zeros(max(n, 1))
        rc = np.zeros(max(n, 1), np.int32)
        ku = np.zeros(max(n // T, 1), np.int32)
        status = L.fedm_plan_fetch(self._h, pred.ctypes.data_as(_dp), rc.ctypes.data_as(_ip), ku.ctypes.data_as(_ip))
        if status != 0:
            raise FastEDMError(f"fedm_plan_fetch failed ({status})")
        return pred[:n], rc[:n], ku[:n // T]

    def result(self):
        L = load_library()
        h = C.c_void_p()
        status = L.fedm_plan_result(self._h, C.byref(h))
        if status != 0:
            raise FastEDMError(f"fedm_plan_result failed ({status})")
        try:
            return _collect(L, h, *self._flags)
        finally:
            L.fedm_result_free(h)

    def close(self):
        if self._h:
            load_library().fedm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A batch prepared for repeated device runs (edm_batch_*): tasks over shared tables L and P, task t
    using library rows lib_lists[t] and prediction rows pred_lists[t].  run() makes every prediction
    with the inputs resident on the GPU; fetch() returns per task (pred [T x Np], rc [T x Np],
    k_used [Np]); route is 2 (persistent launch), 1 (resident task route) or 0 (task by task)."""

    def __init__(self, L, P, opts, lib_lists, pred_lists, ks):
        lib = load_library()
        lib.edm_batch_create.argtypes = [C.c_void_p, C.POINTER(ManifoldView), C.POINTER(ManifoldView),
                                         C.POINTER(Options), C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _i64p,
                                         C.POINTER(C.c_void_p)]
        for name in ("edm_batch_run", "edm_batch_destroy", "edm_batch_route"):
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.edm_batch_fetch.argtypes = [C.c_void_p, _dp, _ip, _ip]
        lib.edm_batch_fetch_coeffs.argtypes = [C.c_void_p, _dp]
        lib.edm_batch_set_graph.argtypes = [C.c_void_p, C.c_int]
        self._keep = []
        keep = self._keep
        Lv, Pv, o = manifold_view(L, keep), manifold_view(P, keep), options(opts, keep)
        self.T = int(np.atleast_1d(opts["thetas"]).size)
        self.Ea = int(np.shape(L["data"])[1])
        self.counts = [len(a) for a in pred_lists]
        lb = np.cumsum([0] + [len(a) for a in lib_lists[:-1]]).astype(np.int32)
        lc = np.array([len(a) for a in lib_lists], np.int32)
        lr = np.concatenate([np.asarray(a, np.int32) for a in lib_lists])
        pb = np.cumsum([0] + self.counts[:-1]).astype(np.int32)
        pc = np.array(self.counts, np.int32)
        pr = np.concatenate([np.asarray(a, np.int32) for a in pred_lists])
        self.ob = np.cumsum([0] + [self.T * n for n in self.counts[:-1]]).astype(np.int64)
        self.total = int(self.T * pc.sum())
        k = np.asarray(ks, np.int32)
        keep += [lb, lc, lr, pb, pc, pr, k, self.ob]
        self._h = C.c_void_p()
        status = lib.edm_batch_create(context(), C.byref(Lv), C.byref(Pv), C.byref(o), len(lib_lists),
                                      lb.ctypes.data_as(_ip), lc.ctypes.data_as(_ip), lr.ctypes.data_as(_ip),
                                      pb.ctypes.data_as(_ip), pc.ctypes.data_as(_ip), pr.ctypes.data_as(_ip),
                                      k.ctypes.data_as(_ip), self.ob.ctypes.data_as(_i64p), C.byref(self._h))
        if status != 0:
            raise FastEDMError(f"edm_batch_create failed ({status})")

    @property
    def route(self):
        return int(load_library().edm_batch_route(self._h))

    def set_graph(self, on=True):
        return int(load_library().edm_batch_set_graph(self._h, int(bool(on))))

    def run(self):
        status = load_library().edm_batch_run(self._h)
        if status != 0:
            raise FastEDMError(f"edm_batch_run failed ({status})")

    def fetch(self):
        pred = np.zeros(max(self.total, 1))
        rc = np.zeros(max(self.total, 1), np.int32)
        ku = np.zeros(max(self.total // self.T, 1), np.int32)
        status = load_library().edm_batch_fetch(self._h, pred.ctypes.data_as(_dp), rc.ctypes.data_as(_ip),
                                                ku.ctypes.data_as(_ip))
        if status != 0:
            raise FastEDMError(f"edm_batch_fetch failed ({status})")
        out = []
        for t, n in enumerate(self.counts):
            o0 = int(self.ob[t])
            out.append((pred[o0:o0 + self.T * n].reshape(self.T, n), rc[o0:o0 + self.T * n].reshape(self.T, n),
                        ku[o0 // self.T:o0 // self.T + n]))
        return out

    def close(self):
        if self._h:
            load_library().edm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def predict_batch(L, P, opts, lib_lists, pred_lists, ks):
    """edm_predict_batch over shared tables L and P: task t uses library rows lib_lists[t] and prediction
    rows pred_lists[t].  Returns per task (pred [T x Np], rc [T x Np], k_used [Np])."""
    lib = load_library()
    keep = []
    Lv, Pv, o = manifold_view(L, keep), manifold_view(P, keep), options(opts, keep)
    T = int(np.atleast_1d(opts["thetas"]).size)
    lb = np.cumsum([0] + [len(a) for a in lib_lists[:-1]]).astype(np.int32)
    lc = np.array([len(a) for a in lib_lists], np.int32)
    lr = np.concatenate([np.asarray(a, np.int32) for a in lib_lists]) if lib_lists else np.zeros(1, np.int32)
    shared = len(pred_lists) > 1 and all(np.array_equal(pred_lists[0], a) for a in pred_lists[1:])
    if shared:
        # one prediction list for every task, as a CCM sweep passes it (the rank-table route needs it)
        pb = np.zeros(len(pred_lists), np.int32)
        pc = np.full(len(pred_lists), len(pred_lists[0]), np.int32)
        pr = np.asarray(pred_lists[0], np.int32) if len(pred_lists[0]) else np.zeros(1, np.int32)
    else:
        pb = np.cumsum([0] + [len(a) for a in pred_lists[:-1]]).astype(np.int32)
        pc = np.array([len(a) for a in pred_lists], np.int32)
        pr = np.concatenate([np.asarray(a, np.int32) for a in pred_lists]) if pred_lists else np.zeros(1, np.int32)
    ob = np.cumsum([0] + [T * len(a) for a in pred_lists[:-1]]).astype(np.int64)
    total = int(T * pc.sum())
    pred = np.zeros(max(total, 1))
    rc = np.zeros(max(total, 1), np.int32)
    ku = np.zeros(max(total // T, 1), np.int32)
    k = np.asarray(ks, np.int32)
    status = lib.edm_predict_batch(context(), C.byref(Lv), C.byref(Pv), C.byref(o), len(lib_lists),
                                   lb.ctypes.data_as(_ip), lc.ctypes.data_as(_ip), lr.ctypes.data_as(_ip),
                                   pb.ctypes.data_as(_ip), pc.ctypes.data_as(_ip), pr.ctypes.data_as(_ip),
                                   k.ctypes.data_as(_ip), ob.ctypes.data_as(_i64p), pred.ctypes.data_as(_dp),
                                   rc.ctypes.data_as(_ip), ku.ctypes.data_as(_ip), None)
    if status != 0:
        raise FastEDMError(f"edm_predict_batch failed ({status})")
    out = []
    for t in range(len(lib_lists)):
        n = len(pred_lists[t])
        o0 = int(ob[t])
        out.append((pred[o0:o0 + T * n].reshape(T, n), rc[o0:o0 + T * n].reshape(T, n), ku[o0 // T:o0 // T + n]))
    return out
