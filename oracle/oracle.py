"""CPU oracle bindings — TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``oracle/liboracle.so`` (a C++ restatement of the reference CPU path) and an
``edm()`` that mirrors the reference R function ``edm()`` (R/edm.R:232-393) argument for argument.
Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MISSING_D = 1.0e100
MISSING_I = -(2**31)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class ManifoldView(C.Structure):
    _fields_ = [("data", _dp), ("targets", _dp), ("point_times", _dp), ("target_times", _dp), ("panel", _ip),
                ("num_points", C.c_int), ("E_x", C.c_int), ("E_dt", C.c_int), ("E_extras", C.c_int),
                ("E_lagged_extras", C.c_int), ("E_actual", C.c_int)]


class Options(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("distance", C.c_int), ("k", C.c_int), ("num_thetas", C.c_int),
                ("thetas", _dp), ("metrics", _ip), ("missing_distance", C.c_double), ("panel_mode", C.c_int),
                ("idw", C.c_double), ("num_panels", C.c_int), ("id_weights", _fp), ("has_missing", C.c_int),
                ("has_categorical", C.c_int), ("use_only_past", C.c_int), ("force_compute", C.c_int),
                ("save_smap_coeffs", C.c_int), ("low_memory", C.c_int), ("aspect_ratio", C.c_double),
                ("precision", C.c_int), ("nthreads", C.c_int)]


class Command(C.Structure):
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
                ("num_extras_lagged", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_library(path):
    """Binds a different build of the same sources (the CPU baseline's -O3 -march=native one)."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_run_command.restype = C.c_void_p
        L.orc_run_command.argtypes = [C.POINTER(Command)]
        for name in ["orc_result_rc", "orc_result_num_stats", "orc_result_matrix_shape", "orc_result_num_manifolds"]:
            getattr(L, name).restype = C.c_int
        L.orc_result_scalar.restype = C.c_double
        L.orc_correlation.restype = C.c_double
        L.orc_mae.restype = C.c_double
        L.orc_emd.restype = C.c_double
        L.orc_mt64_stream.argtypes = [C.c_uint64, C.c_int, _dp]
        L.orc_manifold.argtypes = [C.POINTER(Command), C.c_int, _ip, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.orc_split_masks.argtypes = [C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _ip, _ip, _ip, C.c_int]
        L.orc_generator_state.argtypes = [C.POINTER(Command), C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _ip]
        L.orc_generator_state.restype = C.c_int
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def as_double(v):
    """R numeric vector -> float64 array; NA (None/NaN) stays NaN and becomes MISSING in C++."""
    return np.ascontiguousarray([np.nan if a is None else a for a in v] if isinstance(v, list) else v,
                                dtype=np.float64)


class CommandBuilder:
    """Holds the arrays a Command points to, so they outlive the C call."""

    def __init__(self):
        self.keep = []

    def d(self, v):
        if v is None:
            return None
        a = as_double(v)
        self.keep.append(a)
        return dptr(a)

    def i(self, v):
        if v is None:
            return None
        a = np.ascontiguousarray(v, dtype=np.int32)
        self.keep.append(a)
        return iptr(a)


def _is_nothing(v):
    return v is None or (isinstance(v, float) and np.isnan(v))


def make_command(t, x, y=None, panel=None, E=2, tau=1, theta=1, library=None, k=0, algorithm="simplex", p=None,
                 crossfold=0, full=False, shuffle=False, seed=0, copredict=None, savePredictions=False,
                 saveCoPredictions=False, saveManifolds=False, saveSMAPCoeffs=False, extras=None,
                 allowMissing=False, missingDistance=0.0, dt=False, reldt=False, dtWeight=0.0, numReps=1,
                 panelWeight=0, panelWeights=None, numThreads=1, lowMemory=False, predictWithPast=False,
                 distance="euclidean", aspectRatio=1.0, extrasLagged=0):
    """Argument normalisation of R/edm.R:241-338 and the fixed choices of src/r.cpp:123-159.  extrasLagged
    (no R argument: src/r.cpp:246 fixes it at 0) is the generator's numExtrasLagged, which task-group dumps
    record: the first extrasLagged extras then enter the points with E lags each."""
    n = len(t)
    if len(x) != n:
        raise ValueError("The time and x variables should be the same length")
    y = None if y is None or len(y) == 0 else y
    panel = None if panel is None or len(panel) == 0 else panel
    copredict = None if copredict is None or len(copredict) == 0 else copredict
    if numReps > 1:
        shuffle = True
    explore = y is None
    p = (1 if explore else 0) if p is None else int(p)
    k = 0 if _is_nothing(k) else k
    Es = np.atleast_1d(np.asarray(E, dtype=np.float64))
    if Es.size == 1:
        Es = np.array([2.0]) if np.isnan(Es[0]) else Es
    else:
        Es = Es[~np.isnan(Es)]
    k = -1 if k == np.inf else int(k)
    panelWeight = -1.0 if panelWeight == np.inf else float(panelWeight)
    thetas = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    algo = {"simplex": 0, "smap": 1}.get(algorithm, 9)
    dist = {"mae": 0, "euclidean": 1, "wasserstein": 2}[distance]

    b = CommandBuilder()
    c = Command()
    c.n = n
    c.t = b.d(t)
    c.x = b.d(x)
    c.y = b.d(y)
    c.co_x = b.d(copredict)
    if panel is not None:
        pa = np.asarray(panel, dtype=np.float64)
        c.panel = b.i(np.where(np.isnan(pa), MISSING_I, pa).astype(np.int64).astype(np.int32))
    ex = [] if extras is None else list(extras)
    for e in ex:
        if len(e) != n:
            raise ValueError("An extra variable is not the right size")
    c.num_extras = len(ex)
    c.extras = b.d(np.concatenate([as_double(e) for e in ex]) if ex else None)
    c.num_E = int(Es.size)
    c.E = b.i(Es.astype(np.int32))
    c.tau = int(tau)
    c.num_thetas = int(thetas.size)
    c.thetas = b.d(thetas)
    libs = None if library is None else np.atleast_1d(np.asarray(library)).astype(np.int32)
    c.num_libraries = 0 if libs is None else int(libs.size)
    c.libraries = b.i(libs)
    c.k = k
    c.algorithm = algo
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
        pw = np.ascontiguousarray(panelWeights, dtype=np.float64)
        c.panel_weights = b.d(pw.ravel())
        c.panel_weights_dim = pw.shape[0]
    c.num_threads = int(numThreads)
    c.low_memory = int(bool(lowMemory))
    c.predict_with_past = int(bool(predictWithPast))
    c.distance = dist
    c.aspect_ratio = float(aspectRatio)
    c.precision = 0
    c.num_extras_lagged = int(extrasLagged)
    return c, b, explore


def _na(a):
    a = np.asarray(a, dtype=np.float64)
    return np.where(a == MISSING_D, np.nan, a)


def collect_result(L, prefix, r, copredict_mode, extras_flags):
    """Turns a run_command result handle into R's result list (src/r.cpp:421-457)."""
    f = lambda name: getattr(L, prefix + name)
    r = C.c_void_p(r)
    res = {"rc": f("result_rc")(r)}

    def stats(co):
        m = f("result_num_stats")(r, co)
        E = np.zeros(m, np.int32)
        lib_ = np.zeros(m, np.int32)
        th, rho, mae = np.zeros(m), np.zeros(m), np.zeros(m)
        f("result_stats")(r, co, iptr(E), iptr(lib_), dptr(th), dptr(rho), dptr(mae))
        return {"E": E, "library": lib_, "theta": th, "rho": _na(rho), "mae": _na(mae)}

    res["stats"] = stats(0)
    if copredict_mode:
        res["copredStats"] = stats(1)
    kmin, kmax = C.c_int(), C.c_int()
    f("result_k")(r, C.byref(kmin), C.byref(kmax))
    res["kMin"] = None if kmin.value == MISSING_I else kmin.value
    res["kMax"] = None if kmax.value == MISSING_I else kmax.value
    for which, name in [(0, "predictions"), (1, "copredictions"), (2, "coeffs")]:
        rows, cols = C.c_int(), C.c_int()
        if f("result_matrix_shape")(r, which, C.byref(rows), C.byref(cols)):
            out = np.zeros(rows.value * cols.value)
            f("result_matrix")(r, which, dptr(out))
            res[name] = _na(out.reshape(cols.value, rows.value).T)
    nm = f("result_num_manifolds")(r)
    if nm:
        res["Ms"], res["Mps"] = [], []
        for i in range(nm):
            for ps, key in [(0, "Ms"), (1, "Mps")]:
                rows, cols = C.c_int(), C.c_int()
                f("result_manifold_shape")(r, i, ps, C.byref(rows), C.byref(cols))
                out = np.zeros(rows.value * cols.value)
                f("result_manifold")(r, i, ps, dptr(out))
                res[key].append(_na(out.reshape(rows.value, cols.value)))
    if extras_flags.get("allowMissing"):
        res["missingdistance"] = f("result_scalar")(r, 0)
    if extras_flags.get("dt"):
        res["dtWeight"] = f("result_scalar")(r, 1)
    res["summary"] = summarise(res["stats"])
    f("result_free")(r)
    return res


def summarise(st):
    """stats::aggregate(cbind(rho, mae) ~ E + library + theta, na.omit(stats), mean) (R/edm.R:347-355)."""
    ok = ~(np.isnan(st["rho"]) | np.isnan(st["mae"]))
    keys = {}
    for i in np.nonzero(ok)[0]:
        key = (int(st["E"][i]), int(st["library"][i]), float(st["theta"][i]))
        keys.setdefault(key, []).append(i)
    out = {"E": [], "library": [], "theta": [], "rho": [], "mae": []}
    for key in sorted(keys, key=lambda k: (k[2], k[1], k[0])):
        idx = keys[key]
        out["E"].append(key[0])
        out["library"].append(key[1])
        out["theta"].append(key[2])
        out["rho"].append(float(np.mean(st["rho"][idx])))
        out["mae"].append(float(np.mean(st["mae"][idx])))
    return {k: np.asarray(v) for k, v in out.items()}


def edm(t, x, y=None, **kw):
    """The reference edm() (R/edm.R:232) evaluated on the CPU oracle."""
    c, keep, explore = make_command(t, x, y, **kw)
    L = lib()
    r = L.orc_run_command(C.byref(c))
    return collect_result(L, "orc_", r, kw.get("copredict") is not None and len(kw.get("copredict")) > 0,
                          {"allowMissing": kw.get("allowMissing", False),
                           "dt": kw.get("dt", False) or kw.get("reldt", False)})


def generator_state(t, x, y=None, **kw):
    """The manifold generator's state for an edm() call, as the reference's saveInputs dump keeps it
    (src/manifold.cpp:364-382, src/r.cpp:160-262): normalised series, extras, observation numbers, flags,
    dt weight and the usable rows for the last E."""
    c, keep, explore = make_command(t, x, y, **kw)
    L = lib()
    n = c.n
    maxE = int(c.E[c.num_E - 1])  # src/r.cpp:256: the last E
    ne = c.num_extras
    tt, xx, xm, cx = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    ex = np.zeros(max(ne, 1) * n)
    obs = np.zeros(n, np.int32)
    flags = np.zeros(4, np.int32)
    w = C.c_double()
    usable = np.zeros(n, np.int32)
    L.orc_generator_state(C.byref(c), maxE, dptr(tt), dptr(xx), dptr(xm), dptr(cx), dptr(ex), iptr(obs),
                          iptr(flags), C.byref(w), iptr(usable))
    return {"t": tt, "x": xx, "xmap": xm if flags[2] else None, "co_x": cx if c.co_x else None,
            "extras": ex[:ne * n].reshape(ne, n), "obs": obs, "dt": bool(flags[0]), "reldt": bool(flags[1]),
            "panel_mode": bool(flags[3]), "dt_weight": w.value, "usable": usable, "explore": explore,
            "num_usable": int(usable.sum()), "command": c, "keep": keep}


def predict(M, Mp, opts, nn_cap=0):
    """Task-level oracle: dicts of numpy arrays in, (pred, rc, coeffs, k_used, nn_idx) out."""
    L = lib()
    keep = []
    mv, mpv = manifold_view(M, keep), manifold_view(Mp, keep)
    o = options_struct(opts, keep)
    Np, T, Cc = Mp["data"].shape[0], len(opts["thetas"]), M["data"].shape[1] + 1
    pred = np.zeros(T * Np)
    rc = np.zeros(T * Np, np.int32)
    coeffs = np.zeros(Np * Cc)
    ku = np.zeros(Np, np.int32)
    nn = np.zeros(max(1, Np * nn_cap), np.int32)
    L.orc_predict(C.byref(mv), C.byref(mpv), C.byref(o), dptr(pred), iptr(rc), dptr(coeffs), iptr(ku),
                  iptr(nn) if nn_cap else None, nn_cap)
    return (pred.reshape(T, Np), rc.reshape(T, Np), coeffs.reshape(Np, Cc), ku,
            nn[:Np * nn_cap].reshape(Np, nn_cap) if nn_cap else None)


def distances(M, Mp, opts, i):
    """Potential neighbours of target i before the k selection: (library indices, distances)."""
    L = lib()
    keep = []
    mv, mpv = manifold_view(M, keep), manifold_view(Mp, keep)
    o = options_struct(opts, keep)
    n = M["data"].shape[0]
    inds = np.zeros(max(n, 1), np.int32)
    d = np.zeros(max(n, 1))
    c = L.orc_distances(C.byref(mv), C.byref(mpv), C.byref(o), i, iptr(inds), dptr(d))
    return inds[:c], d[:c]


def manifold_view(m, keep):
    v = ManifoldView()
    data = np.ascontiguousarray(m["data"], np.float64)
    n, Ea = data.shape
    arrs = {k: np.ascontiguousarray(m.get(k, np.zeros(n)), np.float64) for k in ("targets", "point_times",
                                                                                   "target_times")}
    keep.extend([data] + list(arrs.values()))
    v.data = dptr(data)
    v.targets, v.point_times, v.target_times = (dptr(arrs[k]) for k in ("targets", "point_times", "target_times"))
    if m.get("panel") is not None:
        pa = np.ascontiguousarray(m["panel"], np.int32)
        keep.append(pa)
        v.panel = iptr(pa)
    v.num_points = n
    v.E_x = m.get("E_x", Ea)
    v.E_dt = m.get("E_dt", 0)
    v.E_extras = m.get("E_extras", 0)
    v.E_lagged_extras = m.get("E_lagged_extras", 0)
    v.E_actual = Ea
    return v


def options_struct(d, keep):
    o = Options()
    th = np.ascontiguousarray(d["thetas"], np.float64)
    keep.append(th)
    o.algorithm = d.get("algorithm", 0)
    o.distance = d.get("distance", 1)
    o.k = d["k"]
    o.num_thetas = th.size
    o.thetas = dptr(th)
    if d.get("metrics") is not None:
        m = np.ascontiguousarray(d["metrics"], np.int32)
        keep.append(m)
        o.metrics = iptr(m)
    o.missing_distance = d.get("missing_distance", 0.0)
    o.panel_mode = d.get("panel_mode", 0)
    o.idw = d.get("idw", 0.0)
    if d.get("id_weights") is not None:
        w = np.ascontiguousarray(d["id_weights"], np.float32)
        keep.append(w)
        o.id_weights = fptr(w)
        o.num_panels = w.shape[0]
    o.has_missing = d.get("has_missing", 0)
    o.has_categorical = d.get("has_categorical", 0)
    o.use_only_past = d.get("use_only_past", 0)
    o.force_compute = d.get("force_compute", 1)
    o.save_smap_coeffs = d.get("save_smap_coeffs", 0)
    o.low_memory = d.get("low_memory", 0)
    o.aspect_ratio = d.get("aspect_ratio", 1.0)
    o.precision = d.get("precision", 0)
    o.nthreads = d.get("nthreads", 1)
    return o


def mt64_stream(seed, n):
    out = np.zeros(n)
    lib().orc_mt64_stream(seed, n, dptr(out))
    return out


def jacobi_solve(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    n = b.size
    x, sv = np.zeros(n), np.zeros(n)
    r = C.c_int()
    lib().orc_jacobi_solve(n, dptr(A), dptr(b), dptr(x), dptr(sv), C.byref(r))
    return x, sv, r.value


def emd(Cm):
    Cm = np.ascontiguousarray(Cm, np.float64)
    return lib().orc_emd(dptr(Cm), Cm.shape[0], Cm.shape[1])
