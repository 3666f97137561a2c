"""Task-group dumps in the reference's saveInputs format, replayed on the GPU.

The reference's run_command appends one JSON object per call to a dump file when saveInputs names one
(src/r.cpp:275-309, src/r.cpp:30-43): the file holds a JSON array of task groups, each with
  generator  the ManifoldGenerator state (src/manifold.cpp:364-402): _dt, _reldt, _dtWeight, _panel_mode,
             _xmap_mode, _tau, _p, _num_extras, _num_extras_lagged, _x, _xmap, _co_x, _t,
             _observation_number, _extras, _panelIDs;
  opts       the Options fields of src/common.cpp:62-106 (algorithm "Simplex"/"SMap", distance
             "MeanAbsoluteError"/"Euclidean"/"Wasserstein", metrics "Diff"/"CheckSame", ...);
  Es, libraries, k, numReps, crossfold, explore, full, shuffle, saveFinalPredictions,
  saveFinalCoPredictions, saveSMAPCoeffs, copredictMode, usable, rngState.
replay() runs one group through fedm_run_task_group and returns the dictionary edm() returns.
"""
import ctypes as C
import json
import sys

import numpy as np

# src/common.cpp:5-60: a string other than the first spelling decodes to the last alternative
ALGORITHMS = ("Simplex", "SMap")
DISTANCES = ("MeanAbsoluteError", "Euclidean", "Wasserstein")
METRICS = ("Diff", "CheckSame")

GENERATOR_KEYS = ("_dt", "_reldt", "_dtWeight", "_panel_mode", "_xmap_mode", "_tau", "_p", "_num_extras",
                  "_num_extras_lagged", "_x", "_xmap", "_co_x", "_t", "_observation_number", "_extras", "_panelIDs")
OPTION_KEYS = ("copredict", "forceCompute", "savePrediction", "saveSMAPCoeffs", "k", "nthreads", "missingdistance",
               "panelMode", "idw", "thetas", "algorithm", "taskNum", "numTasks", "configNum", "calcRhoMAE",
               "aspectRatio", "distance", "metrics", "cmdLine")
GROUP_KEYS = ("generator", "opts", "Es", "libraries", "k", "numReps", "crossfold", "explore", "full", "shuffle",
              "saveFinalPredictions", "saveFinalCoPredictions", "saveSMAPCoeffs", "copredictMode", "usable",
              "rngState")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class TaskGroup(C.Structure):
    _fields_ = [("n", C.c_int), ("t", _dp), ("x", _dp), ("xmap", _dp), ("co_x", _dp), ("obs", _ip), ("panel", _ip),
                ("extras", _dp), ("num_extras", C.c_int), ("num_extras_lagged", C.c_int), ("dt", C.c_int),
                ("reldt", C.c_int), ("dt_weight", C.c_double), ("tau", C.c_int), ("p", C.c_int),
                ("algorithm", C.c_int), ("distance", C.c_int), ("thetas", _dp), ("num_thetas", C.c_int),
                ("missing_distance", C.c_double), ("panel_mode", C.c_int), ("idw", C.c_double),
                ("extra_metrics", _ip), ("aspect_ratio", C.c_double), ("save_smap_coeffs", C.c_int),
                ("E", _ip), ("num_E", C.c_int), ("libraries", _ip), ("num_libraries", C.c_int), ("k", C.c_int),
                ("num_reps", C.c_int), ("crossfold", C.c_int), ("explore", C.c_int), ("full", C.c_int),
                ("shuffle", C.c_int), ("save_final_predictions", C.c_int), ("save_final_copredictions", C.c_int),
                ("copredict_mode", C.c_int), ("usable", _ip), ("seed", C.c_int), ("precision", C.c_int)]


def decode(value, names):
    """The enum decoding of src/common.cpp: the first spelling, else the following ones, else the last."""
    for i, name in enumerate(names[:-1]):
        if value == name:
            return i
    return len(names) - 1


def read_task_groups(path):
    """The task groups of a dump file (a JSON array; a single object is read as one group)."""
    with open(path) as f:
        data = json.load(f)
    groups = data if isinstance(data, list) else [data]
    for g in groups:
        missing = [k for k in GROUP_KEYS if k not in g]
        missing += ["generator." + k for k in GENERATOR_KEYS if k not in g.get("generator", {})]
        missing += ["opts." + k for k in OPTION_KEYS if k not in g.get("opts", {})]
        if missing:
            raise ValueError(f"not a task group dump: missing {missing}")
    return groups


def group_struct(g, precision=0, seed=0, keep=None):
    """fedm_task_group of one dumped group; `keep` receives the arrays the structure points into."""
    keep = [] if keep is None else keep
    gen, opts = g["generator"], g["opts"]

    def darr(v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        keep.append(a)
        return a

    def iarr(v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
        keep.append(a)
        return a

    s = TaskGroup()
    t = darr(gen["_t"])
    n = t.size
    s.n = n
    s.t = t.ctypes.data_as(_dp)
    s.x = darr(gen["_x"]).ctypes.data_as(_dp)
    if gen["_xmap_mode"] and len(gen["_xmap"]) == n:
        s.xmap = darr(gen["_xmap"]).ctypes.data_as(_dp)
    if len(gen["_co_x"]) == n:
        s.co_x = darr(gen["_co_x"]).ctypes.data_as(_dp)
    s.obs = iarr(gen["_observation_number"]).ctypes.data_as(_ip)
    if gen["_panel_mode"] and len(gen["_panelIDs"]) == n:
        s.panel = iarr(gen["_panelIDs"]).ctypes.data_as(_ip)
    extras = gen["_extras"]
    s.num_extras = len(extras)
    if extras:
        s.extras = darr(np.asarray(extras, dtype=np.float64).reshape(len(extras), n)).ctypes.data_as(_dp)
    s.num_extras_lagged = int(gen["_num_extras_lagged"])
    s.dt, s.reldt = int(bool(gen["_dt"])), int(bool(gen["_reldt"]))
    s.dt_weight = float(gen["_dtWeight"])
    s.tau, s.p = int(gen["_tau"]), int(gen["_p"])
    s.algorithm = decode(opts["algorithm"], ALGORITHMS)
    s.distance = decode(opts["distance"], DISTANCES)
    th = darr(opts["thetas"])
    s.thetas = th.ctypes.data_as(_dp)
    s.num_thetas = th.size
    s.missing_distance = float(opts["missingdistance"])
    s.panel_mode = int(bool(opts["panelMode"]))
    s.idw = float(opts["idw"])
    if opts["metrics"]:
        s.extra_metrics = iarr([decode(m, METRICS) for m in opts["metrics"]]).ctypes.data_as(_ip)
    s.aspect_ratio = float(opts["aspectRatio"])
    s.save_smap_coeffs = int(bool(g["saveSMAPCoeffs"]))
    Es = iarr(g["Es"])
    s.E, s.num_E = Es.ctypes.data_as(_ip), Es.size
    libs = iarr(g["libraries"]) if len(g["libraries"]) else iarr([0])
    s.libraries, s.num_libraries = libs.ctypes.data_as(_ip), len(g["libraries"])
    s.k = int(g["k"])
    s.num_reps, s.crossfold = int(g["numReps"]), int(g["crossfold"])
    s.explore, s.full, s.shuffle = int(bool(g["explore"])), int(bool(g["full"])), int(bool(g["shuffle"]))
    s.save_final_predictions = int(bool(g["saveFinalPredictions"]))
    s.save_final_copredictions = int(bool(g["saveFinalCoPredictions"]))
    s.copredict_mode = int(bool(g["copredictMode"]))
    s.usable = iarr(g["usable"]).ctypes.data_as(_ip)
    s.seed = int(seed)
    s.precision = int(precision)
    return s


def replay(group, precision=0, seed=0):
    """Runs one dumped task group on the GPU; returns edm()'s result dictionary (rc, stats, summary, kMin,
    kMax, and the saved predictions, copredictions and coefficients the group asked for)."""
    pkg = sys.modules[__package__]
    L = pkg.load_library()
    keep = []
    s = group_struct(group, precision, seed, keep)
    f = L.fedm_run_task_group
    f.argtypes = [C.c_void_p, C.POINTER(TaskGroup), C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    status = f(pkg.context(), C.byref(s), C.byref(h))
    if status < 0:
        raise pkg.FastEDMError(f"fedm_run_task_group failed ({status})")
    gen = group["generator"]
    try:
        return pkg._collect(L, h, bool(group["copredictMode"]), float(group["opts"]["missingdistance"]) != 0.0,
                            bool(gen["_dt"]) or bool(gen["_reldt"]))
    finally:
        L.fedm_result_free(h)
