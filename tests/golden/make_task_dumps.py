"""Writes task_dumps.json: edm() calls recorded as the reference's saveInputs dump records them.

The reference appends one task group per run_command call (src/r.cpp:275-309): the manifold generator
after its normalisation (src/manifold.cpp:364-382), the Options that the R front end filled in
(src/r.cpp:123-159, 252-254, 279-281; src/common.cpp:62-83) and the arguments of launch_tasks.  The
generator state and the usable rows come from the oracle's restatement of the same generator; the
missing distance that the front end fills in comes from the oracle's run of the call.  CALLS names the
calls; the file holds one group per call in the order of CALLS, and tests replay each group on the GPU
and compare it with the oracle's run of the call itself.

    python tests/golden/make_task_dumps.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests import scenarios  # noqa: E402

OUT = os.path.join(HERE, "task_dumps.json")


def _shifted(v, s):
    out = np.full(len(v), np.nan)
    out[s:] = v[:-s]
    return out


def calls():
    """name -> keyword arguments of edm() (R/edm.R:232)."""
    t, x, y, _, _ = scenarios.bigger_data(missing=False)
    tm, xm, ym, u, u1 = scenarios.bigger_data(missing=True)
    ts, xs, ys = scenarios.simple_data()
    tp, xp, yp, panel = scenarios.panel_data(drop7=True)
    return {
        "explore_simplex": dict(t=t, x=x, E=[2, 3, 4, 5], savePredictions=True),
        "xmap_libraries_reps": dict(t=t, x=y, y=x, E=3, library=[10, 20, 50, 90], numReps=4, seed=7),
        "smap_coeffs": dict(t=t, x=x, y=y, theta=0.2, algorithm="smap", saveSMAPCoeffs=True,
                            savePredictions=True),
        "smap_thetas": dict(t=t, x=x, E=4, k=np.inf, theta=[0.0, 0.5, 1.0, 2.0, 4.0], algorithm="smap"),
        "copredict": dict(t=t, x=x, y=y, E=3, copredict=_shifted(y, 2), savePredictions=True,
                          saveCoPredictions=True),
        "missing_dt": dict(t=tm, x=xm, y=ym, E=2, allowMissing=True, dt=True, savePredictions=True),
        "extras_missing": dict(t=tm, x=xm, y=ym, E=[3, 4], extras=[u1], allowMissing=True),
        "panel_idw": dict(t=tp, x=xp, y=yp, panel=panel, E=2, panelWeight=0.5),
        "panel_inf": dict(t=tp, x=xp, panel=panel, E=[2, 3], panelWeight=np.inf),
        "crossfold": dict(t=ts, x=xs, E=2, crossfold=5, seed=3),
        "full_k": dict(t=t, x=x, y=y, E=2, k=5, full=True),
        "reldt_explore": dict(t=tm, x=xm, E=3, reldt=True),
        # a lagged extra (numExtrasLagged = 1: the R front end never sets it, src/r.cpp:246, but the
        # generator and its dumps carry it) beside an unlagged one, with missing values
        "extras_lagged": dict(t=tm, x=xm, y=ym, E=[2, 3], extras=[u, u1], extrasLagged=1, allowMissing=True,
                              savePredictions=True),
    }


def _list(a):
    return [float(v) for v in np.asarray(a, dtype=np.float64)]


def task_group(kw):
    """One task group in the dump's layout (src/r.cpp:284-300)."""
    st = oracle.generator_state(**kw)
    c = st["command"]
    n = c.n
    es = [int(c.E[i]) for i in range(c.num_E)]
    libraries = [int(c.libraries[i]) for i in range(c.num_libraries)]
    explore = bool(st["explore"])
    if not explore and not libraries:
        libraries = [st["num_usable"]]
    missing_distance = float(c.missing_distance)
    if c.allow_missing and missing_distance == 0:
        missing_distance = float(oracle.edm(**kw)["missingdistance"])
    thetas = [float(c.thetas[i]) for i in range(c.num_thetas)]
    panel = [] if not st["panel_mode"] else [int(c.panel[i]) for i in range(n)]
    num_tasks = c.num_reps * max(c.crossfold, 1) * len(es) * max(len(libraries), 1)
    generator = {
        "_dt": st["dt"], "_reldt": st["reldt"], "_dtWeight": st["dt_weight"], "_panel_mode": st["panel_mode"],
        "_xmap_mode": st["xmap"] is not None, "_tau": int(c.tau), "_p": int(c.p),
        "_num_extras": int(c.num_extras), "_num_extras_lagged": int(c.num_extras_lagged),
        "_x": _list(st["x"]), "_xmap": [] if st["xmap"] is None else _list(st["xmap"]),
        "_co_x": [] if st["co_x"] is None else _list(st["co_x"]), "_t": _list(st["t"]),
        "_observation_number": [int(v) for v in st["obs"]], "_extras": [_list(e) for e in st["extras"]],
        "_panelIDs": panel,
    }
    opts = {
        "copredict": False, "forceCompute": True, "savePrediction": False,
        "saveSMAPCoeffs": bool(c.save_smap_coeffs), "k": int(c.k), "nthreads": int(c.num_threads),
        "missingdistance": missing_distance, "panelMode": st["panel_mode"], "idw": float(c.panel_weight),
        "thetas": thetas, "algorithm": "Simplex" if c.algorithm == 0 else "SMap", "taskNum": 0,
        "numTasks": num_tasks, "configNum": 0, "calcRhoMAE": True, "aspectRatio": 1.0, "distance": "Euclidean",
        "metrics": ["Diff"] * int(c.num_extras), "cmdLine": "",
    }
    return {
        "generator": generator, "opts": opts, "Es": es, "libraries": libraries, "k": int(c.k),
        "numReps": int(c.num_reps), "crossfold": int(c.crossfold), "explore": explore, "full": bool(c.full),
        "shuffle": bool(c.shuffle), "saveFinalPredictions": bool(c.save_final_predictions),
        "saveFinalCoPredictions": bool(c.save_final_copredictions), "saveSMAPCoeffs": bool(c.save_smap_coeffs),
        "copredictMode": st["co_x"] is not None, "usable": [int(v) for v in st["usable"]], "rngState": "",
    }


def dumps():
    return [task_group(kw) for kw in calls().values()]


def main():
    with open(OUT, "w") as f:
        json.dump(dumps(), f, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
