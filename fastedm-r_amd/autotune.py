"""Autotunes the plans of given workloads on the current GPU and prints the tuner's reports.

    python fastedm-r_amd/autotune.py [ccm] [panel] [lorenz] [--trials N]

ccm     the CCM sweep on coupled logistic maps (T = 20000, E = 3, 40 libraries 10..19000 x 100 replicates)
panel   S-map on 512 logistic-map panels (E = 8 with dt, k = 18, thetas 0, 1, 3), mixed form
lorenz  nothing here: the Lorenz-96 workload runs through a batch, not a plan

The choices go to the tuning cache (FEDM_TUNE_CACHE, else $HOME/.cache/fastedm_hip/tune.txt), where
later plans of the same workload class find them.
"""
import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _package():
    name = "fastedm_r_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def coupled_maps(T):
    x = np.empty(T)
    y = np.empty(T)
    x[0], y[0] = 0.4, 0.2
    for i in range(T - 1):
        x[i + 1] = x[i] * (3.8 - 3.8 * x[i] - 0.02 * y[i])
        y[i + 1] = y[i] * (3.5 - 3.5 * y[i] - 0.1 * x[i])
    return x, y


def ccm_plan(fe, precision):
    x, y = coupled_maps(20000)
    t = np.arange(1, 20001, dtype=float)
    libs = np.round(np.exp(np.linspace(np.log(10), np.log(19000), 40))).astype(int)
    return fe.Plan(t, y, x, E=3, library=libs, numReps=100, precision=precision)


def panel_plan(fe, precision):
    sys.path.insert(0, ROOT)
    from tests import configs
    t, x, panel = configs.cfg4_panels()
    return fe.Plan(t, x, panel=panel, E=8, dt=True, theta=[0, 1, 3], algorithm="smap", precision=precision)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workloads", nargs="*", default=["ccm", "panel"])
    ap.add_argument("--trials", type=int, default=5)
    args = ap.parse_args()
    fe = _package()
    makers = {"ccm": ccm_plan, "panel": panel_plan}
    for w in args.workloads:
        if w not in makers:
            continue
        t0 = time.perf_counter()
        plan = makers[w](fe, fe.MIXED)
        made = time.perf_counter() - t0
        t0 = time.perf_counter()
        n, report = plan.tune(trials=args.trials)
        print(f"== {w} (mixed): plan created in {made:.1f} s, tuned in {time.perf_counter() - t0:.1f} s, "
              f"{n} group(s)\n{report}", flush=True)
        plan.close()


if __name__ == "__main__":
    main()
