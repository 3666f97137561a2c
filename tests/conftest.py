import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def load_product():
    """Imports the product package (its directory name is not a Python identifier)."""
    name = "fastedm_r_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "fastedm-r_amd", "__init__.py"),
                                                  submodule_search_locations=[os.path.join(ROOT, "fastedm-r_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def golden_vectors():
    with open(os.path.join(GOLDEN, "test_ci_vectors.json")) as f:
        return json.load(f)


def logistic_map(obs):
    """R/logistic_map.R:12-31 (coupled logistic maps used by the reference's tests)."""
    r_x, r_y, b_xy, b_yx = 3.625, 3.77, 0.05, 0.4
    x = np.zeros(obs)
    y = np.zeros(obs)
    x[0], y[0] = 0.2, 0.4
    for i in range(1, obs):
        x[i] = x[i - 1] * (r_x * (1 - x[i - 1]) - b_xy * y[i - 1])
        y[i] = y[i - 1] * (r_y * (1 - y[i - 1]) - b_yx * x[i - 1])
    return x, y


def tslag(t, x, lag=1, dt=1):
    """test-ci.R:1-10"""
    t = np.asarray(t, float)
    x = np.asarray(x, float)
    out = np.full(len(t), np.nan)
    for i in range(len(t)):
        lt = t[i] - lag * dt
        if not np.isnan(lt):
            w = np.nonzero(t == lt)[0]
            if w.size:
                out[i] = x[w[0]]
    return out


def tsdiff(t, x, lag=1, dt=1):
    """test-ci.R:12-21"""
    t = np.asarray(t, float)
    x = np.asarray(x, float)
    out = np.full(len(t), np.nan)
    for i in range(len(t)):
        lt = t[i] - lag * dt
        if not np.isnan(x[i]) and not np.isnan(lt):
            w = np.nonzero(t == lt)[0]
            if w.size:
                out[i] = x[i] - x[w[0]]
    return out


def r_sd(v):
    v = np.asarray(v, float)
    v = v[~np.isnan(v)]
    return float(np.std(v, ddof=1))


def expect_approx_equal(a, b):
    """test-ci.R:23-33: same length, all finite, max abs error < 1e-4."""
    a = np.atleast_1d(np.asarray(a, float))
    b = np.atleast_1d(np.asarray(b, float))
    assert a.shape == b.shape, (a, b)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b)), (a, b)
    assert np.max(np.abs(a - b)) < 1e-4, (a, b)


def check_edm_result(res, rho, co_rho=None):
    assert res["rc"] == 0
    expect_approx_equal(res["stats"]["rho"], rho)
    if co_rho is not None:
        expect_approx_equal(res["copredStats"]["rho"], co_rho)


def check_noisy_edm_result(res, lo, hi, co_lo=None, co_hi=None):
    assert res["rc"] == 0
    from oracle.oracle import summarise
    m = summarise(res["stats"])["rho"]
    assert np.all((lo <= m) & (m <= hi)), (m, lo, hi)
    if co_lo is not None:
        cm = summarise(res["copredStats"])["rho"]
        assert np.all((co_lo <= cm) & (cm <= co_hi)), (cm, co_lo, co_hi)
