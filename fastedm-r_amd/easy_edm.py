"""easy_edm: the reference's automated causal-analysis workflow (R/easy_edm.R:40-433) over edm().

The workflow makes a fixed sequence of edm() calls (R/easy_edm.R:163-433):
  1. the embedding dimension E maximising simplex rho over E = 3..10;
  2. a nonlinearity test: S-map over theta = 0, 0.05, ..., 5 with k = all, then 20 replicates at theta = 0
     and at the best theta, compared by a one-sided Kolmogorov-Smirnov test (p < 0.05: nonlinear);
  3. the lag of the effect (-5..6) whose cross-map rho is largest, the effect lagged (nonlinear) or
     differenced by it (linear);
  4. the largest library (the full manifold of the effect);
  5. convergent cross mapping: 1000 shuffled replicates at a small library against the rho at the largest
     library (the "quantile" method, the reference's default), or an exponential fit of rho against the
     library size (the "parametric" method).
It returns True when CCM found evidence of causation.

``engine`` is the edm() implementation the steps call: this package's GPU edm() unless another function
with the same arguments and result dictionary is given.
"""
import math
import sys

import numpy as np


def _engine(engine):
    if engine is not None:
        return engine
    return sys.modules[__package__].edm


def _alert(verbosity, msg):
    if verbosity > 0:
        print(msg)


def scale(v):
    """R's scale() of one column: (v - mean) / sd with the n - 1 sd, missing values ignored."""
    v = np.asarray(v, dtype=float)
    ok = ~np.isnan(v)
    mu = v[ok].mean()
    sd = v[ok].std(ddof=1)
    return (v - mu) / sd


def tslag(t, x, lag=1, dt=1):
    """x at time t - lag * dt for every t, NaN where that time is absent (R/easy_edm.R:218-227)."""
    t = np.asarray(t, dtype=float)
    x = np.asarray(x, dtype=float)
    first = {}
    for i, ti in enumerate(t):
        if not np.isnan(ti) and ti not in first:
            first[ti] = i
    out = np.full(t.size, np.nan)
    for i, ti in enumerate(t):
        lt = ti - lag * dt
        if not np.isnan(lt) and lt in first:
            out[i] = x[first[lt]]
    return out


def r_quantile(x, probs):
    """R's quantile(type = 7)."""
    x = np.sort(np.asarray(x, dtype=float))
    n = x.size
    out = []
    for p in probs:
        h = (n - 1) * p
        lo = math.floor(h)
        hi = min(lo + 1, n - 1)
        out.append(x[lo] + (h - lo) * (x[hi] - x[lo]))
    return out


def ks_less_pvalue(x, y):
    """p-value of R's ks.test(x, y, alternative = "less"): the CDF of x lies below that of y somewhere.
    The statistic is max(F_y - F_x); scipy's two-sample test with alternative "less" is the same test."""
    from scipy import stats as sps
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    x = x[~np.isnan(x)]
    y = y[~np.isnan(y)]
    r = sps.ks_2samp(x, y, alternative="less")
    return float(r.statistic), float(r.pvalue)


def preprocess_inputs(data, cause, effect, time, verbosity, normalize):
    """Series from a data frame (column names) or from vectors (R/easy_edm.R:103-155)."""
    if data is not None:
        _alert(verbosity, "Pulling the time series from the supplied dataframe.")
        cols = list(data.keys()) if isinstance(data, dict) else list(data.columns)
        for name in (cause, effect) + ((time,) if time is not None else ()):
            if name not in cols:
                raise ValueError(f"{name} is not a column in the supplied dataframe.")
        x = np.asarray(data[cause], dtype=float)
        y = np.asarray(data[effect], dtype=float)
        t = np.arange(1, x.size + 1, dtype=float) if time is None else np.asarray(data[time], dtype=float)
    else:
        _alert(verbosity, "Using supplied time series vectors.")
        x = np.asarray(cause, dtype=float)
        y = np.asarray(effect, dtype=float)
        t = np.arange(1, x.size + 1, dtype=float) if time is None else np.asarray(time, dtype=float)
    if not (t.size == x.size == y.size):
        raise ValueError("Time series are not the same length.")
    _alert(verbosity, f"Number of observations is {t.size}")
    if normalize:
        _alert(verbosity, "Normalizing the supplied time series")
        x = scale(x)
        y = scale(y)
    return t, x, y


def find_embedding_dimension(edm, t, x, verbosity=0):
    """E in 3..10 with the largest simplex rho (R/easy_edm.R:159-180)."""
    res = edm(t, x, E=list(range(3, 11)))
    if res["rc"] > 0:
        raise RuntimeError("Search for optimal embedding dimension failed.")
    s = res["summary"]
    if len(s["rho"]) == 0:
        raise RuntimeError("Search for optimal embedding dimension failed (2).")
    E_best = int(s["E"][int(np.argmax(s["rho"]))])
    _alert(verbosity, f"Found optimal embedding dimension E to be {E_best}.")
    return E_best


def test_nonlinearity(edm, t, x, E_best, max_theta=5, num_thetas=100, theta_reps=20, verbosity=0):
    """Best S-map theta and the Kolmogorov-Smirnov nonlinearity test (R/easy_edm.R:184-215)."""
    theta_values = np.linspace(0.0, max_theta, 1 + num_thetas)
    res = edm(t, x, E=E_best, theta=list(theta_values), algorithm="smap", k=np.inf)
    s = res["summary"]
    rho = np.asarray(s["rho"])
    opt = np.nonzero(rho == rho.max())[0]
    opt_theta = [float(v) for v in np.asarray(s["theta"])[opt]]
    _alert(verbosity, f"Found optimal theta to be {opt_theta}, with rho = {list(rho[opt])}.")
    base = edm(t, x, E=E_best, theta=0, numReps=theta_reps, k=20, algorithm="smap")
    best = edm(t, x, E=E_best, theta=opt_theta, numReps=theta_reps, k=20, algorithm="smap")
    ks_stat, ks_p = ks_less_pvalue(best["stats"]["rho"], base["stats"]["rho"])
    _alert(verbosity, f"Found Kolmogorov-Smirnov test statistic to be {ks_stat} with p-value={ks_p}.")
    return opt_theta, ks_p < 0.05


def get_optimal_effect(edm, t, x, y, E_best, is_nonlinear, theta, max_lag=5, verbosity=0):
    """The effect lagged (nonlinear) or differenced (linear) by the lag whose cross-map rho is largest
    (R/easy_edm.R:231-283).  After a negative best lag the reference keeps the rows of its rho table
    selected by a lag/rho matrix; its first row is the best row with a lag >= 0."""
    rows = []
    for lag in range(-max_lag, max_lag + 2):
        res = edm(t, x, tslag(t, y, lag), E=E_best, theta=theta, algorithm="simplex", k=np.inf)
        rows.append((lag, float(np.asarray(res["summary"]["rho"])[0])))
    # order(-rho): stable, missing values last
    order = sorted(range(len(rows)), key=lambda i: (math.isnan(rows[i][1]), -rows[i][1] if not math.isnan(rows[i][1]) else 0))
    ranked = [rows[i] for i in order]
    opt_lag, opt_rho = ranked[0]
    _alert(verbosity, f"Found optimal lag to be {opt_lag} with rho={round(opt_rho, 5)}")
    if opt_lag < 0:
        opt_lag, opt_rho = next(r for r in ranked if r[0] >= 0)
        _alert(verbosity, f"Potential retrocausality, using lag of {opt_lag} with rho={round(opt_rho, 5)}")
    y_lag = tslag(t, y, opt_lag)
    if is_nonlinear:
        _alert(verbosity, f"Lagging time series using optimal lag of {opt_lag}")
        return y_lag
    _alert(verbosity, f"Differencing time series due to failed nonlinearity test (lag={opt_lag})")
    return np.asarray(y, dtype=float) - y_lag


def get_max_library(edm, t, x, y, E_best):
    """Rows of the full library manifold of y (R/easy_edm.R:287-294)."""
    res = edm(t, y, E=E_best, algorithm="simplex", full=True, saveManifolds=True)
    return int(np.asarray(res["Ms"][0]).shape[0])


def test_convergence_dist(edm, t, x, y, E_best, library_max, theta, num_reps=1000, verbosity=0):
    """rho at the largest library against the distribution of rho over shuffled replicates at a small
    library (R/easy_edm.R:395-433).  The reference labels its 0.95 quantile q975 and its 0.975 quantile
    q95; the comparisons below follow its code."""
    library_small = max(E_best + 2, library_max // 10)
    dist = edm(t, y, x, E=E_best, library=library_small, numReps=num_reps, theta=theta, algorithm="simplex",
               k=np.inf, shuffle=True)["stats"]["rho"]
    final = edm(t, y, x, E=E_best, library=library_max, theta=theta, algorithm="simplex", k=np.inf, shuffle=True)
    final_rho = float(np.asarray(final["summary"]["rho"])[0])
    q = r_quantile(dist, [0.95, 0.975])
    q975, q95 = q[0], q[1]
    rho_quantile = float(np.sum(np.asarray(dist) < final_rho)) / len(dist)
    _alert(verbosity, f"At library=E+2, found rho quantiles of {round(q975, 5)} (0.975) and {round(q95, 5)} (0.95)")
    _alert(verbosity, f"At library=max, found final rho was {round(final_rho, 5)}, i.e. quantile={rho_quantile}")
    if final_rho > q975:
        return "Strong evidence"
    if final_rho > q95:
        return "Some evidence"
    return "No evidence"


def cross_mapping(edm, t, x, y, E_best, verbosity=0):
    """CCM over 25 library sizes up to the full manifold (R/easy_edm.R:298-326)."""
    res = edm(t, y, E=E_best, algorithm="simplex", full=True, saveManifolds=True)
    library_max = int(np.asarray(res["Ms"][0]).shape[0])
    _alert(verbosity, f"The maximum library size we can use is {library_max}.")
    start = 100 if library_max >= 500 else 10
    libraries = [int(math.ceil(v)) for v in np.linspace(start, library_max, 25)]
    return edm(t, y, x, E=E_best, library=libraries, algorithm="simplex", k=np.inf, shuffle=True)


def test_convergence_monster(edm, t, x, y, E_best, library_max, theta, verbosity=0):
    """Exponential fit rho = alpha exp(-gamma L) + rho_inf of the CCM curve (R/easy_edm.R:330-391)."""
    from scipy.optimize import curve_fit
    ccm = cross_mapping(edm, t, x, y, E_best, verbosity)
    s = ccm["summary"]
    lib_ = np.asarray(s["library"], dtype=float)
    rho = np.asarray(s["rho"], dtype=float)
    gamma0 = 0.001
    rho_inf0 = rho[-1]
    alpha0 = (rho[0] - rho_inf0) / math.exp(-gamma0 * lib_[0])
    try:
        popt, _ = curve_fit(lambda L, a, g, r: a * np.exp(-g * L) + r, lib_, rho, p0=[alpha0, gamma0, rho_inf0])
        rho_inf = float(popt[2])
    except Exception:
        _alert(verbosity, "Couldn't fit an exponential curve to the rho-L values.")
        rho_inf = float(rho[-1])
    _alert(verbosity, f"The CCM final rho was {rho_inf:.2g}")
    if rho_inf > 0.7:
        return "Strong evidence"
    if rho_inf > 0.5:
        return "Some evidence"
    return "No evidence"


def easy_edm(cause, effect, time=None, data=None, direction="oneway", verbosity=1, showProgressBar=None,
             normalize=True, convergence_method="quantile", engine=None):
    """Automated test of whether `cause` CCM-causes `effect` (R/easy_edm.R:40-98); True when evidence is
    found."""
    edm = _engine(engine)
    max_theta, num_thetas, theta_reps, max_lag = 5, 100, 20, 5
    t, x, y = preprocess_inputs(data, cause, effect, time, verbosity, normalize)
    E_best = find_embedding_dimension(edm, t, x, verbosity)
    opt_theta, nonlinear = test_nonlinearity(edm, t, x, E_best, max_theta, num_thetas, theta_reps, verbosity)
    y_opt = get_optimal_effect(edm, t, x, y, E_best, nonlinear, opt_theta, max_lag, verbosity)
    library_max = get_max_library(edm, t, x, y_opt, E_best)
    if convergence_method == "parametric":
        result = test_convergence_monster(edm, t, x, y_opt, E_best, library_max, opt_theta, verbosity)
    else:
        result = test_convergence_dist(edm, t, x, y_opt, E_best, library_max, opt_theta, verbosity=verbosity)
    found = result != "No evidence"
    if data is not None:
        _alert(verbosity, f"{result} of CCM causation from {cause} to {effect} found.")
    else:
        _alert(verbosity, f"{result} of CCM causation found.")
    return found
