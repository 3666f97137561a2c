"""Pins the CPU oracle where the reference's golden rho values leave it open: closed forms, brute
force on tiny inputs, library routines (numpy/scipy) and the reference's own RNG stream.  Each check
is chosen so that a plausible slip (a dropped term, a wrong index, a transposed operand, a missing
sqrt) fails it."""
import json
import os

import numpy as np
import pytest
import scipy.optimize

from oracle import oracle
from tests.conftest import GOLDEN
from tests.tasks import wasserstein_task

MISSING = 1.0e100


def test_mt64_matches_reference_stream():
    """MtRng64::getReal2 (src/vendor/mersennetwister.h) compiled from the reference: tempered seed first,
    then the MT19937-64 stream."""
    ref = json.load(open(os.path.join(GOLDEN, "mt64_streams.json")))
    for seed, vals in ref.items():
        got = oracle.mt64_stream(int(seed), len(vals))
        np.testing.assert_array_equal(got, np.asarray(vals))


def brute_lp(M, Mp, i, mae=False, missing_distance=0.0, idw=0.0, lp=None, pp=None):
    out = []
    for j, x in enumerate(M):
        acc = 0.0
        if idw > 0:
            acc += idw * (lp[j] != pp[i])
        for a, b in zip(x, Mp[i]):
            if a == MISSING or b == MISSING:
                if missing_distance == 0:
                    acc = None
                    break
                d = missing_distance
            else:
                d = a - b
            acc += abs(d) / len(x) if mae else d * d
        if acc is None or acc == 0:
            continue
        out.append((acc if mae else np.sqrt(acc), j))
    return out


@pytest.mark.parametrize("mae", [False, True])
def test_knn_brute_force(mae):
    rng = np.random.default_rng(3)
    M = rng.integers(0, 3, (40, 2)).astype(float)  # many exact ties
    Mp = rng.integers(0, 3, (15, 2)).astype(float)
    k = 5
    opts = {"algorithm": 0, "k": k, "thetas": [0.0], "distance": 0 if mae else 1}
    pred, rc, _, ku, nn = oracle.predict({"data": M, "targets": np.arange(40.0)}, {"data": Mp}, opts, nn_cap=k)
    for i in range(15):
        cand = brute_lp(M, Mp, i, mae=mae)
        order = sorted(cand)[:k]  # (distance, position) ascending
        assert list(nn[i][:len(order)]) == [j for _, j in order]
        # theta = 0: plain mean of the neighbours' targets (src/edm.cpp:721-730)
        assert pred[0, i] == pytest.approx(np.mean([j for _, j in order]), rel=1e-14)


def test_idw_penalty_sits_inside_the_sqrt():
    M = np.array([[0.0], [3.0]])
    Mp = np.array([[0.0]])
    opts = {"algorithm": 0, "k": 2, "thetas": [1.0], "panel_mode": 1, "idw": 16.0}
    Md = {"data": M, "targets": np.array([1.0, 2.0]), "panel": np.array([1, 0], np.int32)}
    Mpd = {"data": Mp, "panel": np.array([0], np.int32)}
    pred, rc, _, ku, nn = oracle.predict(Md, Mpd, opts, nn_cap=2)
    # d0 = sqrt(16 + 0) = 4, d1 = sqrt(0 + 9) = 3: both kept (k == #valid -> library order)
    w = np.exp(-1.0 * np.array([4.0, 3.0]) / 3.0)
    assert pred[0, 0] == pytest.approx((1.0 * w[0] + 2.0 * w[1]) / w.sum(), rel=1e-14)


def test_missing_coordinates():
    M = np.array([[0.0, MISSING], [1.0, 1.0], [2.0, 2.0]])
    Mp = np.array([[0.0, 0.0]])
    for md, expect in [(0.0, [1, 2]), (0.5, [0, 1, 2])]:
        opts = {"algorithm": 0, "k": 3, "thetas": [0.0], "has_missing": 1, "missing_distance": md}
        _, _, _, ku, nn = oracle.predict({"data": M, "targets": np.zeros(3)}, {"data": Mp}, opts, nn_cap=3)
        assert list(nn[0][:ku[0]]) == expect
        assert brute_lp(M, Mp, 0, missing_distance=md)[0][1] == expect[0]


def test_simplex_k1_is_nearest_target():
    rng = np.random.default_rng(1)
    M = rng.random((30, 3))
    Mp = rng.random((10, 3))
    t = rng.random(30)
    pred, *_ = oracle.predict({"data": M, "targets": t}, {"data": Mp}, {"algorithm": 0, "k": 1, "thetas": [2.0]})
    nearest = [np.argmin(((M - p) ** 2).sum(1)) for p in Mp]
    np.testing.assert_array_equal(pred[0], t[nearest])


def test_smap_theta0_all_is_global_ols():
    """theta = 0, k = all: every weight is 1, so S-map is ordinary least squares on [1, X]."""
    rng = np.random.default_rng(7)
    M = rng.random((60, 3))
    t = 0.3 + M @ np.array([1.0, -2.0, 0.5]) + 0.01 * rng.standard_normal(60)
    Mp = rng.random((8, 3))
    pred, rc, coeffs, ku, _ = oracle.predict({"data": M, "targets": t}, {"data": Mp},
                                             {"algorithm": 1, "k": -1, "thetas": [0.0], "save_smap_coeffs": 1})
    X = np.hstack([np.ones((60, 1)), M])
    beta = np.linalg.lstsq(X, t, rcond=None)[0]
    np.testing.assert_allclose(pred[0], np.hstack([np.ones((8, 1)), Mp]) @ beta, rtol=1e-10)
    np.testing.assert_allclose(coeffs[0], beta, rtol=1e-9)


def test_smap_rank_deficient_gives_minimum_norm():
    """A duplicated column makes X^T X singular: the JacobiSVD solve with Eigen's cutoff returns the
    minimum-norm solution, i.e. pinv(X^T X) X^T y."""
    rng = np.random.default_rng(11)
    a = rng.random(40)
    M = np.stack([a, a, rng.random(40)], 1)
    t = 1.0 + 2.0 * a + rng.random(40) * 0.1
    Mp = rng.random((5, 3))
    pred, rc, coeffs, *_ = oracle.predict({"data": M, "targets": t}, {"data": Mp},
                                          {"algorithm": 1, "k": -1, "thetas": [0.0], "save_smap_coeffs": 1})
    X = np.hstack([np.ones((40, 1)), M])
    A = X.T @ X
    beta = np.linalg.pinv(A, rcond=4 * np.finfo(float).eps) @ (X.T @ t)
    np.testing.assert_allclose(coeffs[0], beta, rtol=1e-7, atol=1e-9)
    assert coeffs[0][1] == pytest.approx(coeffs[0][2], rel=1e-9)  # the two equal columns share weight


def test_jacobi_svd_singular_values_and_solve():
    rng = np.random.default_rng(5)
    for n in [2, 5, 17, 41]:
        X = rng.standard_normal((n + 3, n))
        A = X.T @ X
        b = rng.standard_normal(n)
        x, sv, rank = oracle.jacobi_solve(A, b)
        np.testing.assert_allclose(sv, np.linalg.svd(A, compute_uv=False), rtol=1e-10)
        assert rank == n
        np.testing.assert_allclose(A @ x, b, rtol=1e-8, atol=1e-10)
    # cutoff: singular values below s0 * n * eps are dropped
    A = np.diag([1.0, 1e-20, 0.5])
    x, sv, rank = oracle.jacobi_solve(A, np.array([1.0, 1.0, 1.0]))
    assert rank == 2
    np.testing.assert_allclose(x, [1.0, 0.0, 2.0])


def test_correlation_and_mae_against_numpy():
    rng = np.random.default_rng(2)
    a, b = rng.random(300), rng.random(300)
    assert oracle.lib().orc_correlation(300, oracle.dptr(a), oracle.dptr(b)) == pytest.approx(
        np.corrcoef(a, b)[0, 1], rel=1e-12)
    assert oracle.lib().orc_mae(300, oracle.dptr(a), oracle.dptr(b)) == pytest.approx(np.mean(np.abs(a - b)),
                                                                                    rel=1e-12)
    c = np.full(10, 0.5)
    assert oracle.lib().orc_correlation(10, oracle.dptr(c), oracle.dptr(a[:10].copy())) == MISSING


def test_emd_against_assignment_and_linprog():
    rng = np.random.default_rng(9)
    for n in [2, 3, 5]:
        C = rng.random((n, n))
        r, c = scipy.optimize.linear_sum_assignment(C)
        assert oracle.emd(C) == pytest.approx(C[r, c].sum() / n, rel=1e-12)
    for li, lj in [(2, 3), (4, 3), (1, 5)]:
        C = rng.random((li, lj))
        A_eq = []
        for i in range(li):
            row = np.zeros(li * lj)
            row[i * lj:(i + 1) * lj] = 1
            A_eq.append(row)
        for j in range(lj):
            row = np.zeros(li * lj)
            row[j::lj] = 1
            A_eq.append(row)
        b_eq = [1 / li] * li + [1 / lj] * lj
        lp = scipy.optimize.linprog(C.ravel(), A_eq=np.array(A_eq), b_eq=b_eq, bounds=(0, None))
        assert oracle.emd(C) == pytest.approx(lp.fun, rel=1e-9)


def test_wasserstein_lags_is_sorted_matching():
    """Lags only: the cost |x_a - y_b| on the line makes the EMD with uniform weights the 1-D W1 distance
    of the two empirical distributions (scipy.stats.wasserstein_distance, sorted matching)."""
    from scipy.stats import wasserstein_distance
    M, Mp, o = wasserstein_task(1, 40, 6, 7, 4)
    for i in range(6):
        inds, d = oracle.distances(M, Mp, o, i)
        ref = np.array([wasserstein_distance(M["data"][j], Mp["data"][i]) for j in range(40)])
        np.testing.assert_array_equal(inds, np.arange(40))
        np.testing.assert_allclose(d, ref, rtol=1e-12)


def test_wasserstein_skipped_missing_columns():
    """missingdistance == 0: columns with a MISSING entry are dropped, so len_i != len_j; the value is
    the W1 distance of the remaining values (unequal sample sizes)."""
    from scipy.stats import wasserstein_distance
    M, Mp, o = wasserstein_task(2, 60, 6, 6, 4, missing=0.25)
    for i in range(6):
        inds, d = oracle.distances(M, Mp, o, i)
        y = Mp["data"][i][Mp["data"][i] != MISSING]
        exp_i, exp_d = [], []
        for j in range(60):
            x = M["data"][j][M["data"][j] != MISSING]
            if x.size and y.size:
                exp_i.append(j)
                exp_d.append(wasserstein_distance(x, y))
        np.testing.assert_array_equal(inds, exp_i)
        np.testing.assert_allclose(d, exp_d, rtol=1e-12)


def test_wasserstein_constant_terms():
    """Unlagged extras, the panel penalty and a missingdistance column add constants: uniform marginals
    carry unit mass, so each constant shifts the EMD by itself."""
    from scipy.stats import wasserstein_distance
    E = 5
    M, Mp, o = wasserstein_task(3, 50, 5, E, 4, unlagged=2, panels=2, idw=0.35)
    X, Y = M["data"], Mp["data"]
    for i in range(5):
        inds, d = oracle.distances(M, Mp, o, i)
        ref = np.array([wasserstein_distance(X[j, :E], Y[i, :E]) + np.abs(X[j, E:] - Y[i, E:]).sum()
                        + 0.35 * (M["panel"][j] != Mp["panel"][i]) for j in range(50)])
        np.testing.assert_allclose(d, ref, rtol=1e-12)
    # a MISSING column with missingdistance m: that column costs m against every column, so the optimum
    # pairs it with the column that is left over
    M, Mp, o = wasserstein_task(4, 1, 1, 2, 1, missing_distance=0.8)
    M["data"][0] = [0.1, MISSING]
    Mp["data"][0] = [0.3, 0.9]
    inds, d = oracle.distances(M, Mp, o, 0)
    C = np.array([[0.2, 0.8], [0.8, 0.8]])
    r, c = scipy.optimize.linear_sum_assignment(C)
    assert d[0] == pytest.approx(C[r, c].sum() / 2, rel=1e-14)


def test_wasserstein_categorical_row_and_skip_other():
    """A CheckSame extra row adds [a != b] to cell (a, b); with idw < 0 other panels are skipped."""
    E = 3
    M, Mp, o = wasserstein_task(5, 30, 4, E, 4, lagged=1, categorical=True, panels=3, idw=-1.0)
    X, Y = M["data"], Mp["data"]
    for i in range(4):
        inds, d = oracle.distances(M, Mp, o, i)
        same = np.nonzero(M["panel"] == Mp["panel"][i])[0]
        exp_i, exp_d = [], []
        for j in same:
            Cm = np.abs(X[j, :E][:, None] - Y[i, :E][None, :]) + (X[j, E:2 * E][:, None] != Y[i, E:2 * E][None, :])
            r, c = scipy.optimize.linear_sum_assignment(Cm)
            v = Cm[r, c].sum() / E
            if v != 0:
                exp_i.append(j)
                exp_d.append(v)
        np.testing.assert_array_equal(inds, exp_i)
        np.testing.assert_allclose(d, exp_d, rtol=1e-12)


def test_wasserstein_dt_row_scaled_by_aspect():
    """dt row: cell (a, b) adds gamma |dt_a - dt_b|, gamma = aspect (max x - min x + 1e-6)/(max dt + 1e-6)
    of the library point (src/distances.cpp:344-366)."""
    E = 4
    M, Mp, o = wasserstein_task(6, 20, 3, E, 4, dt=True, aspect_ratio=0.6)
    X, Y = M["data"], Mp["data"]
    for i in range(3):
        _, d = oracle.distances(M, Mp, o, i)
        for j in range(20):
            x, tx, y, ty = X[j, :E], X[j, E:], Y[i, :E], Y[i, E:]
            g = 0.6 * (x.max() - x.min() + 1e-6) / (tx.max() + 1e-6)
            Cm = np.abs(x[:, None] - y[None, :]) + g * np.abs(tx[:, None] - ty[None, :])
            r, c = scipy.optimize.linear_sum_assignment(Cm)
            assert d[j] == pytest.approx(Cm[r, c].sum() / E, rel=1e-12)


def command(t, x, **kw):
    c, keep, _ = oracle.make_command(t, x, **kw)
    return c, keep


def test_manifold_lags_and_targets_by_hand():
    t = np.arange(1.0, 11.0)
    x = 10.0 * np.arange(1.0, 11.0)
    c, keep = command(t, x, E=3, p=1)
    n = 10
    f = np.ones(n, np.int32)
    out = np.zeros(n * 3)
    tg = np.zeros(n)
    m = oracle.lib().orc_manifold(oracle.C.byref(c), 3, oracle.iptr(f), 1, 0, oracle.dptr(out), oracle.dptr(tg), n)
    pts = out.reshape(n, 3)
    assert m == 10
    np.testing.assert_array_equal(pts[5], [60.0, 50.0, 40.0])  # lag j of row i is row i - j
    assert pts[0][1] == MISSING and pts[1][2] == MISSING
    np.testing.assert_array_equal(tg[:9], x[1:])  # target p = 1 step ahead
    assert tg[9] == MISSING


def test_manifold_lagged_extras_by_hand():
    """Point layout [x lags | lagged extras, E lags each | unlagged extras] (src/manifold.cpp:225-233): with
    numExtrasLagged = 1 the first extra enters at every lag, the second once, a missing lag missing in both."""
    t = np.arange(1.0, 9.0)
    x = np.arange(1.0, 9.0)
    a = 100.0 + np.arange(8.0)
    b = 200.0 + np.arange(8.0)
    c, keep = command(t, x, E=3, p=1, extras=[a, b], extrasLagged=1)
    n, Ea = 8, 3 + 3 + 1
    f = np.ones(n, np.int32)
    out = np.zeros(n * Ea)
    tg = np.zeros(n)
    oracle.lib().orc_manifold(oracle.C.byref(c), 3, oracle.iptr(f), 1, 0, oracle.dptr(out), oracle.dptr(tg), n)
    pts = out.reshape(n, Ea)
    np.testing.assert_array_equal(pts[5], [6.0, 5.0, 4.0, 105.0, 104.0, 103.0, 205.0])
    np.testing.assert_array_equal(pts[1][[2, 5]], [MISSING, MISSING])  # lag 2 of row 1 is absent
    assert pts[1][6] == 201.0


def test_manifold_panel_boundary_and_gap():
    t = np.array([1.0, 2, 3, 4, 1, 2, 3, 5])
    x = np.arange(1.0, 9.0)
    panel = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    c, keep = command(t, x, E=2, panel=panel, p=1)
    f = np.ones(8, np.int32)
    out = np.zeros(16)
    tg = np.zeros(8)
    oracle.lib().orc_manifold(oracle.C.byref(c), 2, oracle.iptr(f), 1, 0, oracle.dptr(out), oracle.dptr(tg), 8)
    pts = out.reshape(8, 2)
    assert pts[4][1] == MISSING  # first row of panel 1 has no lag across the boundary
    assert pts[7][1] == MISSING  # t=5 has no t=4 observation in its panel
    assert tg[3] == MISSING and tg[6] == MISSING  # no target across the boundary or the gap


def test_dt_columns_constant_on_regular_grid():
    t = np.arange(1.0, 21.0) * 2.0
    x = np.sin(np.arange(20.0)) + 2.0
    c, keep = command(t, x, E=3, dt=True, dtWeight=1.0, p=1)
    f = np.ones(20, np.int32)
    out = np.zeros(20 * 6)
    tg = np.zeros(20)
    oracle.lib().orc_manifold(oracle.C.byref(c), 3, oracle.iptr(f), 1, 0, oracle.dptr(out), oracle.dptr(tg), 20)
    pts = out.reshape(20, 6)
    np.testing.assert_array_equal(pts[10][3:], [2.0, 2.0, 2.0])  # t_target - t_now, then t_prev - t_lag


def test_fixed_size_library_draws_exactly_library_rows():
    n = 300
    usable = np.ones(n, np.int32)
    usable[:5] = 0
    libs = np.array([10, 50, 290], np.int32)
    lib_out = np.zeros(3 * 2 * n, np.int32)
    pred_out = np.zeros(3 * 2 * n, np.int32)
    ntask = oracle.lib().orc_split_masks(n, oracle.iptr(usable), 0, 0, 1, 0, 0, 2, 1, 3, oracle.iptr(libs),
                                         oracle.iptr(lib_out), oracle.iptr(pred_out), 6)
    assert ntask == 6
    lib_out = lib_out.reshape(6, n)
    pred_out = pred_out.reshape(6, n)
    assert list(lib_out.sum(1)) == [10, 50, 290, 10, 50, 290]
    assert (pred_out == usable).all()
    assert not lib_out[:, :5].any()
    # the draw is the library smallest of the uniforms (src/library_prediction_split.h:236-263)
    u = oracle.mt64_stream(0, 295)
    np.testing.assert_array_equal(np.nonzero(lib_out[0][5:])[0], np.sort(np.argsort(u)[:10]))


def test_crossfold_groups_partition_usable_rows():
    n = 103
    usable = np.ones(n, np.int32)
    lib_out = np.zeros(5 * n, np.int32)
    pred_out = np.zeros(5 * n, np.int32)
    oracle.lib().orc_split_masks(n, oracle.iptr(usable), 1, 0, 0, 5, 0, 1, 1, 0, None, oracle.iptr(lib_out),
                                 oracle.iptr(pred_out), 5)
    pred_out = pred_out.reshape(5, n)
    lib_out = lib_out.reshape(5, n)
    assert (pred_out.sum(0) == 1).all()  # every usable row is predicted exactly once
    assert ((lib_out + pred_out) == 1).all()


def _exp_grid():
    """Dense arguments over the weights' domain: exp(-theta d / d_ref) takes x <= 0 down to where it
    underflows (-745.13), plus the reduction's case boundaries and the special values."""
    rng = np.random.default_rng(31)
    x = np.concatenate([
        -np.logspace(-12, np.log10(745.2), 120000),
        rng.uniform(-745.2, 0.0, 60000),
        rng.uniform(-1.2, 0.0, 40000),
        -np.arange(1, 1076) * np.log(2.0),          # multiples of ln 2 (k changes)
        np.array([0.0, -0.0, -3.725290298461914e-09, -3.7252902e-09, -0.5 * np.log(2.0), -1.5 * np.log(2.0),
                  -708.3964185322641, -709.0, -744.44007192138126, -745.0, -745.13321910194110842,
                  -745.1332191019412, -746.0, -1e308, 1e-300, 0.5, 2.0, 700.0]),
    ])
    return x


def test_weight_exp_within_one_ulp_of_libm():
    """The oracle's FDLIBM exp (the weights' exp on the CPU and the GPU) against math.exp over a dense grid
    of the weights' domain: within 1 ulp everywhere, exact at the special values."""
    import math
    L = oracle.lib()
    L.orc_weight_exp.restype = oracle.C.c_double
    L.orc_weight_exp.argtypes = [oracle.C.c_double]
    x = _exp_grid()
    got = np.array([L.orc_weight_exp(float(v)) for v in x])
    ref = np.array([math.exp(float(v)) for v in x])
    ulp = np.spacing(np.abs(ref))
    ulp[ref == 0] = np.spacing(0.0)
    assert np.all(np.abs(got - ref) <= ulp), float(np.max(np.abs(got - ref) / ulp))
    assert L.orc_weight_exp(0.0) == 1.0 and L.orc_weight_exp(-0.0) == 1.0
    assert L.orc_weight_exp(-746.0) == 0.0 and L.orc_weight_exp(-1e308) == 0.0
    assert np.isnan(L.orc_weight_exp(float("nan")))
    assert L.orc_weight_exp(float("-inf")) == 0.0
    # FDLIBM's method is within 1 ulp but not correctly rounded: about 6 % of these arguments are one ulp
    # away from glibc's value (a weight's last bit, which the reference leaves to its libm/Eigen exp)
    assert np.mean(got != ref) < 0.1


def test_panel_weight_matrix_brute_force():
    """The panel-weight matrix branch (src/distances.cpp:182-183, keys built at src/r.cpp:204-220): the
    penalty W[(library panel, prediction panel)] joins the squared distance inside the sqrt, as a float,
    keyed by position in the run-length-distinct panel ids, a later run of a panel overwriting an earlier
    one.  Panel 5 appears in two runs and W is asymmetric, so a transposed key or the first run's
    values fail the check."""
    rng = np.random.default_rng(21)
    blocks = [(5, 30), (2, 25), (5, 20), (9, 28)]
    t, x, panel = [], [], []
    for pid, n in blocks:
        t += list(range(1, n + 1))
        x += list(rng.uniform(0.1, 0.9, n))
        panel += [pid] * n
    t, x, panel = np.array(t, float), np.array(x), np.array(panel)
    uniq = [5, 2, 5, 9]
    W = rng.uniform(0.0, 0.4, (4, 4))
    k, theta = 4, 1.3
    res = oracle.edm(t, x, panel=panel, E=2, k=k, theta=[theta], panelWeights=W, savePredictions=True)
    assert res["rc"] == 0
    key = {}
    for i, a in enumerate(uniq):
        for j, b in enumerate(uniq):
            key[(a, b)] = float(np.float32(W[i, j]))
    n = len(t)
    same = lambda r, q: 0 <= q < n and panel[q] == panel[r] and t[q] == t[r] + (q - r)
    usable = [r for r in range(n) if same(r, r - 1)]
    nu = len(usable)
    lib = [r for r in usable[:nu // 2] if same(r, r + 1)]
    pred_rows = usable[nu // 2:]
    expect = np.full(n, np.nan)
    for i in pred_rows:
        cand = []
        for pos, j in enumerate(lib):
            d2 = key[(panel[j], panel[i])]
            d2 += (x[j] - x[i]) ** 2
            d2 += (x[j - 1] - x[i - 1]) ** 2
            if d2 != 0:
                cand.append((np.sqrt(d2), pos, j))
        nn = sorted(cand)[:k]
        dmin = nn[0][0]
        w = np.array([np.exp(-theta * d / dmin) for d, _, _ in nn])
        expect[i] = sum(x[j + 1] * (wi / w.sum()) for (_, _, j), wi in zip(nn, w))
    got = res["predictions"][:, 0]
    ok = ~np.isnan(expect)
    assert np.array_equal(np.isnan(got), ~ok)
    np.testing.assert_allclose(got[ok], expect[ok], rtol=1e-13, atol=0)
    # the transposed key gives a different answer (the check can fail)
    res_t = oracle.edm(t, x, panel=panel, E=2, k=k, theta=[theta], panelWeights=W.T.copy(), savePredictions=True)
    assert np.max(np.abs(res_t["predictions"][ok, 0] - expect[ok])) > 1e-6
