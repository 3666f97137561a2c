"""The batch routes beyond the persistent simplex kernels, graph replay and interruption, on the GPU.

  * the resident task route (S-map, missing values, panels, wide k, MAE distances, coefficients):
    tasks gathered and uploaded once, kernels launched per task; results against the oracle's edm();
  * hipGraph replay of a plan (persistent route and resident route): the second and later runs replay
    the captured launches; every replay, with other GPU work in between, must equal the eager run bit
    for bit (predictions, rc, k_used);
  * keep_going at the command and plan levels (src/edm.cpp:134, 377, 419): a stop returns BREAK_HIT
    and no statistics, and the next full run is unaffected.
Bar against the oracle: rc and k bit-exact, predictions, rho and MAE within 1e-12 + 1e-10 |ref|."""
import numpy as np
import pytest

from oracle import oracle
from tests import scenarios
from tests.conftest import load_product

pytestmark = pytest.mark.gpu


def _close(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    assert np.all(np.abs(a[ok] - b[ok]) <= 1e-12 + 1e-10 * np.abs(b[ok])), float(np.max(np.abs(a[ok] - b[ok])))


def _series(n=600, seed=7, missing=0.0):
    X, Y = scenarios.logistic_map(n)
    t = np.arange(1, n + 1, dtype=float)
    if missing:
        rng = np.random.default_rng(seed)
        X = X.copy()
        X[rng.random(n) < missing] = np.nan
    return t, X, Y


ROUTE_CASES = {
    # (command keywords, precision)
    "smap_ccm_libraries": (lambda t, X, Y: dict(x=Y, y=X, E=3, library=[20, 60, 200, 500], numReps=3,
                                                algorithm="smap", theta=[0.0, 1.0, 4.0]), 1),
    "smap_E_grid_k_all": (lambda t, X, Y: dict(x=X, E=[2, 3, 4], k=np.inf, theta=[0.5, 2.0], algorithm="smap"), 0),
    "smap_coeffs_mixed": (lambda t, X, Y: dict(x=X, y=Y, E=5, theta=0.7, algorithm="smap", saveSMAPCoeffs=True), 1),
    "simplex_wide_k": (lambda t, X, Y: dict(x=X, E=3, k=80, theta=[1.0, 3.0]), 0),
    "simplex_missing": (lambda t, X, Y: dict(x=X, E=[2, 4], allowMissing=True), 1),
    "simplex_mae": (lambda t, X, Y: dict(x=X, E=3, distance="mae"), 0),
}


@pytest.mark.parametrize("name", sorted(ROUTE_CASES))
def test_resident_route_matches_oracle(name):
    fe = load_product()
    make, precision = ROUTE_CASES[name]
    t, X, Y = _series(missing=0.05 if "missing" in name else 0.0)
    kw = dict(make(t, X, Y), savePredictions=True)
    plan = fe.Plan(t, precision=precision, **kw)
    assert plan.resident_batches >= 1 and plan.device_batches == 0
    assert plan.run() == 0
    a = plan.result()
    plan.close()
    b = oracle.edm(t, **kw)
    assert a["rc"] == b["rc"] and (a["kMin"], a["kMax"]) == (b["kMin"], b["kMax"])
    for key in ("E", "library", "theta"):
        np.testing.assert_array_equal(a["stats"][key], b["stats"][key])
    _close(a["stats"]["rho"], b["stats"]["rho"])
    _close(a["stats"]["mae"], b["stats"]["mae"])
    for key in ("predictions", "coeffs"):
        if key in b:
            _close(a[key], b[key])


def _panel_command():
    rng = np.random.default_rng(12)
    n = 900
    t = np.tile(np.arange(1, 301, dtype=float), 3)
    x = np.concatenate([scenarios.logistic_map(300)[0], scenarios.logistic_map(320)[1][20:],
                        rng.random(300)])
    panel = np.repeat([1, 2, 3], 300)
    return t, x, panel


def test_resident_route_panels_matches_oracle():
    fe = load_product()
    t, x, panel = _panel_command()
    for extra in (dict(panelWeight=0.3), dict(panelWeight=np.inf), dict(panelWeights=np.array(
            [[0.0, 0.2, 0.5], [0.2, 0.0, 0.1], [0.5, 0.1, 0.0]]))):
        kw = dict(panel=panel, E=3, algorithm="smap", theta=[0.0, 2.0], savePredictions=True, **extra)
        plan = fe.Plan(t, x, precision=1, **kw)
        assert plan.resident_batches == 1
        plan.run()
        a = plan.result()
        plan.close()
        b = oracle.edm(t, x, **kw)
        assert a["rc"] == b["rc"]
        _close(a["stats"]["rho"], b["stats"]["rho"])
        _close(a["predictions"], b["predictions"])


def _replays(fe, plan, other):
    """Eager run, then three graph replays with other GPU work in between; every replay's outputs must
    equal the eager run's bit for bit."""
    assert plan.run() == 0
    ref = [np.copy(v) for v in plan.fetch()]
    assert plan.set_graph(True) >= 1
    for i in range(3):
        assert plan.run() == 0
        got = plan.fetch()
        for g, r in zip(got, ref):
            assert np.array_equal(np.asarray(g).view(np.uint8), np.asarray(r).view(np.uint8)), f"replay {i}"
        other()
    return ref


def test_graph_replay_persistent_route():
    fe = load_product()
    t, X, Y = _series(3000)
    kw = dict(E=3, library=[10, 40, 200, 1000, 2900], numReps=6)
    plan = fe.Plan(t, Y, X, precision=1, **kw)
    assert plan.device_batches == 1
    rng = np.random.default_rng(0)
    M = {"data": rng.random((500, 3)), "targets": rng.random(500)}
    Mp = {"data": rng.random((100, 3))}

    def other():
        fe.predict(M, Mp, {"algorithm": 0, "k": 4, "thetas": [1.0]})

    _replays(fe, plan, other)
    res = plan.result()
    b = oracle.edm(t, Y, X, **kw)
    _close(res["stats"]["rho"], b["stats"]["rho"])
    plan.close()


def test_graph_replay_resident_route():
    fe = load_product()
    t, X, Y = _series(800)
    kw = dict(E=[2, 3], library=[30, 120, 600], numReps=2, algorithm="smap", theta=[0.0, 3.0], x=Y, y=X)
    plan = fe.Plan(t, precision=0, **kw)
    assert plan.resident_batches >= 1
    other_plan = fe.Plan(t, X, E=4, k=np.inf, theta=[1.0], algorithm="smap")

    def other():
        other_plan.run()

    _replays(fe, plan, other)
    res = plan.result()
    b = oracle.edm(t, **kw)
    _close(res["stats"]["rho"], b["stats"]["rho"])
    _close(res["stats"]["mae"], b["stats"]["mae"])
    plan.close()
    other_plan.close()


@pytest.mark.parametrize("Ea", [12, 20, 40])
def test_graph_replay_matrix_core_route(Ea):
    """Mixed-precision S-map batches whose k-NN takes the matrix-core route (E_a >= 8): the two-phase form
    with its overflow path (E_a = 12, an integer table, so that targets with more candidates than the
    screen records take the wider seed, the streaming kernel and the merge) and the seeded streaming form
    (E_a = 20 and 40), replayed as a graph with other GPU work between the replays.  Every replay's
    predictions, codes and k used equal the eager run's bit for bit, and the eager run equals the plain
    (FP64) batch's."""
    fe = load_product()
    rng = np.random.default_rng(Ea)
    NL, NP, k = 3000, 500, Ea + 2
    if Ea == 12:
        data = 1.0 + rng.integers(0, 4, size=(NL + NP, Ea)).astype(float)
    else:
        data = 0.5 + rng.random((NL + NP, Ea))
    table = {"data": data, "targets": 0.5 + rng.random(NL + NP)}
    opts = {"algorithm": 1, "k": k, "thetas": [0.0, 2.0]}
    libs, preds = [np.arange(NL)], [np.arange(NL, NL + NP)]
    print(f"graph replay, matrix-core route: Ea={Ea} k={k}", flush=True)
    mixed = fe.Batch(table, table, dict(opts, precision=fe.MIXED), libs, preds, [k])
    plain = fe.Batch(table, table, dict(opts, precision=fe.PLAIN), libs, preds, [k])
    mixed.run()
    ref = [np.copy(a) for a in mixed.fetch()[0]]
    plain.run()
    for g, r in zip(plain.fetch()[0], ref):
        assert np.array_equal(np.asarray(g).view(np.uint8), np.asarray(r).view(np.uint8))
    assert np.all(ref[1] == 0)
    assert mixed.set_graph(True) >= 1
    for i in range(3):
        mixed.run()
        for g, r in zip(mixed.fetch()[0], ref):
            assert np.array_equal(np.asarray(g).view(np.uint8), np.asarray(r).view(np.uint8)), f"replay {i}"
        plain.run()
    mixed.close()
    plain.close()


def _stopper(after):
    calls = {"n": 0}

    def keep_going():
        calls["n"] += 1
        return calls["n"] <= after

    return keep_going, calls


@pytest.mark.parametrize("route", ["persistent", "resident"])
def test_keep_going_stops_a_command(route):
    fe = load_product()
    t, X, Y = _series(2000)
    if route == "persistent":
        kw = dict(x=Y, y=X, E=3, library=[10, 100, 1000, 1990], numReps=3)
    else:
        kw = dict(x=X, E=[2, 3, 4], algorithm="smap", theta=[0.0, 1.0])
    for after in (0, 1, 2):
        keep_going, calls = _stopper(after)
        res = fe.edm(t, keep_going=keep_going, **kw)
        assert res["rc"] == 1 and len(res["stats"]["rho"]) == 0 and calls["n"] == after + 1
    # a plan stopped once, then run in full
    keep_going, calls = _stopper(1)
    plan = fe.Plan(t, keep_going=keep_going, **kw)
    assert plan.run() == 1
    assert plan.result()["rc"] == 1
    calls["n"] = -10 ** 9
    assert plan.run() == 0
    a = plan.result()
    plan.close()
    b = oracle.edm(t, **kw)
    assert a["rc"] == b["rc"] == 0
    _close(a["stats"]["rho"], b["stats"]["rho"])
