"""The batched route (edm_predict_batch: one persistent launch over many tasks sharing a point table)
against the CPU oracle task by task, in the plain (FP64) and mixed (FP32 screening + FP64 re-rank)
forms.  Bar for both: k_used and retcodes bit-exact, predictions within 1e-12 + 1e-10 |ref|."""
import zlib

import numpy as np
import pytest

from oracle import oracle
from tests.conftest import load_product
from tests.tasks import assert_predictions_close

pytestmark = pytest.mark.gpu


def ccm_like(seed, n, Ea, sizes, reps, integer=False):
    rng = np.random.default_rng(seed)
    table = rng.integers(0, 5, size=(n, Ea)).astype(float) if integer else rng.random((n, Ea))
    targets = rng.random(n)
    L = {"data": table, "targets": targets}
    libs, preds = [], []
    for r in range(reps):
        for s in sizes:
            libs.append(np.sort(rng.choice(n, size=min(s, n), replace=False)))
            preds.append(np.arange(n))
    return L, libs, preds


@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
@pytest.mark.parametrize("case", ["E3_k4", "E5_k6_integer", "E8_k9_thetas", "tiny_libraries"])
def test_batch_matches_oracle(precision, case):
    fe = load_product()
    spec = {"E3_k4": (3, 4, [10, 37, 150, 900], 3, False, (1.0,)),
            "E5_k6_integer": (5, 6, [12, 80, 400], 2, True, (1.0,)),
            "E8_k9_thetas": (8, 9, [20, 300, 1200], 2, False, (0.0, 1.0, 4.0)),
            "tiny_libraries": (2, 3, [1, 2, 3, 4, 6], 2, False, (1.0,))}[case]
    Ea, k, sizes, reps, integer, thetas = spec
    L, libs, preds = ccm_like(zlib.crc32(case.encode()) % 997, 1500, Ea, sizes, reps, integer)
    opts = {"algorithm": 0, "k": k, "thetas": np.asarray(thetas), "precision": precision}
    print(f"case {case} precision {precision}: {len(libs)} tasks", flush=True)
    got = fe.predict_batch(L, L, opts, libs, preds, [k] * len(libs))
    for t, (lib, pr) in enumerate(zip(libs, preds)):
        M = {"data": L["data"][lib], "targets": L["targets"][lib]}
        Mp = {"data": L["data"][pr]}
        rp, rr, _, rk, _ = oracle.predict(M, Mp, dict(opts, k=k))
        gp, gr, gk = got[t]
        np.testing.assert_array_equal(gr, rr)
        np.testing.assert_array_equal(gk, rk)
        assert_predictions_close(gp, rp)


@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
def test_ccm_command_plan_matches_oracle(precision):
    """edm(t, y, x, library=..., numReps=...) through the plan route vs the oracle's run_command."""
    fe = load_product()
    from tests.scenarios import logistic_map
    X, Y = logistic_map(800)
    t = np.arange(1, 801, dtype=float)
    kw = dict(E=3, library=[10, 25, 60, 150, 400, 790], numReps=4, savePredictions=True)
    plan = fe.Plan(t, Y, X, precision=precision, **kw)
    assert plan.device_batches == 1
    plan.run()
    plan.run()
    a = plan.result()
    b = oracle.edm(t, Y, X, **kw)
    assert a["rc"] == b["rc"] and (a["kMin"], a["kMax"]) == (b["kMin"], b["kMax"])
    np.testing.assert_array_equal(a["stats"]["library"], b["stats"]["library"])
    assert np.all(np.abs(a["stats"]["rho"] - b["stats"]["rho"]) <= 1e-12 + 1e-10 * np.abs(b["stats"]["rho"]))
    assert np.all(np.abs(a["stats"]["mae"] - b["stats"]["mae"]) <= 1e-12 + 1e-10 * np.abs(b["stats"]["mae"]))
    x, y = a["predictions"], b["predictions"]
    assert np.array_equal(np.isnan(x), np.isnan(y))
    ok = ~np.isnan(y)
    assert np.all(np.abs(x[ok] - y[ok]) <= 1e-12 + 1e-10 * np.abs(y[ok]))
