"""The CCM sweep at its full size on the GPU against the CPU oracle.

Workload: coupled logistic maps x' = x(3.8 - 3.8x - 0.02y), y' = y(3.5 - 3.5y - 0.1x), x0 = 0.4, y0 = 0.2,
T = 20000; edm(t, y, x, E=3, library = 40 log-spaced sizes 10..19000, numReps = 100): 4000 simplex tasks
(k = 4) of 19998 targets, 79 992 000 predictions, through the plan route (fedm_plan_*), which sends the
long libraries to the FP32-screened rank tables and the short ones to the direct FP64 kernel.

Checked:
  * the mixed form's outputs (prediction, rc, k_used) equal the plain batched form's bit for bit, all
    79 992 000 of them;
  * the library and prediction rows of the first two replicates' 80 tasks equal the oracle splitter's
    (src/library_prediction_split.h, MT19937-64 stream of src/vendor/mersennetwister.h);
  * for replicates 1 and 100 (80 tasks, every library size), 200 sampled targets per task against
    oracle.predict on the oracle's own manifolds: rc and k_used bit-exact, predictions within
    1e-12 + 1e-10 |ref|;
  * rho and MAE of the first two replicates' 80 tasks against the oracle's edm() whole, to the same bar.
"""
import time

import numpy as np
import pytest

from oracle import oracle
from tests import configs
from tests.conftest import load_product
from tests.tasks import assert_predictions_close

pytestmark = pytest.mark.gpu

N = 20000
NP = 19998
E = 3
K = 4


def _inputs():
    x, y = configs.coupled_maps(N)
    t = np.arange(1, N + 1, dtype=float)
    return t, x, y, configs.ccm_libraries()


def test_ccm_sweep_full_size_matches_oracle():
    fe = load_product()
    t, x, y, libs = _inputs()
    kw = dict(E=E, library=libs, numReps=100)
    outs, stats, sample_rows = {}, {}, {}
    sample_tasks = list(range(0, 40)) + list(range(3960, 4000))
    for name, prec in (("mixed", fe.MIXED), ("plain", fe.PLAIN)):
        print(f"ccm sweep: {name} plan", flush=True)
        plan = fe.Plan(t, y, x, precision=prec, **kw)
        assert plan.device_batches == 1
        assert plan.num_predictions == 100 * 40 * NP
        t0 = time.perf_counter()
        assert plan.run() == 0
        run_s = time.perf_counter() - t0
        outs[name] = plan.fetch()
        res = plan.result()
        assert res["rc"] == 0 and (res["kMin"], res["kMax"]) == (K, K)
        stats[name] = res["stats"]
        if name == "plain":
            tasks = plan.tasks(rows=True, indices=sample_tasks)
            sample_rows = {i: d for i, d in zip(sample_tasks, tasks)}
        print(f"ccm sweep: {name} run {run_s * 1e3:.1f} ms, redone {fe.exact_redos()}", flush=True)
        plan.close()

    # mixed against plain batched: every output bit for bit
    (pm, rm, km), (pp, rp, kp) = outs["mixed"], outs["plain"]
    assert pm.size == 79_992_000
    np.testing.assert_array_equal(rm, rp)
    np.testing.assert_array_equal(km, kp)
    assert np.array_equal(pm.view(np.int64), pp.view(np.int64)), \
        f"{int(np.sum(pm != pp))} predictions differ, max |d| {np.max(np.abs(pm - pp)):.3e}"
    assert (rp == 0).all() and (kp == K).all()

    # the first two replicates' splits against the oracle splitter
    c, keep, _ = oracle.make_command(t, y, x, E=E, library=libs, numReps=2)
    usable = np.zeros(N, np.int32)
    nu = oracle.lib().orc_usable(oracle.C.byref(c), E, 0, oracle.iptr(usable))
    assert nu == NP
    lib_m = np.zeros(80 * N, np.int32)
    pred_m = np.zeros(80 * N, np.int32)
    L = libs.astype(np.int32)
    ntasks = oracle.lib().orc_split_masks(N, oracle.iptr(usable), 0, 0, 1, 0, 0, 2, 1, 40, oracle.iptr(L),
                                          oracle.iptr(lib_m), oracle.iptr(pred_m), 80)
    assert ntasks == 80
    lib_m = lib_m.reshape(80, N)
    pred_m = pred_m.reshape(80, N)
    for i in range(40):
        d = sample_rows[i]
        assert d["library"] == libs[i] and d["num_lib"] == libs[i] and d["num_pred"] == NP
        np.testing.assert_array_equal(d["lib_rows"], np.nonzero(lib_m[i])[0])
        np.testing.assert_array_equal(d["pred_rows"], np.nonzero(pred_m[i])[0])

    # sampled targets of replicates 1 and 100 against oracle.predict on the oracle's manifolds
    def orc_manifold(rows, prediction_set):
        f = np.zeros(N, np.int32)
        f[rows] = 1
        out = np.zeros(N * E)
        tg = np.zeros(N)
        m = oracle.lib().orc_manifold(oracle.C.byref(c), E, oracle.iptr(f), int(prediction_set), 0,
                                      oracle.dptr(out), oracle.dptr(tg), N)
        return out[:m * E].reshape(m, E), tg[:m]

    rng = np.random.default_rng(3)
    worst = 0.0
    for i in sample_tasks:
        d = sample_rows[i]
        M, Mt = orc_manifold(d["lib_rows"], False)
        Mp, _ = orc_manifold(d["pred_rows"], True)
        assert M.shape[0] == d["num_lib"] and Mp.shape[0] == NP
        pick = np.sort(rng.choice(NP, size=200, replace=False))
        rp_, rr_, _, rk_, _ = oracle.predict({"data": M, "targets": Mt}, {"data": Mp[pick]},
                                             {"algorithm": 0, "k": K, "thetas": [1.0]})
        o = i * NP
        np.testing.assert_array_equal(rp[o + pick], rr_[0])
        np.testing.assert_array_equal(kp[o + pick], rk_)
        worst = max(worst, assert_predictions_close(pp[o + pick], rp_[0]))

    # rho and MAE of the first two replicates against the oracle's edm()
    t0 = time.perf_counter()
    b = oracle.edm(t, y, x, E=E, library=libs, numReps=2, numThreads=16)
    cpu_s = time.perf_counter() - t0
    assert b["rc"] == 0
    for key in ("rho", "mae"):
        a_ = stats["plain"][key][:80]
        b_ = b["stats"][key]
        assert np.all(np.abs(a_ - b_) <= 1e-12 + 1e-10 * np.abs(b_)), key
        np.testing.assert_array_equal(stats["mixed"][key], stats["plain"][key])
    np.testing.assert_array_equal(stats["plain"]["library"][:80], b["stats"]["library"])
    print(f"ccm sweep: 79 992 000 predictions mixed == plain bit for bit; 80 tasks x 200 sampled targets vs "
          f"the oracle, max |d yhat| {worst:.2e}; rho/MAE of 80 tasks vs the oracle's edm ({cpu_s:.1f} s on "
          f"16 threads)", flush=True)
