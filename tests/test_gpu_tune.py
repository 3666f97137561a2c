"""The autotuner (fedm_plan_tune) on the GPU: it searches result-invariant launch and route choices,
keeps a candidate only when its outputs equal the default's bit for bit, persists the choice per
workload class in the tuning cache, and later plans of the same class are created with it."""
import numpy as np
import pytest

from oracle import oracle
from tests import scenarios
from tests.conftest import load_product

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a).view(np.uint8)


def _check_tuned(fe, make_plan, tmp_path, expect_axes):
    plan = make_plan()
    assert plan.run() == 0
    before = [np.copy(v) for v in plan.fetch()]
    n, report = plan.tune(trials=2)
    print(report, flush=True)
    assert n >= 1 and "best:" in report and "rejected" not in report
    for axis in expect_axes:
        assert axis + ":" in report
    assert plan.run() == 0
    for a, b in zip(plan.fetch(), before):
        assert np.array_equal(_bits(a), _bits(b))
    knobs = plan.knobs()
    lines = (tmp_path / "tune.txt").read_text().splitlines()
    assert len(lines) >= 1
    # a new plan of the same workload class is created with the cached choice
    plan2 = make_plan()
    assert plan2.knobs() == knobs
    assert plan2.run() == 0
    for a, b in zip(plan2.fetch(), before):
        assert np.array_equal(_bits(a), _bits(b))
    plan.close()
    plan2.close()
    return knobs


def test_tune_persistent_route(tmp_path, monkeypatch):
    monkeypatch.setenv("FEDM_TUNE_CACHE", str(tmp_path / "tune.txt"))
    fe = load_product()
    X, Y = scenarios.logistic_map(6000)
    t = np.arange(1, 6001, dtype=float)
    kw = dict(E=3, library=[20, 80, 300, 1500, 5900], numReps=3)

    def make():
        p = fe.Plan(t, Y, X, precision=1, **kw)
        assert p.device_batches == 1
        return p

    _check_tuned(fe, make, tmp_path, ["rank_scale", "direct_alt_waves", "direct_blocks_per_cu"])
    res = make()
    res.run()
    a = res.result()
    b = oracle.edm(t, Y, X, numThreads=16, **kw)
    np.testing.assert_array_equal(a["stats"]["rho"], b["stats"]["rho"])
    res.close()


def test_tune_resident_route(tmp_path, monkeypatch):
    monkeypatch.setenv("FEDM_TUNE_CACHE", str(tmp_path / "tune.txt"))
    fe = load_product()
    X, Y = scenarios.logistic_map(3000)
    t = np.arange(1, 3001, dtype=float)
    kw = dict(E=8, algorithm="smap", theta=[0.0, 2.0], k=12)

    def make():
        p = fe.Plan(t, X, precision=1, **kw)
        assert p.resident_batches == 1
        return p

    _check_tuned(fe, make, tmp_path, ["list_blocks", "use_mfma", "mfma_blocks", "smap_lds"])
