"""The HIP path against the reference's own golden values (tests/testthat/test-ci.R), and against
the oracle on the same commands (predictions, rho and MAE within 1e-10)."""
import numpy as np
import pytest

from oracle import oracle
from tests import scenarios
from tests.conftest import load_product

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scenario", scenarios.ALL, ids=[s.__name__ for s in scenarios.ALL])
def test_gpu_matches_reference_goldens(scenario):
    print(f"scenario {scenario.__name__}", flush=True)
    scenario(load_product().edm)


def _same_stats(a, b):
    for key in ("E", "library", "theta"):
        np.testing.assert_array_equal(a[key], b[key])
    for key in ("rho", "mae"):
        x, y = np.asarray(a[key]), np.asarray(b[key])
        assert np.array_equal(np.isnan(x), np.isnan(y))
        ok = ~np.isnan(x)
        assert np.all(np.abs(x[ok] - y[ok]) <= 1e-12 + 1e-10 * np.abs(y[ok])), (x, y)


COMMANDS = {
    "simplex_explore": lambda t, x, y: dict(t=t, x=x, E=[2, 3, 4, 5, 6], savePredictions=True),
    "smap_theta": lambda t, x, y: dict(t=t, x=x, E=4, k=np.inf, theta=[0, .5, 1, 2, 4, 8], algorithm="smap"),
    "xmap_libraries": lambda t, x, y: dict(t=t, x=y, y=x, E=3, library=[10, 20, 50, 100], numReps=5),
    "smap_coeffs": lambda t, x, y: dict(t=t, x=x, y=y, theta=0.2, algorithm="smap", saveSMAPCoeffs=True),
}


@pytest.mark.parametrize("name", sorted(COMMANDS))
def test_command_parity(name):
    fe = load_product()
    X, Y = scenarios.logistic_map(500)
    t = np.arange(1, 501, dtype=float)
    kw = COMMANDS[name](t, X, Y)
    a = fe.edm(**kw)
    b = oracle.edm(**kw)
    assert a["rc"] == b["rc"]
    assert (a["kMin"], a["kMax"]) == (b["kMin"], b["kMax"])
    _same_stats(a["stats"], b["stats"])
    for key in ("predictions", "coeffs"):
        if key in b:
            x, y = a[key], b[key]
            assert np.array_equal(np.isnan(x), np.isnan(y))
            ok = ~np.isnan(y)
            assert np.all(np.abs(x[ok] - y[ok]) <= 1e-12 + 1e-10 * np.abs(y[ok]))
