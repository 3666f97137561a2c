"""The benchmark configurations on the GPU against the CPU oracle.

cfg1 and cfg2 are compared whole (stats and predictions).  cfg4 (512 panels, S-map, dt, E=8) and
cfg5 (Lorenz-96, E_a = 40, k = 42) run at full size on the GPU; the oracle checks the manifolds whole
and the predictions of a sample of targets against the full library.  Both precision forms are held to
the same bar (the mixed form changes only the neighbour screening): predictions within
1e-12 + 1e-10 |ref|; rho and MAE likewise; manifolds and k bit-exact."""
import time

import numpy as np
import pytest

from oracle import oracle
from tests import configs
from tests.conftest import load_product

pytestmark = pytest.mark.gpu

REL, ABS = 1e-10, 1e-12


def close(a, b, rel=REL, abs_=ABS):
    """Max |a - b| after checking |a - b| <= abs_ + rel |b| and the NA pattern."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    d = np.abs(a[ok] - b[ok])
    assert np.all(d <= abs_ + rel * np.abs(b[ok])), float(d.max())
    return float(d.max()) if d.size else 0.0


def compare_commands(a, b):
    assert a["rc"] == b["rc"]
    assert (a["kMin"], a["kMax"]) == (b["kMin"], b["kMax"])
    for key in ("E", "library", "theta"):
        np.testing.assert_array_equal(a["stats"][key], b["stats"][key])
    close(a["stats"]["rho"], b["stats"]["rho"])
    close(a["stats"]["mae"], b["stats"]["mae"])


@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
def test_cfg1_simplex_explore(precision):
    fe = load_product()
    t, x = configs.cfg1_series()
    kw = dict(E=[2, 3, 4, 5, 6], savePredictions=True)
    a = fe.edm(t, x, precision=precision, **kw)
    b = oracle.edm(t, x, **kw)
    compare_commands(a, b)
    close(a["predictions"], b["predictions"])
    print("cfg1 rho per E:", np.round(a["stats"]["rho"], 5), "mae:", np.round(a["stats"]["mae"], 5))


def test_cfg2_smap_theta_sweep():
    fe = load_product()
    x, y = configs.coupled_maps(5000)
    t = np.arange(1, 5001, dtype=float)
    kw = dict(E=4, k=np.inf, theta=[0, 0.5, 1, 2, 4, 8], algorithm="smap", savePredictions=True)
    t0 = time.perf_counter()
    a = fe.edm(t, x, **kw)
    gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = oracle.edm(t, x, numThreads=16, **kw)
    cpu_s = time.perf_counter() - t0
    compare_commands(a, b)
    d = close(a["predictions"], b["predictions"])
    print(f"cfg2: 14994 predictions, GPU {gpu_s:.3f} s, oracle (16 threads) {cpu_s:.3f} s, max |d yhat| {d:.2e}; "
          f"rho {np.round(a['stats']['rho'], 5)}")


def sample_rows(n, m, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=min(m, n), replace=False))


@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
def test_cfg4_panels_smap_dt(precision):
    fe = load_product()
    t, x, panel = configs.cfg4_panels()
    kw = dict(panel=panel, E=8, dt=True, theta=[0, 1, 3], algorithm="smap", savePredictions=True,
              saveManifolds=True, precision=precision)
    t0 = time.perf_counter()
    a = fe.edm(t, x, **kw)
    gpu_s = time.perf_counter() - t0
    stages = fe.timings()
    assert a["rc"] == 0
    # manifolds against the oracle's, whole and bit for bit
    c, keep, _ = oracle.make_command(t, x, panel=panel, E=8, dt=True, theta=[0, 1, 3], algorithm="smap")
    n = len(t)
    usable = np.zeros(n, np.int32)
    nu = oracle.lib().orc_usable(oracle.C.byref(c), 8, 0, oracle.iptr(usable))
    rows = np.nonzero(usable)[0]
    lib_rows, pred_rows = rows[:nu // 2], rows[nu // 2:]

    def orc_manifold(sel, prediction_set):
        f = np.zeros(n, np.int32)
        f[sel] = 1
        out = np.zeros(n * 16)
        tg = np.zeros(n)
        m = oracle.lib().orc_manifold(oracle.C.byref(c), 8, oracle.iptr(f), int(prediction_set), 0,
                                      oracle.dptr(out), oracle.dptr(tg), n)
        return out[:m * 16].reshape(m, 16), tg[:m]

    M, Mt = orc_manifold(lib_rows, False)
    Mp, _ = orc_manifold(pred_rows, True)
    Mnan = np.where(M == 1e100, np.nan, M)
    Mpnan = np.where(Mp == 1e100, np.nan, Mp)
    np.testing.assert_array_equal(a["Ms"][0], Mnan)
    np.testing.assert_array_equal(a["Mps"][0], Mpnan)
    # predictions of sampled targets against the full library
    pick = sample_rows(len(pred_rows), 160, 4)
    opts = {"algorithm": 1, "k": 18, "thetas": [0.0, 1.0, 3.0], "panel_mode": 1, "nthreads": 16}
    panel_i = panel.astype(np.int32)
    Md = {"data": M, "targets": Mt, "panel": panel_i[lib_rows]}
    Mpd = {"data": Mp[pick], "panel": panel_i[pred_rows[pick]]}
    rp, rr, _, rk, _ = oracle.predict(Md, Mpd, opts)
    assert (rk == 18).all()
    d = close(a["predictions"][pred_rows[pick], 0], rp[2])
    print(f"cfg4 precision={precision}: {len(lib_rows)} library x {len(pred_rows)} targets, E_a=16, k=18, 3 thetas: GPU {gpu_s:.2f} s "
          f"(device: select {stages[0]:.1f} ms, merge {stages[1]:.1f} ms, S-map {stages[2]:.1f} ms); "
          f"sampled max |d yhat| {d:.2e}; rho {np.round(a['stats']['rho'], 5)}")


@pytest.mark.parametrize("precision", [0, 1], ids=["plain", "mixed"])
def test_cfg5_lorenz96_smap(precision):
    fe = load_product()
    series = configs.lorenz96()
    lib, pred = configs.cfg5_manifold(series)
    opts = {"algorithm": 1, "k": 42, "thetas": [2.0], "save_smap_coeffs": 1, "nthreads": 16, "precision": precision}
    t0 = time.perf_counter()
    gp, gr, gc, gk, gn, status = fe.predict(lib, pred, opts)
    gpu_s = time.perf_counter() - t0
    stages = fe.timings()
    assert status == 0 and (gr == 0).all() and (gk == 42).all()
    pick = sample_rows(pred["data"].shape[0], 120, 5)
    rp, rr, rc_, rk, _ = oracle.predict(lib, {"data": pred["data"][pick]}, opts)
    d = close(gp[0, pick], rp[0])
    dc = close(np.where(gc[pick] == 1e100, np.nan, gc[pick]), np.where(rc_ == 1e100, np.nan, rc_))
    yt = pred["targets"]
    rho = np.corrcoef(gp[0], yt)[0, 1]
    print(f"cfg5 precision={precision}: {lib['data'].shape[0]} library x {pred['data'].shape[0]} targets, E_a=40, k=42: GPU {gpu_s:.2f} s "
          f"(device: select {stages[0]:.1f} ms, merge {stages[1]:.1f} ms, S-map {stages[2]:.1f} ms); "
          f"sampled max |d yhat| {d:.2e}, max |d coeff| {dc:.2e}; rho {rho:.5f}")
