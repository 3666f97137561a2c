"""The reference's end-to-end test scenarios (tests/testthat/test-ci.R), written against any
``edm`` callable that mirrors R's edm().  Each scenario asserts the golden rho values of the
reference test, with its tolerance (max |d rho| < 1e-4 absolute, test-ci.R:23-33) or its bands."""
import numpy as np

from tests.conftest import (check_edm_result, check_noisy_edm_result, expect_approx_equal, golden_vectors,
                            logistic_map, r_sd, tsdiff, tslag)


def simple_data():
    # test-ci.R:81-86
    X, Y = logistic_map(500)
    x, y = X[299:500], Y[299:500]
    t = 299.0 + np.arange(1, len(x) + 1)
    return t, x, y


def missing_data():
    # test-ci.R:219-238
    t, x, y = simple_data()
    u = np.asarray(golden_vectors()["u_seed12345678"])
    keep = u >= 0.1
    t, x, y, u = t[keep].copy(), x[keep].copy(), y[keep].copy(), u[keep].copy()
    x[u < 0.2] = np.nan
    t[np.fmod(t, 19) == 1] = np.nan
    return t, x, y, u


def bigger_data(missing):
    # test-ci.R:293-330
    X, Y = logistic_map(100)
    t = np.arange(1, 101, dtype=float)
    x, y = X.copy(), Y.copy()
    if not missing:
        return t, x, y, None, None
    g = golden_vectors()
    u = np.asarray(g["u_bigger"])
    u1 = np.asarray(g["u1_seed1"])
    keep = u >= 0.1
    t, x, y, u, u1 = t[keep].copy(), x[keep].copy(), y[keep].copy(), u[keep].copy(), u1[keep].copy()
    x[u < 0.2] = np.nan
    u1[np.fmod(t, 7) == 1] = np.nan
    t[np.fmod(t, 19) == 1] = np.nan
    return t, x, y, u, u1


def panel_data(drop7):
    # test-ci.R:430-436, 477-491
    X, Y = logistic_map(100)
    t = np.arange(1, 101, dtype=float)
    panel = (np.arange(1, 101) > 100 / 3) * 1.0
    x = X.copy()
    y = Y.copy()
    if drop7:
        keep = np.fmod(t, 7) != 0
        x, panel, t = x[keep], panel[keep], t[keep]
        y = y[keep]
    return t, x, y, panel


# --------------------------------------------------------------------------- scenarios


def no_seed_predict_with_past(edm):
    t = np.arange(1, 11, dtype=float)
    x = np.arange(1, 11, dtype=float)
    edm(t, x, shuffle=True)
    old = edm(t, x, full=True, predictWithPast=False, savePredictions=True)
    new = edm(t, x, full=True, predictWithPast=True, savePredictions=True)
    check_edm_result(old, .9722)
    check_edm_result(new, .99866)


def simple_explore_E(edm):
    t, x, y = simple_data()
    res = edm(t, x, E=list(range(2, 11)))
    check_edm_result(res, [.99893, .99879, .99835, .99763, .99457, .99385, .991, .98972, .98572])


def simple_xmap_k5(edm):
    t, x, y = simple_data()
    check_edm_result(edm(t, x, y, k=5), .55861)
    check_edm_result(edm(t, y, x, k=5), .94454)


def simple_xmap_lib8(edm):
    t, x, y = simple_data()
    check_edm_result(edm(t, x, y, E=6, library=8), .3362)
    check_edm_result(edm(t, y, x, E=6, library=8), .51116)


def simple_crossfold(edm):
    t, x, y = simple_data()
    res = edm(t, x, k=5, crossfold=10)
    expect_approx_equal(np.mean(res["stats"]["rho"]), .99946)


def simple_smap_theta(edm):
    t, x, y = simple_data()
    res = edm(t, x, theta=0.2 + np.arange(19) * 0.1, algorithm="smap")
    expect_approx_equal(res["stats"]["rho"][0], .99874)
    expect_approx_equal(res["stats"]["rho"][-1], .99882)


def simple_smap_xmap_coeffs(edm):
    t, x, y = simple_data()
    res1 = edm(t, x, y, theta=0.2, algorithm="smap", saveSMAPCoeffs=True)
    res2 = edm(t, y, x, theta=0.2, algorithm="smap", saveSMAPCoeffs=True)
    check_edm_result(res1, .66867)
    check_edm_result(res2, .98487)
    beta1 = res1["coeffs"]
    assert np.sum(np.isnan(beta1[0, :])) == beta1.shape[1]
    assert np.sum(np.isnan(beta1[1:, :])) == 0


def simple_predictions_copredict(edm):
    t, x, y = simple_data()
    res = edm(t, y, x, savePredictions=True)
    x2 = res["predictions"][:, 0]
    check_edm_result(res, .94272)
    assert np.isnan(x2[0]) and np.sum(np.isnan(x2[1:])) == 0

    res = edm(t, x, copredict=y, saveCoPredictions=True)
    teste = res["copredictions"][:, 0]
    check_edm_result(res, .9989, co_rho=.78002)
    assert np.isnan(teste[0]) and np.sum(np.isnan(teste[1:])) == 0

    zx = (x - np.mean(x)) / r_sd(x)
    check_edm_result(edm(t, zx, p=10), .90235)
    check_edm_result(edm(t, y, x, p=10), .89554)

    res = edm(t, y, x, p=10, copredict=x2, saveCoPredictions=True)
    testx = res["copredictions"][:, 0]
    check_edm_result(res, .89554, co_rho=.67401)
    assert np.sum(np.isnan(testx[:2])) == 2 and np.sum(np.isnan(testx[2:])) == 0

    zx2 = (x2 - np.nanmean(x2)) / r_sd(x2)
    res = edm(t, y, x, p=10, copredict=zx2, saveCoPredictions=True)
    testx2 = res["copredictions"][:, 0]
    check_edm_result(res, .89554, co_rho=.93837)
    assert np.sum(np.isnan(testx2[:2])) == 2 and np.sum(np.isnan(testx2[2:])) == 0

    u1 = np.asarray(golden_vectors()["u1_seed12345678"])
    res = edm(t, y, x, extras=[u1], p=10, copredict=zx2, saveCoPredictions=True)
    testx3 = res["copredictions"][:, 0]
    check_edm_result(res, .37011, co_rho=.9364)
    assert np.sum(np.isnan(testx3[:2])) == 2 and np.sum(np.isnan(testx3[2:])) == 0


def simple_consistency(edm):
    t, x, y = simple_data()
    resXmap = edm(t, tslag(t, x), x)
    check_edm_result(resXmap, .99939)
    resExplore = edm(t, x, full=True)
    check_edm_result(resExplore, .99939)
    expect_approx_equal(resXmap["stats"]["rho"], resExplore["stats"]["rho"])
    check_edm_result(edm(t, x, y), .54213)
    check_edm_result(edm(t, y, x), .94272)
    res = edm(t, x, E=[2, 3], theta=[0, 1])
    check_edm_result(res, [.99863, .99895, .99734, .99872])
    check_edm_result(edm(t, x, lowMemory=True), .9989)
    check_edm_result(edm(t, x, numThreads=4), .9989)


def missing_basic(edm):
    t, x, y, u = missing_data()
    check_edm_result(edm(t, x), .99814)
    check_edm_result(edm(t, x, dt=True, saveManifolds=True, dtWeight=1), .95569)
    check_edm_result(edm(t, x, allowMissing=True), .99766)
    check_edm_result(edm(t, x, allowMissing=True, missingDistance=1.0), .99765)


def missing_xmap_allow(edm):
    t, x, y, u = missing_data()
    check_edm_result(edm(t, x, tslag(t, x), allowMissing=True), .99983)
    check_edm_result(edm(t, tslag(t, x), x, allowMissing=True), .99864)


def missing_smap_dt_extras(edm):
    t, x, y, u = missing_data()
    res1 = edm(t, x, tslag(t, x), extras=[u], dt=True, algorithm="smap", saveSMAPCoeffs=True, E=5)
    res2 = edm(t, tslag(t, x), x, extras=[u], dt=True, algorithm="smap", saveSMAPCoeffs=True, E=5)
    check_edm_result(res1, 1.0)
    check_edm_result(res2, .77523)
    res = edm(t, x, tslag(t, x, 3), extras=[u], dt=True, algorithm="smap", saveSMAPCoeffs=True, E=5)
    check_edm_result(res, .36976)


def missing_crossfold_diff(edm):
    t, x, y, u = missing_data()
    res = edm(t, x, extras=[u], allowMissing=True, dt=True, crossfold=5)
    expect_approx_equal(np.mean(res["stats"]["rho"]), .92512)
    check_edm_result(edm(t, tsdiff(t, x), dt=True), .89192)


def missing_replicates(edm):
    t, x, y, u = missing_data()
    check_noisy_edm_result(edm(t, x, numReps=20), .99225, .9981)
    check_noisy_edm_result(edm(t, x, y, library=50, numReps=20), .35556, .40613)
    check_noisy_edm_result(edm(t, y, x, library=50, numReps=20), .82245, .85151)


def bigger_crossfold_extras(edm):
    t, x, y, _, _ = bigger_data(False)
    res = edm(t, x, E=2, crossfold=2, k=np.inf, allowMissing=True)
    expect_approx_equal(np.mean(res["stats"]["rho"]), .98175)
    res = edm(t, x, E=2, crossfold=10, k=np.inf, allowMissing=True)
    expect_approx_equal(np.mean(res["stats"]["rho"]), .98325)
    res = edm(t, x, E=5, extras=[tsdiff(t, y)], full=True, allowMissing=True)
    check_edm_result(res, .95266)


def bigger_library_sweep(edm):
    t, x, y, u, u1 = bigger_data(True)
    libs = list(range(10, 71, 5))
    res1 = edm(t, x, y, allowMissing=True, dt=True, library=libs)
    res2 = edm(t, y, x, allowMissing=True, dt=True, library=libs)
    check_edm_result(res1, [0.20492, 0.11316, 0.15244, 0.18469, 0.2577, 0.28964, 0.29208, 0.33099, 0.39233,
                            0.41628, 0.37522, 0.36816, 0.40495])
    check_edm_result(res2, [0.39118, 0.55506, 0.6788, 0.70348, 0.71176, 0.72476, 0.75539, 0.78565, 0.80807,
                            0.83358, 0.83503, 0.85401, 0.85847])


def bigger_negative_p_copredict(edm):
    t, x, y, u, u1 = bigger_data(True)
    check_edm_result(edm(t, x, p=-1), .99751)
    check_edm_result(edm(t, x, y, p=-1), .26842)
    check_edm_result(edm(t, y, x, p=-1), .8974)
    check_edm_result(edm(t, x, copredict=y), .99237, co_rho=.67756)
    check_edm_result(edm(t, x, copredict=y, full=True), .99416, co_rho=.77599)
    check_edm_result(edm(t, x, y, copredict=u1), .30789, co_rho=.42901)
    check_edm_result(edm(t, y, x, copredict=u1), .90401, co_rho=.5207)


def bigger_E_theta_grid(edm):
    t, x, y, u, u1 = bigger_data(True)
    res = edm(t, x, E=[2, 3, 4, 5], theta=[0, 1], copredict=y)
    rho = np.array([0.90482, 0.95631, 0.88553, 0.95751, 0.90482, 0.95652, np.nan, 0.95565])
    co_rho = np.array([0.47353, 0.51137, 0.41523, 0.50186, 0.27504, 0.42485, np.nan, 0.48008])
    assert res["rc"] == 0
    keep = np.arange(8) != 6
    assert np.max(np.abs(res["stats"]["rho"][keep] - rho[keep])) < 1e-4
    assert np.max(np.abs(res["copredStats"]["rho"][keep] - co_rho[keep])) < 1e-4


def bigger_libraries_copredict(edm):
    t, x, y, u, u1 = bigger_data(True)
    check_edm_result(edm(t, x, y, library=[5, 10, 20, 40], copredict=u1), [0.18385, 0.085223, 0.085659, 0.22313],
                     [0.18429, 0.26729, 0.37307, 0.36359])
    check_edm_result(edm(t, y, x, library=[5, 10, 20, 40], copredict=u1), [0.43651, 0.49275, 0.71521, 0.84646],
                     [0.63167, 0.63089, 0.50528, 0.3571])


def bigger_replicates(edm):
    t, x, y, u, u1 = bigger_data(True)
    check_noisy_edm_result(edm(t, x, copredict=y, numReps=20), .97335, .99339, .67584, .71214)
    res1 = edm(t, x, y, library=[5, 10, 20, 40], copredict=u1, numReps=100)
    res2 = edm(t, y, x, library=[5, 10, 20, 40], copredict=u1, numReps=100)
    check_noisy_edm_result(res1, np.array([-.24081, -.040511, .055321, .22091]), np.array([.74934, .63791, .53273,
                                                                                          .4299]),
                           np.array([-.40689, .1645, .15547, .045263]), np.array([1.1666, .93126, .86293, .70085]))
    check_noisy_edm_result(res2, np.array([-.15045, .25476, .53274, .72057]), np.array([.83649, .78668, .83097,
                                                                                       .90064]),
                           np.array([-.38777, .062654, .32211, .39681]), np.array([1.0946, .97872, .7964, .67552]))


def panel_E40(edm):
    t, x, y, panel = panel_data(False)
    check_edm_result(edm(t, x, panel=panel, E=40), .86964)
    check_edm_result(edm(t, x, panel=panel, E=40, allowMissing=True), .92115)
    check_edm_result(edm(t, x, panel=panel, E=40, panelWeight=np.inf), .86964)
    check_edm_result(edm(t, x, panel=panel, E=40, panelWeight=np.inf, allowMissing=True), .91768)


def panel_E40_xmap(edm):
    t, x, y, panel = panel_data(False)
    check_edm_result(edm(t, x, y, panel=panel, E=40), .76444)
    check_edm_result(edm(t, y, x, panel=panel, E=40), .83836)
    check_edm_result(edm(t, x, y, panel=panel, E=40, allowMissing=True), .63174)
    check_edm_result(edm(t, y, x, panel=panel, E=40, allowMissing=True), .81394)
    check_edm_result(edm(t, x, y, panel=panel, E=40, panelWeight=np.inf), .76444)
    check_edm_result(edm(t, y, x, panel=panel, E=40, panelWeight=np.inf), .83836)
    check_edm_result(edm(t, x, y, panel=panel, E=40, panelWeight=np.inf, allowMissing=True), .55937)
    check_edm_result(edm(t, y, x, panel=panel, E=40, panelWeight=np.inf, allowMissing=True), .75815)


def panel_missing_obs(edm):
    t, x, y, panel = panel_data(True)
    check_edm_result(edm(t, x, panel=panel, E=5), .95118)
    check_edm_result(edm(t, x, panel=panel, E=5, allowMissing=True), .95905)
    check_edm_result(edm(t, x, panel=panel, E=5, panelWeight=np.inf), .92472)
    check_edm_result(edm(t, x, panel=panel, E=5, panelWeight=np.inf, allowMissing=True), .93052)
    check_edm_result(edm(t, x, panel=panel, E=5, panelWeight=np.inf, k=np.inf), .92472)
    check_edm_result(edm(t, x, panel=panel, E=5, reldt=True), .90239)
    check_edm_result(edm(t, x, panel=panel, E=5, reldt=True, allowMissing=True), .9085)
    check_edm_result(edm(t, x, panel=panel, E=5, panelWeight=np.inf, reldt=True), .78473)
    check_edm_result(edm(t, x, panel=panel, E=5, panelWeight=np.inf, reldt=True, allowMissing=True), .75709)


def bad_inputs(edm):
    t, x, y = simple_data()
    assert edm(t, x, y, E=np.nan)["rc"] == 0
    assert edm(t, x, y, E=[2, 3, np.nan])["rc"] == 0
    assert edm(t, x, y, k=np.nan)["rc"] == 0


ALL = [no_seed_predict_with_past, simple_explore_E, simple_xmap_k5, simple_xmap_lib8, simple_crossfold,
       simple_smap_theta, simple_smap_xmap_coeffs, simple_predictions_copredict, simple_consistency, missing_basic,
       missing_xmap_allow, missing_smap_dt_extras, missing_crossfold_diff, missing_replicates,
       bigger_crossfold_extras, bigger_library_sweep, bigger_negative_p_copredict, bigger_E_theta_grid,
       bigger_libraries_copredict, bigger_replicates, panel_E40, panel_E40_xmap, panel_missing_obs, bad_inputs]
