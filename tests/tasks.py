"""Seeded task-level inputs shared by the parity tests (oracle vs HIP path)."""
import numpy as np

MISSING_D = 1.0e100


def random_task(seed, NL, NP, Ea, k, algorithm=0, thetas=(1.0,), integer=False, missing=0.0, missing_distance=0.0,
                panels=0, idw=0.0, id_weights=False, past=False, low_memory=False, categorical=False, mae=False,
                duplicate_targets=False, save_coeffs=False, precision=0):
    rng = np.random.default_rng(seed)
    if integer:
        M = rng.integers(0, 4, size=(NL, Ea)).astype(np.float64)
        Mp = rng.integers(0, 4, size=(NP, Ea)).astype(np.float64)
    else:
        M = rng.random((NL, Ea))
        Mp = rng.random((NP, Ea))
    if duplicate_targets:
        # some prediction points coincide with library points: zero distances must be dropped
        idx = rng.integers(0, NL, size=NP // 3)
        Mp[: NP // 3] = M[idx]
    if missing > 0:
        M[rng.random(M.shape) < missing] = MISSING_D
        Mp[rng.random(Mp.shape) < missing] = MISSING_D
    tM = rng.random(NL) * 2 - 0.5
    lib_pt = np.arange(NL, dtype=np.float64)
    lib_tt = lib_pt + 1 + rng.integers(0, 3, NL)
    pred_pt = rng.integers(0, NL + 2, NP).astype(np.float64)
    Md = {"data": M, "targets": tM, "point_times": lib_pt, "target_times": lib_tt}
    Mpd = {"data": Mp, "targets": rng.random(NP), "point_times": pred_pt, "target_times": pred_pt + 1}
    opts = {"algorithm": algorithm, "distance": 0 if mae else 1, "k": k, "thetas": np.asarray(thetas, float),
            "missing_distance": missing_distance, "has_missing": int(missing > 0), "use_only_past": int(past),
            "low_memory": int(low_memory), "save_smap_coeffs": int(save_coeffs), "force_compute": 1,
            "precision": precision}
    metrics = np.zeros(Ea, np.int32)
    if categorical:
        metrics[-1] = 1
        opts["has_categorical"] = 1
    opts["metrics"] = metrics
    if panels:
        Md["panel"] = rng.integers(0, panels, NL).astype(np.int32)
        Mpd["panel"] = rng.integers(0, panels, NP).astype(np.int32)
        opts["panel_mode"] = 1
        opts["idw"] = idw
        if id_weights:
            opts["id_weights"] = rng.random((panels, panels)).astype(np.float32)
            opts["idw"] = 0.0
    return Md, Mpd, opts


def assert_predictions_close(gpu, ref, rtol=1e-10, atol=1e-12):
    """Same MISSING pattern; elsewhere |gpu - ref| <= atol + rtol * |ref| (the plain form's bar)."""
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    gm = gpu == MISSING_D
    rm = ref == MISSING_D
    assert np.array_equal(gm, rm), f"missing pattern differs at {np.argwhere(gm != rm)[:5]}"
    d = np.abs(gpu[~gm] - ref[~rm])
    tol = atol + rtol * np.abs(ref[~rm])
    bad = d > tol
    assert not bad.any(), f"{bad.sum()} values differ, max |d| {d.max():.3e}"
    return float(d.max()) if d.size else 0.0


def wasserstein_task(seed, NL, NP, E, k, dt=False, lagged=0, unlagged=0, missing=0.0, missing_distance=0.0,
                     categorical=False, panels=0, idw=0.0, id_weights=False, algorithm=0, thetas=(1.0,),
                     aspect_ratio=1.0, save_coeffs=False):
    """A task with the Wasserstein point layout [x lags | dt lags | lagged extras | unlagged extras]
    (src/distances.cpp:318-326) and its per-variable metrics (src/distances.cpp:37-47)."""
    rng = np.random.default_rng(seed)
    rows = 1 + int(dt) + lagged
    Ea = rows * E + unlagged

    def points(n):
        P = np.empty((n, Ea))
        P[:, :E] = rng.random((n, E))
        r = 1
        if dt:
            P[:, E:2 * E] = rng.integers(1, 4, (n, E)).astype(np.float64)
            r = 2
        for _ in range(lagged):
            P[:, r * E:(r + 1) * E] = (rng.integers(0, 3, (n, E)).astype(np.float64) if categorical
                                       else rng.random((n, E)))
            r += 1
        P[:, rows * E:] = (rng.integers(0, 2, (n, unlagged)).astype(np.float64) if categorical
                           else rng.random((n, unlagged)))
        if missing > 0:
            P[rng.random(P.shape) < missing] = MISSING_D
        return P

    M, Mp = points(NL), points(NP)
    dims = {"E_x": E, "E_dt": E if dt else 0, "E_extras": lagged * E + unlagged, "E_lagged_extras": lagged * E}
    Md = dict(data=M, targets=rng.random(NL) * 2 - 0.5, **dims)
    Mpd = dict(data=Mp, targets=rng.random(NP), **dims)
    metrics = np.zeros(rows + unlagged, np.int32)
    if categorical:
        metrics[1 + int(dt):] = 1
    opts = {"algorithm": algorithm, "distance": 2, "k": k, "thetas": np.asarray(thetas, float),
            "missing_distance": missing_distance, "has_missing": int(missing > 0), "metrics": metrics,
            "has_categorical": int(categorical), "aspect_ratio": aspect_ratio, "save_smap_coeffs": int(save_coeffs),
            "force_compute": 1}
    if panels:
        Md["panel"] = rng.integers(0, panels, NL).astype(np.int32)
        Mpd["panel"] = rng.integers(0, panels, NP).astype(np.int32)
        opts["panel_mode"] = 1
        opts["idw"] = idw
        if id_weights:
            opts["id_weights"] = rng.random((panels, panels)).astype(np.float32)
            opts["idw"] = 0.0
    return Md, Mpd, opts
