"""HIP path (edm_predict through the C ABI) against the CPU oracle on seeded tasks.

Bar: neighbour indices, k_used and retcodes bit-exact; predictions and S-map coefficients within
|d| <= 1e-12 + 1e-10 |ref| (FP64 plain form)."""
import zlib

import numpy as np
import pytest

from oracle import oracle
from tests.conftest import load_product
from tests.tasks import assert_predictions_close, random_task, wasserstein_task

pytestmark = pytest.mark.gpu

CASES = {
    "simplex_E3_k4": dict(NL=500, NP=300, Ea=3, k=4),
    "simplex_duplicates": dict(NL=300, NP=200, Ea=2, k=3, duplicate_targets=True),
    "simplex_integer_ties": dict(NL=400, NP=150, Ea=2, k=5, integer=True),
    "simplex_all": dict(NL=300, NP=120, Ea=3, k=-1, thetas=(0.0, 1.0, 3.0)),
    "simplex_k_exceeds_valid": dict(NL=6, NP=20, Ea=2, k=8),
    "simplex_slices_merge": dict(NL=6000, NP=1500, Ea=4, k=7, thetas=(0.5, 1.0)),
    "smap_all_thetas": dict(NL=600, NP=300, Ea=4, k=-1, algorithm=1, thetas=(0, .5, 1, 2, 4, 8), save_coeffs=True),
    "smap_k6_coeffs": dict(NL=400, NP=200, Ea=5, k=6, algorithm=1, thetas=(0.2,), save_coeffs=True),
    "missing_with_distance": dict(NL=300, NP=150, Ea=4, k=5, missing=0.05, missing_distance=0.3),
    "missing_dropped": dict(NL=300, NP=150, Ea=4, k=5, missing=0.05),
    "panels_skip_other": dict(NL=300, NP=150, Ea=3, k=4, panels=3, idw=-1.0),
    "panels_idw": dict(NL=300, NP=150, Ea=3, k=4, panels=3, idw=0.7),
    "panels_id_weights": dict(NL=300, NP=150, Ea=3, k=4, panels=4, id_weights=True),
    "only_past": dict(NL=300, NP=150, Ea=3, k=4, past=True),
    "low_memory_ignores_past": dict(NL=300, NP=150, Ea=3, k=4, past=True, low_memory=True),
    "categorical": dict(NL=300, NP=150, Ea=3, k=4, categorical=True, integer=True),
    "mae_distance": dict(NL=300, NP=150, Ea=3, k=4, mae=True),
    "smap_E16_k18": dict(NL=2000, NP=300, Ea=16, k=18, algorithm=1, thetas=(0, 1, 3)),
    # the register-row S-map kernels by width: n = 8 (8 + 0), 13 (16 + 0), 20 (16 + 4), 23 (16 + 8)
    "smap_E7_k9": dict(NL=800, NP=200, Ea=7, k=9, algorithm=1, thetas=(0, 1.5), save_coeffs=True),
    "smap_E12_k20": dict(NL=800, NP=200, Ea=12, k=20, algorithm=1, thetas=(0.5,), save_coeffs=True),
    "smap_E19_k21": dict(NL=1000, NP=150, Ea=19, k=21, algorithm=1, thetas=(0, 2.0), save_coeffs=True),
    "smap_E22_k30": dict(NL=1000, NP=150, Ea=22, k=30, algorithm=1, thetas=(1.0,), save_coeffs=True),
    # the LDS-row S-map kernel: two systems per wavefront (n = 29), one (n = 34)
    "smap_E28_k31": dict(NL=1200, NP=130, Ea=28, k=31, algorithm=1, thetas=(0, 2.0), save_coeffs=True),
    "smap_E33_k36": dict(NL=1200, NP=100, Ea=33, k=36, algorithm=1, thetas=(0.5,), save_coeffs=True),
    "smap_E40_k42": dict(NL=1500, NP=120, Ea=40, k=42, algorithm=1, thetas=(2.0,), save_coeffs=True),
    "simplex_E40_k41": dict(NL=1500, NP=200, Ea=40, k=41),
    "invalid_algorithm": dict(NL=100, NP=50, Ea=2, k=3, algorithm=7),
    # k wider than the register lists: radix select + sort on the dense rows
    "simplex_wide_k100": dict(NL=3000, NP=200, Ea=3, k=100, thetas=(0.5, 2.0)),
    "simplex_wide_ties_k90": dict(NL=2000, NP=150, Ea=2, k=90, integer=True),
    "simplex_wide_duplicates_k80": dict(NL=2000, NP=150, Ea=3, k=80, duplicate_targets=True),
    "simplex_wide_global_sort_k5000": dict(NL=9000, NP=60, Ea=3, k=5000),
    "smap_wide_k150": dict(NL=2000, NP=150, Ea=5, k=150, algorithm=1, thetas=(0, 1, 3), save_coeffs=True),
    "wide_missing_panels_k120": dict(NL=2000, NP=100, Ea=4, k=120, missing=0.05, missing_distance=0.3, panels=3,
                                     idw=0.5),
    "wide_invalid_algorithm": dict(NL=300, NP=50, Ea=2, k=70, algorithm=7),
    # points wider than 64 coordinates: the row-wise distance kernel, exact selection on the dense rows,
    # and the block-wide S-map solve for 65..128 unknowns
    "simplex_E80_k4": dict(NL=600, NP=60, Ea=80, k=4, thetas=(1.0, 3.0)),
    "simplex_E80_all": dict(NL=300, NP=40, Ea=80, k=-1, thetas=(0.5, 2.0)),
    "simplex_E300_k5": dict(NL=300, NP=30, Ea=300, k=5),
    "wide_options_E90": dict(NL=500, NP=40, Ea=90, k=6, missing=0.05, missing_distance=0.3, panels=3, idw=0.5),
    "smap_E64_k70": dict(NL=400, NP=30, Ea=64, k=70, algorithm=1, thetas=(0.5,), save_coeffs=True),
    "smap_E80_k82": dict(NL=500, NP=30, Ea=80, k=82, algorithm=1, thetas=(0, 2.0), save_coeffs=True),
    "smap_E80_k82_mixed": dict(NL=500, NP=30, Ea=80, k=82, algorithm=1, thetas=(1.0,), precision=1),
    "smap_E80_all": dict(NL=200, NP=20, Ea=80, k=-1, algorithm=1, thetas=(1.0,), save_coeffs=True),
    "smap_E127_k130": dict(NL=400, NP=12, Ea=127, k=130, algorithm=1, thetas=(1.0,), save_coeffs=True),
    # all-neighbour S-map systems formed ahead of the solve (one size class of entries per lane each)
    "smap_all_E1_8thetas": dict(NL=700, NP=90, Ea=1, k=-1, algorithm=1, thetas=(0, .1, .5, 1, 2, 3, 4, 8),
                                save_coeffs=True),
    "smap_all_E8_missing_panels": dict(NL=1000, NP=150, Ea=8, k=-1, algorithm=1, thetas=(0, 1, 3), missing=0.05,
                                       missing_distance=0.3, panels=3, save_coeffs=True),
    "smap_all_E12": dict(NL=500, NP=60, Ea=12, k=-1, algorithm=1, thetas=(0.5, 2.0), save_coeffs=True),
    "smap_all_E20": dict(NL=400, NP=40, Ea=20, k=-1, algorithm=1, thetas=(1.0, 3.0), save_coeffs=True),
    "smap_all_E36": dict(NL=300, NP=30, Ea=36, k=-1, algorithm=1, thetas=(1.0,), save_coeffs=True),
    "smap_all_E50": dict(NL=300, NP=20, Ea=50, k=-1, algorithm=1, thetas=(0.5,), save_coeffs=True),
    "smap_all_E63": dict(NL=260, NP=20, Ea=63, k=-1, algorithm=1, thetas=(1.0,), save_coeffs=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_task_parity(name):
    fe = load_product()
    spec = CASES[name]
    M, Mp, opts = random_task(zlib.crc32(name.encode()) % 1000, **spec)
    cap = max(64, spec["k"])  # the whole neighbour list
    print(f"case {name}: NL={spec['NL']} NP={spec['NP']} Ea={spec['Ea']} k={spec['k']}", flush=True)
    rp, rr, rc_, rk, rn = oracle.predict(M, Mp, opts, nn_cap=cap)
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=cap)
    assert status == 0
    np.testing.assert_array_equal(gr, rr)
    np.testing.assert_array_equal(gk, rk)
    np.testing.assert_array_equal(gn, rn)
    assert_predictions_close(gp, rp)
    if opts.get("save_smap_coeffs"):
        assert_predictions_close(gc, rc_)


@pytest.mark.parametrize("Ea", [1, 3, 7, 8, 12, 15, 16, 17, 19, 22, 23, 32, 40, 47])
def test_smap_register_rows_match_lds_rows(Ea, monkeypatch):
    """For n = E_a + 1 <= 24 the S-map solve keeps U and V rows in registers: FEDM_SMAP_BLOCK=0, one row
    per lane; 1, block-wide with each rotation's 2x2 step made once per system; 2 and 3, two rows per lane
    with the 2x2 steps made once per system (fewer or more systems per wavefront); 4, the default choice.
    For 33 <= n <= 48 U's rows are in registers; FEDM_SMAP_LDS_ROWS=1 keeps them in LDS.  All run the same
    floating-point operations in the same order, so every prediction and coefficient must agree to the
    bit (k near n gives rank-deficient and ill-conditioned systems)."""
    fe = load_product()
    for k in (Ea + 1, Ea + 3):
        M, Mp, opts = random_task(100 + 7 * Ea + k, NL=700, NP=150, Ea=Ea, k=k, algorithm=1, thetas=(0, 1, 4),
                                  save_coeffs=True, integer=(k == Ea + 1))
        print(f"smap rows: Ea={Ea} k={k}", flush=True)
        monkeypatch.delenv("FEDM_SMAP_LDS_ROWS", raising=False)
        monkeypatch.delenv("FEDM_SMAP_BLOCK", raising=False)
        gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=64)
        runs = []
        for var, val in (("FEDM_SMAP_LDS_ROWS", "1"), ("FEDM_SMAP_BLOCK", "0"), ("FEDM_SMAP_BLOCK", "1"),
                         ("FEDM_SMAP_BLOCK", "2"), ("FEDM_SMAP_BLOCK", "3")):
            monkeypatch.setenv(var, val)
            runs.append(fe.predict(M, Mp, opts, nn_cap=64))
            monkeypatch.delenv(var)
        assert status == 0
        for lp, lr, lc, lk, ln, lstatus in runs:
            assert lstatus == 0
            np.testing.assert_array_equal(gr, lr)
            np.testing.assert_array_equal(np.ascontiguousarray(gp).view(np.uint64),
                                          np.ascontiguousarray(lp).view(np.uint64))
            np.testing.assert_array_equal(np.ascontiguousarray(gc).view(np.uint64),
                                          np.ascontiguousarray(lc).view(np.uint64))


def test_smap_width_limit():
    """The S-map solve takes at most 128 unknowns: E_a = 128 is refused with -95 (and rc 8000), while the
    same task with simplex runs."""
    fe = load_product()
    M, Mp, opts = random_task(7, NL=300, NP=20, Ea=128, k=131, algorithm=1)
    print("smap width limit: Ea=128", flush=True)
    with pytest.raises(fe.FastEDMError, match=r"\(-95\)"):
        fe.predict(M, Mp, opts, nn_cap=0)
    opts["algorithm"] = 0
    opts["k"] = 5
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=0)
    assert status == 0


WASS_CASES = {
    "wass_E5_k4": dict(NL=400, NP=100, E=5, k=4),
    "wass_dt_aspect": dict(NL=300, NP=80, E=6, k=5, dt=True, aspect_ratio=0.7, thetas=(0.5, 2.0)),
    "wass_skip_missing": dict(NL=300, NP=80, E=6, k=4, missing=0.15),
    "wass_missing_distance": dict(NL=300, NP=80, E=6, k=4, missing=0.15, missing_distance=0.4),
    "wass_extras_categorical": dict(NL=300, NP=80, E=4, k=6, lagged=1, unlagged=2, categorical=True),
    "wass_panels_idw": dict(NL=300, NP=80, E=5, k=4, panels=3, idw=0.3),
    "wass_panels_skip_other": dict(NL=300, NP=80, E=5, k=4, panels=3, idw=-1.0),
    "wass_panels_id_weights": dict(NL=300, NP=80, E=5, k=4, panels=3, id_weights=True),
    "wass_smap_all": dict(NL=200, NP=60, E=4, k=-1, algorithm=1, thetas=(0.0, 1.0, 2.0), save_coeffs=True),
    "wass_smap_k8": dict(NL=300, NP=60, E=5, k=8, algorithm=1, thetas=(1.0,), dt=True),
    "wass_simplex_all": dict(NL=200, NP=60, E=4, k=-1, thetas=(0.0, 1.0)),
    "wass_wide_k80": dict(NL=400, NP=40, E=5, k=80),
    "wass_E16": dict(NL=200, NP=40, E=16, k=5),
    "wass_E32": dict(NL=60, NP=10, E=32, k=3),
    # two transport nodes per lane (series of 33..64 observations)
    "wass_E33": dict(NL=40, NP=8, E=33, k=3),
    "wass_E48_dt_missing": dict(NL=40, NP=8, E=48, k=3, dt=True, missing=0.1, thetas=(0.5, 2.0)),
    "wass_E64_missing_distance": dict(NL=30, NP=6, E=64, k=3, missing=0.1, missing_distance=0.4, unlagged=1),
}


@pytest.mark.parametrize("name", sorted(WASS_CASES))
def test_wasserstein_parity(name):
    """Wasserstein distances (exact transport per pair on the GPU) against the oracle's exact solver:
    neighbour sets bit-exact, predictions within the plain bar."""
    fe = load_product()
    spec = WASS_CASES[name]
    M, Mp, opts = wasserstein_task(zlib.crc32(name.encode()) % 1000, **spec)
    cap = max(64, spec["k"])
    print(f"case {name}: NL={spec['NL']} NP={spec['NP']} E={spec['E']} k={spec['k']}", flush=True)
    rp, rr, rc_, rk, rn = oracle.predict(M, Mp, opts, nn_cap=cap)
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=cap)
    assert status == 0
    np.testing.assert_array_equal(gr, rr)
    np.testing.assert_array_equal(gk, rk)
    np.testing.assert_array_equal(gn, rn)
    assert_predictions_close(gp, rp)
    if opts.get("save_smap_coeffs"):
        assert_predictions_close(gc, rc_)


def test_keep_going_interrupts():
    fe = load_product()
    M, Mp, opts = random_task(5, NL=200, NP=100, Ea=3, k=4)
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, keep_going=lambda: 0)
    assert status == 1
    assert (gr[0] == 1).all()
    assert (gp == 1.0e100).all()


def test_mfma_lane_layout():
    """The FP32 matrix-core tile layout the mixed route relies on."""
    fe = load_product()
    out = np.zeros(256, np.float32)
    assert fe.load_library().edm_selftest_mfma(out.ctypes.data_as(fe._fp)) == 0
    for lane in range(64):
        g, j = lane // 16, lane % 16
        for r in range(4):
            i = 4 * g + r
            assert out[lane * 4 + r] == 100.0 * (i + 1) * (j + 1), (lane, r, out[lane * 4 + r])


MIXED_CASES = {
    "mixed_smap_E16_k18": dict(NL=3000, NP=400, Ea=16, k=18, algorithm=1, thetas=(0, 1, 3)),
    "mixed_smap_E40_k42": dict(NL=2000, NP=150, Ea=40, k=42, algorithm=1, thetas=(2.0,), save_coeffs=True),
    "mixed_simplex_E8_k9": dict(NL=5000, NP=700, Ea=8, k=9, thetas=(0.5, 2.0)),
    "mixed_integer_ties_E8": dict(NL=2000, NP=300, Ea=8, k=6, integer=True),
    "mixed_duplicates_E12": dict(NL=1500, NP=300, Ea=12, k=5, duplicate_targets=True),
    "mixed_k_exceeds_valid_E9": dict(NL=7, NP=40, Ea=9, k=10),
}


def logistic_task(seed, NL, NP, E, k, algorithm=0, thetas=(1.0,), save_coeffs=False, panels=1):
    """Lagged embeddings of logistic-map series (r = 3.9, a thin attractor) split at random into
    library and targets."""
    rng = np.random.default_rng(seed)
    n = NL + NP + E * panels
    rows = []
    for p in range(panels):
        x = np.empty(n // panels + E)
        x[0] = rng.uniform(0.1, 0.9)
        for t in range(1, len(x)):
            x[t] = 3.9 * x[t - 1] * (1.0 - x[t - 1])
        rows.append(np.stack([x[i:len(x) - E + 1 + i] for i in range(E)], axis=1))
    X = np.concatenate(rows)[: NL + NP]
    perm = rng.permutation(NL + NP)
    M, Mp = X[perm[:NL]].copy(), X[perm[NL:]].copy()
    Md, Mpd, opts = random_task(seed, NL=NL, NP=NP, Ea=E, k=k, algorithm=algorithm, thetas=thetas,
                                save_coeffs=save_coeffs, precision=1)
    Md["data"], Mpd["data"] = M, Mp
    return Md, Mpd, opts


CLUSTERED_CASES = {
    "logistic_simplex_E8_k9": dict(NL=20000, NP=3000, E=8, k=9, thetas=(0.5, 2.0)),
    "logistic_smap_E16_k18": dict(NL=12000, NP=2500, E=16, k=18, algorithm=1, thetas=(0, 1, 3), save_coeffs=True),
    "logistic_panels_E12_k13": dict(NL=15000, NP=2000, E=12, k=13, panels=7),
}


@pytest.mark.parametrize("name", sorted(CLUSTERED_CASES))
def test_mixed_clustered_parity(name):
    """Mixed route on a thin attractor (many near neighbours inside the FP32 screening bound) against
    the FP64 oracle: indices exact."""
    fe = load_product()
    M, Mp, opts = logistic_task(zlib.crc32(name.encode()) % 1000, **CLUSTERED_CASES[name])
    print(f"case {name}", flush=True)
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=64)
    rp, rr, rc_, rk, rn = oracle.predict(M, Mp, opts, nn_cap=64)
    assert status == 0
    np.testing.assert_array_equal(gr, rr)
    np.testing.assert_array_equal(gk, rk)
    np.testing.assert_array_equal(gn, rn)
    assert_predictions_close(gp, rp)
    if opts.get("save_smap_coeffs"):
        assert_predictions_close(gc, rc_)


@pytest.mark.parametrize("name", sorted(MIXED_CASES))
def test_mixed_task_parity(name):
    """Mixed route (FP32 matrix-core screening + FP64 re-rank) against the FP64 oracle: indices exact."""
    fe = load_product()
    spec = MIXED_CASES[name]
    M, Mp, opts = random_task(zlib.crc32(name.encode()) % 1000, precision=1, **spec)
    print(f"case {name}", flush=True)
    rp, rr, rc_, rk, rn = oracle.predict(M, Mp, opts, nn_cap=64)
    gp, gr, gc, gk, gn, status = fe.predict(M, Mp, opts, nn_cap=64)
    assert status == 0
    np.testing.assert_array_equal(gr, rr)
    np.testing.assert_array_equal(gk, rk)
    np.testing.assert_array_equal(gn, rn)
    # the mixed form changes only the screening; the simplex weights and the S-map solve run in FP64 in
    # the reference's order, so the plain bar holds
    d = assert_predictions_close(gp, rp)
    print(f"max |d yhat| {d:.3e}")
    if opts.get("save_smap_coeffs"):
        assert_predictions_close(gc, rc_)


def test_device_exp_matches_oracle_bitwise():
    """The device exp of the simplex and S-map weights (FDLIBM's method, fdlibm_exp.h) against the
    oracle's over the dense grid of the weights' domain that pins the oracle's to within 1 ulp of libm:
    identical bits everywhere, so the two copies of the method cannot drift apart."""
    import ctypes as C
    from tests.test_oracle_math import _exp_grid
    fe = load_product()
    L = fe.load_library()
    L.edm_selftest_exp.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
    x = np.ascontiguousarray(_exp_grid())
    out = np.zeros_like(x)
    assert L.edm_selftest_exp(x.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)),
                              x.size) == 0
    O = oracle.lib()
    O.orc_weight_exp.restype = C.c_double
    O.orc_weight_exp.argtypes = [C.c_double]
    ref = np.array([O.orc_weight_exp(float(v)) for v in x])
    assert np.array_equal(out.view(np.int64), ref.view(np.int64))


@pytest.mark.parametrize("case", ["clustered_dt", "ties", "logistic"])
def test_mfma_seed_leaves_results_unchanged(case, monkeypatch):
    """The matrix-core route starts each target's FP32 screening threshold from an exact k-th distance
    over library rows near it in a k-d order (FEDM_MFMA_SEED=1: its 256-row window here; 4 096 rows around
    it, the k-th of each lane's four smallest, above 16 coordinates), or screens every pair against the k-th over its 256-row window, re-ranks
    the recorded candidates once and hands targets with more than 256 candidates, after a second seed over
    1 024 rows, to the streaming kernel (FEDM_MFMA_SEED=2; 3, the default, chooses); the screen of the
    two-phase form skips the library chunks whose box lies beyond its targets' thresholds unless
    FEDM_MFMA_PRUNE=0, and shares each target block's chunks out over FEDM_MFMA_SPLIT blocks (default:
    chosen from the target count).  Each only removes candidates that cannot enter the list, so indices,
    k_used, codes and predictions must equal the unseeded run's (FEDM_MFMA_SEED=0) bit for bit and the
    oracle's.  Cases: a dt-like layout (a column equal to a constant on most rows), an integer table with
    many tied and zero distances (targets whose candidates overflow), and the lag embedding of a chaotic
    logistic map with dt columns (a low-dimensional manifold, where most chunks are skipped; targets not a
    multiple of the 128-target blocks, library not a multiple of the 64-row chunks)."""
    fe = load_product()
    rng = np.random.default_rng(11 if case == "ties" else 12)
    NL, NP, Ea, k = 3000, 400, 12, 9
    if case == "ties":
        data = rng.integers(0, 4, size=(NL + NP, Ea)).astype(float)
    elif case == "logistic":
        NL, NP, Ea, k = 20000, 3001, 16, 18
        x = np.empty(NL + NP + 8)
        x[0] = 0.4
        for t in range(x.size - 1):
            x[t + 1] = 3.9 * x[t] * (1.0 - x[t])
        lags = np.stack([x[7 - j:x.size - 1 - j] for j in range(8)], axis=1)
        dt = rng.choice([1.0, 1.25, 2.5], size=(NL + NP, 8), p=[0.9, 0.05, 0.05])
        data = np.hstack([lags, dt])
    else:
        base = rng.random((NL + NP, Ea))
        base[:, 6:] = 1.25
        gaps = rng.random((NL + NP, 6)) < 0.05
        base[:, 6:][gaps] = 2.5
        data = base
    M = {"data": data[:NL].copy(), "targets": rng.random(NL)}
    Mp = {"data": data[NL:].copy()}
    opts = {"algorithm": 0, "k": k, "thetas": [0.0, 1.0], "precision": fe.MIXED}
    print(f"mfma seed: {case}", flush=True)
    outs = []
    for seed, prune, split in (("0", "1", "0"), ("1", "1", "0"), ("2", "1", "0"), ("3", "1", "0"), ("2", "0", "0"),
                               ("2", "1", "1"), ("2", "1", "5")):
        monkeypatch.setenv("FEDM_MFMA_SEED", seed)
        monkeypatch.setenv("FEDM_MFMA_PRUNE", prune)
        monkeypatch.setenv("FEDM_MFMA_SPLIT", split)
        outs.append(fe.predict(M, Mp, opts, nn_cap=k))
    for v in ("FEDM_MFMA_SEED", "FEDM_MFMA_PRUNE", "FEDM_MFMA_SPLIT"):
        monkeypatch.delenv(v)
    p0, r0, c0, k0, n0, s0 = outs[0]
    assert s0 == 0
    for p1, r1, c1, k1, n1, s1 in outs[1:]:
        assert s1 == 0
        np.testing.assert_array_equal(n0, n1)
        np.testing.assert_array_equal(k0, k1)
        np.testing.assert_array_equal(r0, r1)
        np.testing.assert_array_equal(np.ascontiguousarray(p0).view(np.uint64),
                                      np.ascontiguousarray(p1).view(np.uint64))
    rp, rr, rc_, rk, rn = oracle.predict(M, Mp, dict(opts, precision=0), nn_cap=k)
    np.testing.assert_array_equal(n1, rn)
    np.testing.assert_array_equal(k1, rk)
    assert_predictions_close(p1, rp)
