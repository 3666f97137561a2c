"""Seeded generators of the benchmark configurations' inputs (synthetic data of the stated shapes).

cfg1  logistic map x' = 3.79 x (1 - x), x0 = 0.2, T = 500, plus N(0, sd 0.01) noise (seed 1)
cfg2  coupled logistic maps x' = x(3.8 - 3.8x - 0.02y), y' = y(3.5 - 3.5y - 0.1x), x0 = 0.4, y0 = 0.2, T = 5000
cfg3  the same maps at T = 20000 (bench.py)
cfg4  512 logistic-map panels, lengths U{200..2000}, r ~ U[3.6, 3.95], x0 ~ U(0.1, 0.9), 5 % NaN (seed 4)
cfg5  Lorenz-96, N = 40, F = 8, RK4 dt = 0.01, x = F + N(0,1) (seed 5), 1000 transient steps dropped
"""
import numpy as np


def cfg1_series():
    rng = np.random.default_rng(1)
    x = np.empty(500)
    x[0] = 0.2
    for i in range(499):
        x[i + 1] = 3.79 * x[i] * (1 - x[i])
    return np.arange(1, 501, dtype=float), x + rng.normal(0.0, 0.01, 500)


def coupled_maps(T):
    x = np.empty(T)
    y = np.empty(T)
    x[0], y[0] = 0.4, 0.2
    for i in range(T - 1):
        x[i + 1] = x[i] * (3.8 - 3.8 * x[i] - 0.02 * y[i])
        y[i + 1] = y[i] * (3.5 - 3.5 * y[i] - 0.1 * x[i])
    return x, y


def ccm_libraries():
    """40 log-spaced integer library sizes from 10 to 19000 (they sum to 107 910)."""
    return np.round(np.exp(np.linspace(np.log(10), np.log(19000), 40))).astype(int)


def cfg4_panels(num_panels=512, seed=4):
    rng = np.random.default_rng(seed)
    ts, xs, ps = [], [], []
    for p in range(num_panels):
        n = int(rng.integers(200, 2001))
        r = rng.uniform(3.6, 3.95)
        v = np.empty(n)
        v[0] = rng.uniform(0.1, 0.9)
        for i in range(n - 1):
            v[i + 1] = r * v[i] * (1 - v[i])
        ts.append(np.arange(1, n + 1, dtype=float))
        xs.append(v)
        ps.append(np.full(n, p))
    t = np.concatenate(ts)
    x = np.concatenate(xs)
    panel = np.concatenate(ps)
    x[rng.random(x.size) < 0.05] = np.nan
    return t, x, panel


def lorenz96(steps=100000, N=40, F=8.0, dt=0.01, transient=1000, seed=5):
    rng = np.random.default_rng(seed)
    s = F + rng.standard_normal(N)

    ip1 = (np.arange(N) + 1) % N
    im1 = (np.arange(N) - 1) % N
    im2 = (np.arange(N) - 2) % N

    def f(v):
        return (v[ip1] - v[im2]) * v[im1] - v + F

    out = np.empty((steps, N))
    for i in range(transient + steps):
        k1 = f(s)
        k2 = f(s + 0.5 * dt * k1)
        k3 = f(s + 0.5 * dt * k2)
        k4 = f(s + dt * k3)
        s = s + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        if i >= transient:
            out[i - transient] = s
    return out


def cfg5_manifold(series, E=10, extras=30):
    """Points [x0 lags 0..E-1 | x1..x_extras unlagged], target x0 one step ahead; 80/20 rows."""
    x0 = series[:, 0]
    n = series.shape[0]
    rows = np.arange(E - 1, n - 1)
    pts = np.empty((rows.size, E + extras))
    for j in range(E):
        pts[:, j] = x0[rows - j]
    pts[:, E:] = series[rows][:, 1:1 + extras]
    targets = x0[rows + 1]
    cut = int(0.8 * rows.size)
    lib = {"data": pts[:cut], "targets": targets[:cut], "E_x": E, "E_extras": extras}
    pred = {"data": pts[cut:], "targets": targets[cut:], "E_x": E, "E_extras": extras}
    return lib, pred
