"""Seeded synthetic series for the benchmark configurations and parity tests.

All generators iterate in float64 and return the requested dtype; z-normalised
series have mean 0 and standard deviation 1 (population std, ddof=0).
"""
import numpy as np


def znorm(x):
    x = np.asarray(x, dtype=np.float64)
    return (x - x.mean()) / x.std()


def logistic(n, seed=0, r=4.0, dtype=np.float32):
    """x_{k+1} = r x_k (1 - x_k), x_0 ~ U(0.1, 0.9) from the seed."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(0.1, 0.9)
    out = np.empty(n)
    for k in range(n):
        out[k] = v
        v = r * v * (1.0 - v)
    return out.astype(dtype)


def _rk4(f, y0, dt, steps, keep_every=1):
    y = np.array(y0, dtype=np.float64)
    out = np.empty((steps, y.size))
    for k in range(steps):
        out[k] = y
        k1 = f(y)
        k2 = f(y + 0.5 * dt * k1)
        k3 = f(y + 0.5 * dt * k2)
        k4 = f(y + dt * k3)
        y = y + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
    return out


def lorenz_x(n, seed=0, dt=0.01, transient=10000, dtype=np.float32):
    """x-component of the Lorenz system (sigma 10, rho 28, beta 8/3), RK4,
    seeded N(0,1)^3 initial condition, transient dropped, z-normalised."""
    s, r, b = 10.0, 28.0, 8.0 / 3.0

    def f(y):
        return np.array([s * (y[1] - y[0]), y[0] * (r - y[2]) - y[1], y[0] * y[1] - b * y[2]])

    y0 = np.random.default_rng(seed).standard_normal(3)
    tr = _rk4(f, y0, dt, n + transient)
    return znorm(tr[transient:, 0]).astype(dtype)


def rossler_x(n, seed=0, dt=0.05, transient=20000, dtype=np.float32):
    """x-component of the Roessler system (a = b = 0.2, c = 5.7), RK4,
    seeded initial condition, transient dropped, z-normalised."""
    a, b, c = 0.2, 0.2, 5.7

    def f(y):
        return np.array([-y[1] - y[2], y[0] + a * y[1], b + y[2] * (y[0] - c)])

    y0 = np.random.default_rng(seed).standard_normal(3)
    tr = _rk4(f, y0, dt, n + transient)
    return znorm(tr[transient:, 0]).astype(dtype)


def ar1_sine(n, seed=0, phi=0.95, amp=0.3, period=1000.0, dtype=np.float32):
    """AR(1) x_t = phi x_{t-1} + xi_t, xi ~ N(0,1), plus amp*sin(2 pi t/period), z-normalised."""
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal(n)
    x = np.empty(n)
    v = 0.0
    for t in range(n):
        v = phi * v + xi[t]
        x[t] = v
    x = x + amp * np.sin(2 * np.pi * np.arange(n) / period)
    return znorm(x).astype(dtype)


def log_spaced(lo, hi, k):
    return np.exp(np.linspace(np.log(lo), np.log(hi), k))


# Configurations (sizes and threshold grids) -----------------------------------

def config1():
    """Logistic map, N = 10,000, tau 1, m 3, Euclidean, 5 thresholds, l_min 2."""
    return dict(x=logistic(10000, seed=1), tau=[1], emb=[3], thr=[0.05, 0.1, 0.15, 0.2, 0.25], lmin=[2], norms=[1])


def config2(n=131072, seed=2):
    """Lorenz x, N = 131,072, tau 8, m 3, both norms, 20 thresholds log-spaced 0.02..0.5 std."""
    return dict(x=lorenz_x(n, seed=seed), tau=[8], emb=[3], thr=list(log_spaced(0.02, 0.5, 20)), lmin=[2],
                norms=[1, 2])


def config3(n=262144, m=16, seed=3, dtype=np.float32):
    """AR(1)+sine, N = 262,144, tau 1, m in {16, 32, 64}, 16 thresholds linspace(0.1, 1.6)*sqrt(m), Euclidean."""
    return dict(x=ar1_sine(n, seed=seed, dtype=dtype), tau=[1], emb=[m],
                thr=list(np.linspace(0.1, 1.6, 16) * np.sqrt(m)), lmin=[2], norms=[1])


def config5(n=1048576, seed=5):
    """Roessler x, N = 1,048,576, tau 1, m 10, Euclidean, 32 thresholds log-spaced 0.01..0.3 std."""
    return dict(x=rossler_x(n, seed=seed), tau=[1], emb=[10], thr=list(log_spaced(0.01, 0.3, 32)), lmin=[2],
                norms=[1])


def config4_series(k, seed=4, nseries=4096):
    """Series k of the batched configuration: lengths log-uniform in
    [512, 32768]; a seeded mix of one third AR(1) with phi ~ U(-0.9, 0.9), one
    third logistic maps with r ~ U(3.5, 4.0), one third Gaussian white noise;
    fp32.  Each series is drawn from its own seed, so any subset can be made
    without the others."""
    rng = np.random.default_rng([seed, k])
    n = int(np.exp(rng.uniform(np.log(512), np.log(32768))))
    kind = k % 3
    if kind == 0:
        phi, xi, x, v = rng.uniform(-0.9, 0.9), rng.standard_normal(n), np.empty(n), 0.0
        for t in range(n):
            v = phi * v + xi[t]
            x[t] = v
    elif kind == 1:
        r, v, x = rng.uniform(3.5, 4.0), rng.uniform(0.1, 0.9), np.empty(n)
        for t in range(n):
            x[t] = v
            v = r * v * (1.0 - v)
    else:
        x = rng.standard_normal(n)
    return x.astype(np.float32)


def config4(sample=None, seed=4, nseries=4096):
    """4,096 fp32 series (config4_series), grid tau in {1, 2, 4} x m in
    {2, 3, 5} x 8 thresholds {0.05, ..., 0.40} x the series' std, both norms,
    l_min 2.  sample: indices of the series to make (all by default)."""
    idx = range(nseries) if sample is None else sample
    series = [config4_series(k, seed, nseries) for k in idx]
    return dict(series=series, tau=[1, 2, 4], emb=[2, 3, 5], thr_rel=[0.05 * (i + 1) for i in range(8)],
                lmin=[2], norms=[1, 2])
