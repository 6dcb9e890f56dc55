"""Seeded synthetic workloads (numpy PCG64), float64 coordinates and charges.

uniform     N points in [0,1]^2 (Laplace, p=32, cutoff 50 in the reference's CPU-runnable case)
clustered   70% of N from 8 isotropic Gaussians (centres U[0.1,0.9]^2, sigma log-uniform in
            [1e-3, 5e-2], redrawn until inside [0,1]^2) followed by 30% uniform points
annulus     uniform in the annulus 0.2 < |x - (0.5,0.5)| < 0.5
star_curve  points equispaced in theta on r(theta) = 1 + 0.3 cos(5 theta + phi)
"""
import numpy as np


def uniform(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.random(n)
    y = rng.random(n)
    tau = rng.standard_normal(n)
    return x, y, tau


def clustered(n, seed=2, n_clusters=8, frac=0.7):
    rng = np.random.default_rng(seed)
    nc = int(round(frac * n))
    per = [nc // n_clusters + (1 if i < nc % n_clusters else 0) for i in range(n_clusters)]
    xs, ys = [], []
    for k in range(n_clusters):
        cx, cy = rng.uniform(0.1, 0.9, size=2)
        sig = float(np.exp(rng.uniform(np.log(1e-3), np.log(5e-2))))
        px = np.empty(0)
        py = np.empty(0)
        need = per[k]
        while need > 0:
            a = rng.normal(cx, sig, size=2 * need + 16)
            b = rng.normal(cy, sig, size=2 * need + 16)
            ok = (a >= 0) & (a <= 1) & (b >= 0) & (b <= 1)
            a, b = a[ok][:need], b[ok][:need]
            px = np.concatenate([px, a])
            py = np.concatenate([py, b])
            need = per[k] - px.size
        xs.append(px)
        ys.append(py)
    nu = n - nc
    xs.append(rng.random(nu))
    ys.append(rng.random(nu))
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    tau = rng.standard_normal(n)
    return x, y, tau


def annulus(n, seed=3, nrhs=1):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(0.04, 0.25, size=n))
    th = rng.uniform(0.0, 2 * np.pi, size=n)
    x = 0.5 + r * np.cos(th)
    y = 0.5 + r * np.sin(th)
    tau = rng.standard_normal((nrhs, n)) if nrhs > 1 else rng.standard_normal(n)
    return x, y, tau


def star_curve(n, phi):
    th = 2 * np.pi * np.arange(n) / n
    r = 1.0 + 0.3 * np.cos(5 * th + phi)
    return r * np.cos(th), r * np.sin(th)


def star_curves(n_curves=64, seed=5, n_min=2e3, n_max=5e5):
    """Closed star curves r(theta) = 1 + 0.3 cos(5 theta + phi), phi ~ U[0, 2 pi), with
    n = round(exp(U[ln n_min, ln n_max])) points equispaced in theta: a list of (x, y)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_curves):
        phi = rng.uniform(0.0, 2 * np.pi)
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        out.append(star_curve(n, phi))
    return out


def sample_targets(n, m, seed=12):
    """m distinct point indices for sampled error checks."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=min(m, n), replace=False))
