"""Seeded test fields shared by the CPU and GPU tests."""
import numpy as np


def periodic_field(cfg, seed, amp=0.05):
    """A smooth periodic field plus seeded noise, evaluated at i mod im (and
    j, k alike), so the duplicate planes i = 0, im hold identical values and
    qswap's averages leave it unchanged.  Returns q (5, km+1, jm+1, im+1)."""
    rng = np.random.default_rng(seed)
    im, jm, km = cfg.im, cfg.jm, cfg.km
    ii = np.arange(im + 1) % im
    jj = np.arange(jm + 1) % jm
    kk = np.arange(km + 1) % km if km > 0 else np.zeros(1, dtype=int)
    x = 2 * np.pi * ii / im
    y = 2 * np.pi * jj / jm
    z = 2 * np.pi * kk / km if km > 0 else np.zeros(1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    nz = lambda: rng.standard_normal((max(km, 1), jm, im))  # noqa: E731
    ext = lambda a: a[np.ix_(kk, jj, ii)]  # noqa: E731
    rho = 1.0 + amp * ext(nz())
    u = np.sin(X) * np.cos(Y) * np.cos(Z) + amp * ext(nz())
    v = -np.cos(X) * np.sin(Y) * np.cos(Z) + amp * ext(nz())
    w = (amp * ext(nz())) if km > 0 else np.zeros_like(u)
    p = 1.0 / (1.4 * cfg.mach ** 2) + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16.0
    q = np.empty((5,) + u.shape)
    q[0] = rho
    q[1], q[2], q[3] = rho * u, rho * v, rho * w
    q[4] = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return q


def double_shock_tube(im, jm, km=0):
    """A periodic double shock tube along x: Sod's states, rho = 1, p = 1 for
    x in [pi/2, 3 pi/2) and rho = 0.125, p = 0.1 elsewhere, at rest, the
    same on every y (and z) line.  Returns q (5, km+1, jm+1, im+1) for the
    nondimensional E = p/(gamma-1) + rho |u|^2/2 with gamma = 1.4."""
    ii = np.arange(im + 1) % im
    x = 2 * np.pi * ii / im
    inside = (x >= np.pi / 2) & (x < 3 * np.pi / 2)
    shape = (km + 1, jm + 1, im + 1)
    rho = np.broadcast_to(np.where(inside, 1.0, 0.125), shape)
    p = np.broadcast_to(np.where(inside, 1.0, 0.1), shape)
    q = np.zeros((5,) + shape)
    q[0] = rho
    q[4] = p / 0.4
    return q
