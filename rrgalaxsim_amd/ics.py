"""Seeded initial conditions for the benchmark and test workloads.

Generators use numpy's PCG64 (``np.random.default_rng(seed)``), so every
machine produces the same doubles for the same seed.  Units: G = 1.

* ``plummer``: Plummer sphere sampled as in Aarseth, Henon & Wielen (1974):
  radius from the inverse cumulative mass, speed by von Neumann rejection on
  g(q) = q^2 (1 - q^2)^(7/2) times the local escape speed, isotropic
  directions, then recentred to zero centre-of-mass position and velocity.
* ``exp_disk``: exponential disk, surface density ~ exp(-R) (scale 1),
  sech^2 vertical profile z = h atanh(2u - 1), circular speed
  v_c^2 = M(<R)/R with M(<R) = M (1 - (1 + R) e^-R) (plus an optional extra
  enclosed-mass term), plus isotropic Gaussian dispersion sigma = f v_c;
  radii are truncated at ``r_max`` and |z| at ``z_max``.

Also the reference's ini text format (header "N t", rows "x y z vx vy vz";
solver/lib/rrgsim_io/src/io_particles_csv.cpp:11-87) for the CLI.
"""
from __future__ import annotations

import numpy as np


def _isotropic(rng, n):
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)


def plummer(n, *, seed=0, mass=1.0, a=1.0, mfrac_max=0.999, recenter=True, rng=None):
    """Plummer sphere (M, a, G=1).  Returns pos (n,3), vel (n,3), m (n,)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    x1 = rng.uniform(0.0, mfrac_max, n)
    x1 = np.maximum(x1, 1e-12)
    r = a / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _isotropic(rng, n)
    # speed: q = v / v_esc with density g(q) = q^2 (1-q^2)^3.5, max 0.092
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 * mass / a) * (1.0 + (r / a) ** 2) ** (-0.25)
    vel = (q * vesc)[:, None] * _isotropic(rng, n)
    m = np.full(n, mass / n)
    if recenter:
        pos -= np.average(pos, axis=0, weights=m)
        vel -= np.average(vel, axis=0, weights=m)
    return pos, vel, m


def _exp_disk_radius(u):
    """Invert M(<R)/M = 1 - (1+R) e^-R = u by Newton's method (vectorised)."""
    R = np.where(u < 0.5, np.sqrt(2.0 * u), -np.log1p(-u) + 1.0)
    for _ in range(60):
        f = 1.0 - (1.0 + R) * np.exp(-R) - u
        df = R * np.exp(-R)
        step = f / np.maximum(df, 1e-300)
        R = np.maximum(R - step, 1e-12)
        if np.max(np.abs(step)) < 1e-15:
            break
    return R


def exp_disk(n, *, seed=1, mass=1.0, h=0.1, sigma_frac=0.1, r_max=7.5, z_max=None, extra_mass=None, rng=None):
    """Exponential disk (scale 1).  ``extra_mass(R)`` adds enclosed mass from
    another component to the circular speed.  Returns pos, vel, m."""
    rng = np.random.default_rng(seed) if rng is None else rng
    umax = 1.0 - (1.0 + r_max) * np.exp(-r_max)
    R = _exp_disk_radius(rng.uniform(0.0, umax, n))
    th = rng.uniform(0.0, 2.0 * np.pi, n)
    if z_max is None:
        z_max = 7.5 * h
    zu = np.tanh(z_max / h)
    z = h * np.arctanh(rng.uniform(-zu, zu, n))
    pos = np.stack([R * np.cos(th), R * np.sin(th), z], axis=1)
    menc = mass * (1.0 - (1.0 + R) * np.exp(-R))
    if extra_mass is not None:
        menc = menc + extra_mass(R)
    vc = np.sqrt(menc / R)
    vel = np.stack([-vc * np.sin(th), vc * np.cos(th), np.zeros(n)], axis=1)
    vel += (sigma_frac * vc)[:, None] * rng.standard_normal((n, 3))
    m = np.full(n, mass / n)
    return pos, vel, m


def disk_bulge(n_disk=196_608, n_bulge=65_536, *, seed=4, bulge_mass=0.2, bulge_a=0.2):
    """Disk (M=1) + Plummer bulge (M=0.2, a=0.2); the disk's circular speed
    includes the bulge's enclosed mass M_b R^3 / (R^2 + a^2)^(3/2)."""
    rng = np.random.default_rng(seed)
    bp, bv, bm = plummer(n_bulge, mass=bulge_mass, a=bulge_a, rng=rng, recenter=False)
    dp, dv, dm = exp_disk(n_disk, rng=rng,
                          extra_mass=lambda R: bulge_mass * R ** 3 / (R * R + bulge_a * bulge_a) ** 1.5)
    pos = np.concatenate([dp, bp]); vel = np.concatenate([dv, bv]); m = np.concatenate([dm, bm])
    pos -= np.average(pos, axis=0, weights=m)
    vel -= np.average(vel, axis=0, weights=m)
    return pos, vel, m


def ensemble_sizes(k=64, *, seed=2, lo=12, hi=16):
    """N_g = round(2^U(lo, hi)) for k galaxies."""
    rng = np.random.default_rng(seed)
    return np.rint(2.0 ** rng.uniform(lo, hi, k)).astype(np.int64)


def write_ini(path, pos, vel, t=0.0):
    """Write the reference's ini text format at full double precision."""
    pos = np.asarray(pos, dtype=np.float64); vel = np.asarray(vel, dtype=np.float64)
    with open(path, "w") as f:
        f.write(f"{len(pos)} {t:.6f}\n")
        np.savetxt(f, np.hstack([pos, vel]), fmt="%.17g")
