"""Pins the oracle's reverse-mode gradient of the PM run (oracle/rrg_oracle.cpp
pm_adjoint; forward map = pm_run, reference rrgsim_integration_rect.cuh:53-112)
by central finite differences of pm_run itself, on both boundary types, and the
gradient transpose by the dot-product identity <D phi, g> = <phi, D^T g>.

The loss is linear in the final state, L = <gx, x_T> + <gv, v_T>, and each
check compares a directional derivative along a seeded random direction in
x0, v0 or m.  The deposit's fixed-point rounding (2^-60 here) is far below the
finite-difference resolution.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

C = 4574.337022617616


def _cfg(periodic, **kw):
    cfg = dict(nx=16, dx=1.0, dt=0.05, diss_base=0.785, wave_speed=C, setup_iterations=20, iterations=6,
               periodic=periodic, fixed_e=60)
    cfg.update(kw)
    return cfg


def _loss(orc, pos, vel, m, steps, gx, gv, cfg):
    o = orc.pm_run(pos, vel, m, steps, **cfg)
    return float(np.sum(gx * o["pos"]) + np.sum(gv * o["vel"])), o["iterations"]


@pytest.mark.parametrize("periodic", [True, False])
def test_pm_adjoint_matches_finite_differences(orc, periodic):
    n, steps = 300, 5
    pos, vel, m = ics.exp_disk(n, seed=2)
    cfg = _cfg(periodic)
    rng = np.random.default_rng(7)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    gx0, gv0, gm = orc.pm_adjoint(pos, vel, m, steps, gx, gv, **cfg)
    assert np.all(np.isfinite(gx0)) and np.all(np.isfinite(gm))
    for what in ("x", "v", "m"):
        d = rng.standard_normal((n, 3)) if what != "m" else m * rng.standard_normal(n)
        h = 1e-6
        args = {"x": lambda s: (pos + s * d, vel, m), "v": lambda s: (pos, vel + s * d, m),
                "m": lambda s: (pos, vel, m + s * d)}[what]
        lp, _ = _loss(orc, *args(h), steps, gx, gv, cfg)
        lm, _ = _loss(orc, *args(-h), steps, gx, gv, cfg)
        fd = (lp - lm) / (2 * h)
        an = float(np.sum({"x": gx0, "v": gv0, "m": gm}[what] * d))
        assert abs(fd - an) < 1e-6 * (abs(fd) + 1e-3), (what, fd, an)


def test_pm_adjoint_mass_gradient_nonzero_and_per_particle(orc):
    """Single-particle derivatives (not only directional sums) for a few
    particles, so a transposed index would show."""
    n, steps = 120, 4
    pos, vel, m = ics.exp_disk(n, seed=4)
    cfg = _cfg(True, iterations=4)
    rng = np.random.default_rng(3)
    gx = rng.standard_normal((n, 3))
    gv = np.zeros((n, 3))
    gx0, gv0, gm = orc.pm_adjoint(pos, vel, m, steps, gx, None, **cfg)
    for q, c in ((0, 0), (37, 2), (101, 1)):
        for what, grad in (("x", gx0), ("v", gv0)):
            e = np.zeros((n, 3)); e[q, c] = 1.0
            h = 1e-6
            a = (pos + h * e, vel, m) if what == "x" else (pos, vel + h * e, m)
            b = (pos - h * e, vel, m) if what == "x" else (pos, vel - h * e, m)
            fd = (_loss(orc, *a, steps, gx, gv, cfg)[0] - _loss(orc, *b, steps, gx, gv, cfg)[0]) / (2 * h)
            assert abs(fd - grad[q, c]) < 1e-6 * (abs(fd) + 1e-3), (what, q, c, fd, grad[q, c])
        e = np.zeros(n); e[q] = 1e-6
        fd = (_loss(orc, pos, vel, m + e, steps, gx, gv, cfg)[0] -
              _loss(orc, pos, vel, m - e, steps, gx, gv, cfg)[0]) / 2e-6
        assert abs(fd - gm[q]) < 1e-5 * (abs(fd) + 1e-2), (q, fd, gm[q])
    assert np.max(np.abs(gm)) > 0


def test_pm_adjoint_with_steady_check_uses_recorded_counts(orc):
    """With a tolerance the stencil stops early; the gradient is that of the
    run with the recorded iteration counts (checked where +-h keep them)."""
    n, steps = 200, 3
    pos, vel, m = ics.exp_disk(n, seed=6)
    cfg = _cfg(True, setup_iterations=400, iterations=60, tolerance=2e-3, check_step=5, hold=2)
    rng = np.random.default_rng(11)
    gx = rng.standard_normal((n, 3))
    gv = np.zeros((n, 3))
    base = orc.pm_run(pos, vel, m, steps, **cfg)["iterations"]
    assert base[0] < 400  # the check did stop the setup solve
    gx0, _, _ = orc.pm_adjoint(pos, vel, m, steps, gx, None, **cfg)
    d = rng.standard_normal((n, 3))
    h = 1e-6
    lp, ip = _loss(orc, pos + h * d, vel, m, steps, gx, gv, cfg)
    lm, im = _loss(orc, pos - h * d, vel, m, steps, gx, gv, cfg)
    assert np.array_equal(ip, base) and np.array_equal(im, base)
    fd = (lp - lm) / (2 * h)
    assert abs(fd - np.sum(gx0 * d)) < 1e-6 * (abs(fd) + 1e-3)


@pytest.mark.parametrize("nx", [2, 3, 9])
def test_gradient_transpose_dot_product(orc, nx):
    rng = np.random.default_rng(nx)
    phi = rng.standard_normal(nx ** 3)
    g = rng.standard_normal((nx ** 3, 3))
    lhs = np.sum(orc.gradient(nx, 0.7, phi) * g)
    rhs = np.sum(phi * orc.gradient_t(g, nx, 0.7))
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)


def test_pm_adjoint_zero_dirichlet_with_absorbing_layer(orc):
    """Off-periodic grids carry per-cell coefficients (the absorbing layer),
    so the transposed stencil is L(K_main mu) rather than K_main L(mu) and the
    source adjoint is scaled per cell."""
    n, steps = 300, 4
    pos, vel, m = ics.exp_disk(n, seed=12)
    cfg = _cfg(False, bc_frac=0.5, diss_extra=4.0)
    rng = np.random.default_rng(13)
    gx, gv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    gx0, gv0, gm = orc.pm_adjoint(pos, vel, m, steps, gx, gv, **cfg)
    for what in ("x", "v", "m"):
        d = rng.standard_normal((n, 3)) if what != "m" else m * rng.standard_normal(n)
        h = 1e-6
        args = {"x": lambda s: (pos + s * d, vel, m), "v": lambda s: (pos, vel + s * d, m),
                "m": lambda s: (pos, vel, m + s * d)}[what]
        fd = (_loss(orc, *args(h), steps, gx, gv, cfg)[0] - _loss(orc, *args(-h), steps, gx, gv, cfg)[0]) / (2 * h)
        an = float(np.sum({"x": gx0, "v": gv0, "m": gm}[what] * d))
        assert abs(fd - an) < 1e-6 * (abs(fd) + 1e-3), (what, fd, an)
    # the layer matters for the gradient
    gx0_plain, _, _ = orc.pm_adjoint(pos, vel, m, steps, gx, gv, **_cfg(False))
    assert np.max(np.abs(gx0 - gx0_plain)) > 1e-6 * np.max(np.abs(gx0))


def test_pm_adjoint_rejects_monopole_boundaries(orc):
    pos, vel, m = ics.exp_disk(50, seed=1)
    cfg = _cfg(True)
    cfg["boundary"] = "monopole_dirichlet"
    with pytest.raises(ValueError):
        orc.pm_adjoint(pos, vel, m, 2, np.ones((50, 3)), None, **cfg)
