"""Pins the CPU oracle's PM restatement (oracle/rrg_oracle.cpp; reference
rm_p2mesh.cu:176-290, rm_mesh2p.cu:48-114, wave.cu:106-297,
wave_steady_checker.cu:95-143) against what the mathematics fixes:

* CIC: single-particle weights at a cell centre / cell corner, mass
  conservation, clamping at the faces, and the transpose identity
  <CIC(m), g> = <m, gather(g)> with the trilinear gather (same weights).
* fixed point: agrees with the float deposit to the fixed resolution and is
  bitwise independent of particle order.
* wave: the periodic scheme's fixed point solves the 7-point discrete Poisson
  equation, checked against numpy.fft with eigenvalues (2cos(2 pi k/n)-2)/dx^2;
  a uniform field with rho = rho_mean is steady; the scheme is linear.
* gradient: exact on linear fields (faces too) and on quadratics inside.
* gather: reproduces linear fields exactly between the outer cell centres.
* steady check: a synthetic pair with a known ratio.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

C = 4574.337022617616


def centres(nx, dx):
    lo = -0.5 * nx * dx
    return lo + (np.arange(nx) + 0.5) * dx


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cic_cell_centre_and_corner(orc, dtype):
    nx, dx = 8, 0.5
    c = centres(nx, dx)
    # particle on the centre of cell (2, 5, 3): all mass there
    pos = np.array([[c[2], c[5], c[3]]])
    g = orc.cic_float(pos, np.array([2.0]), nx, dx, dtype=dtype).reshape(nx, nx, nx)
    assert g[3, 5, 2] == 2.0 and np.count_nonzero(g) == 1
    # particle on the corner shared by cells (2..3, 5..6, 3..4): 1/8 each
    pos = np.array([[c[2] + dx / 2, c[5] + dx / 2, c[3] + dx / 2]])
    g = orc.cic_float(pos, np.array([8.0]), nx, dx, dtype=dtype).reshape(nx, nx, nx)
    assert np.count_nonzero(g) == 8
    assert np.all(g[3:5, 5:7, 2:4] == 1.0)


def test_cic_clamps_outside(orc):
    nx, dx = 4, 1.0
    pos = np.array([[-10.0, -10.0, -10.0], [10.0, 10.0, 10.0]])
    g = orc.cic_float(pos, np.array([1.0, 3.0]), nx, dx).reshape(nx, nx, nx)
    assert g[0, 0, 0] == 1.0
    assert np.isclose(g[3, 3, 3], 3.0 * (1 - 1e-6 + 1e-6) ** 3, rtol=1e-5)
    assert np.isclose(g.sum(), 4.0, rtol=1e-12)


def test_cic_mass_and_transpose_identity(orc):
    nx, dx = 16, 0.5
    rng = np.random.default_rng(0)
    half = nx * dx / 2 - dx * 0.6  # strictly between the outer cell centres
    pos = rng.uniform(-half, half, (2000, 3))
    m = rng.uniform(0.5, 2.0, 2000)
    grid = orc.cic_float(pos, m, nx, dx)
    assert np.isclose(grid.sum(), m.sum(), rtol=1e-13)
    f = rng.standard_normal((nx ** 3, 3))
    lhs = grid @ f[:, 0]
    rhs = m @ orc.gather(pos, nx, dx, f)[:, 0]
    assert np.isclose(lhs, rhs, rtol=1e-12)


def test_fixed_point_matches_float_and_is_order_free(orc):
    pos, _, m = ics.exp_disk(20_000, seed=1)
    e = 60
    gf = orc.cic_float(pos, m, 32, 0.5)
    gx = orc.cic_fixed(pos, m, 32, 0.5, e)
    back = orc.fixed_to_real(gx, e)
    assert np.max(np.abs(back - gf)) < 1e-14
    perm = np.random.default_rng(2).permutation(len(m))
    assert np.array_equal(orc.cic_fixed(pos[perm], m[perm], 32, 0.5, e), gx)


def _fft_poisson(src, nx, dx):
    k = 2 * np.pi * np.fft.fftfreq(nx)
    lam = (2 * np.cos(k) - 2) / dx ** 2
    L = lam[:, None, None] + lam[None, :, None] + lam[None, None, :]
    s = np.fft.fftn(src.reshape(nx, nx, nx))
    L[0, 0, 0] = 1.0
    phi = s / L
    phi[0, 0, 0] = 0.0
    return np.real(np.fft.ifftn(phi)).ravel()


def test_periodic_wave_fixed_point_is_discrete_poisson(orc):
    nx, dx = 12, 16.0 / 12
    rng = np.random.default_rng(1)
    mass = rng.uniform(0.5, 1.5, nx ** 3)
    mass /= mass.sum()  # total mass 1, as the live configuration assumes
    dt = orc.wave_dt(dx, C)
    rm = orc.rho_mean(1.0, nx, dx)
    z = np.zeros(nx ** 3)
    prev, curr, it = orc.solve_wave(nx, dx, 2.0, C, dt, True, rm, z, z, mass, 20_000, 1e-13, 50, 3)
    assert it < 20_000
    rho = mass / dx ** 3 - rm
    ref = _fft_poisson(4 * np.pi * rho, nx, dx)
    got = curr - curr.mean()
    assert np.max(np.abs(got - ref)) < 1e-9 * np.max(np.abs(ref))


def test_wave_uniform_steady_and_linear(orc):
    nx, dx = 6, 1.0
    dt = orc.wave_dt(dx, C)
    rm = orc.rho_mean(1.0, nx, dx)
    mass = np.full(nx ** 3, rm * dx ** 3)
    phi = np.full(nx ** 3, -0.3)
    out = orc.wave_iteration(nx, dx, 0.5, C, dt, True, rm, phi, phi, mass)
    assert np.allclose(out, phi, rtol=1e-14, atol=0)
    rng = np.random.default_rng(3)
    p1, o1, p2, o2 = (rng.standard_normal(nx ** 3) for _ in range(4))
    zero_src = np.full(nx ** 3, rm * dx ** 3)  # rho - rho_mean = 0
    w = lambda a, b: orc.wave_iteration(nx, dx, 0.5, C, dt, True, rm, a, b, zero_src)  # noqa: E731
    assert np.allclose(w(p1 + 2 * p2, o1 + 2 * o2), w(p1, o1) + 2 * w(p2, o2), rtol=1e-12, atol=1e-12)


def test_wave_zero_dirichlet_edges(orc):
    nx, dx = 6, 1.0
    rng = np.random.default_rng(0)
    phi = rng.standard_normal(nx ** 3)
    out = orc.wave_iteration(nx, dx, 0.5, C, orc.wave_dt(dx, C), False, 0.0, phi, phi,
                             rng.uniform(0, 1, nx ** 3)).reshape(nx, nx, nx)
    assert np.all(out[0] == 0) and np.all(out[-1] == 0) and np.all(out[:, 0] == 0) and np.all(out[:, :, -1] == 0)
    assert np.all(out[1:-1, 1:-1, 1:-1] != 0)


def test_wave_coefficients_cancel_wave_speed(orc):
    """With the CFL step dt = dx / (2 sqrt(3) c), K = c^2 dt^2 / dx^2 = 1/12."""
    for dx in (0.125, 0.5):
        dt = orc.wave_dt(dx, C)
        k = orc.wave_coef(0.0, C, dt, dx)
        assert np.isclose(k["k_main"], 1.0 / 12, rtol=1e-12)
        assert np.isclose(k["k_actual"], 2 * (1 - 3 / 12), rtol=1e-12)
        assert k["k_old"] == -1.0


def test_gradient_linear_and_quadratic(orc):
    nx, dx = 10, 0.25
    c = centres(nx, dx)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")  # X varies along axis 0 = z index in x-fastest layout
    # x-fastest layout: flat index = ix + iy*nx + iz*nx^2 -> array [iz, iy, ix]
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    lin = (2.0 * xx - 3.0 * yy + 0.5 * zz).ravel()
    a = orc.gradient(nx, dx, lin)
    assert np.allclose(a, [-2.0, 3.0, -0.5], rtol=1e-12, atol=1e-12)
    quad = (xx ** 2).ravel()
    a = orc.gradient(nx, dx, quad).reshape(nx, nx, nx, 3)
    inner = a[:, :, 1:-1, 0]
    assert np.allclose(inner, -2.0 * xx[:, :, 1:-1], rtol=1e-12, atol=1e-12)
    # one-sided at the faces: -(phi[1]-phi[0])/dx = -(x0 + x1)
    assert np.allclose(a[:, :, 0, 0], -(c[0] + c[1]), rtol=1e-12)


def test_gather_reproduces_linear_field(orc):
    nx, dx = 8, 0.5
    c = centres(nx, dx)
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    A = np.array([[1.0, 2.0, -1.0], [0.5, 0.0, 3.0], [-2.0, 1.0, 1.0]])
    b = np.array([0.1, -0.2, 0.3])
    X = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    field = X @ A.T + b
    rng = np.random.default_rng(0)
    pos = rng.uniform(c[0], c[-1], (500, 3))
    got = orc.gather(pos, nx, dx, field)
    assert np.allclose(got, pos @ A.T + b, rtol=1e-12, atol=1e-12)


def test_steady_err_known_ratio(orc):
    a = np.array([1.0, -4.0, 2.0, 0.5])
    b = a.copy()
    b[2] += 0.02
    assert np.isclose(orc.steady_err(a, b), 0.02 / 4.0, rtol=1e-14)


def test_pm_run_conserves_momentum_roughly(orc):
    """End-to-end sanity: a PM run of a centred disk keeps its centre of mass."""
    pos, vel, m = ics.exp_disk(5000, seed=1)
    o = orc.pm_run(pos, vel, m, 3, nx=32, dx=0.5, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=300,
                   iterations=50)
    assert np.all(np.isfinite(o["pos"])) and np.all(np.isfinite(o["phi"]))
    com0 = np.average(pos, axis=0, weights=m)
    com1 = np.average(o["pos"], axis=0, weights=m)
    assert np.max(np.abs(com1 - com0)) < 1e-3
    assert o["iterations"][0] == 300 and np.all(o["iterations"][1:] == 50)
