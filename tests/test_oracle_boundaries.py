"""Pins the oracle's non-periodic boundaries (oracle/rrg_oracle.cpp,
wave_iteration_model; reference wave.cu:18-193, co_boundary.cuh:4-27,
nbody.cu:294-439) against what the mathematics fixes:

* absorbing layer: it changes only the dissipation, so the iteration's fixed
  point is still the discrete Poisson solution with the boundary values
  (checked against scipy's sparse solve), while transients decay faster;
* monopole Dirichlet: the fixed point is the Poisson solution whose boundary
  values are -M/r about the centre of mass;
* monopole radiative: a field equal to the monopole everywhere stays put, and
  an outgoing spherical wave psi = f(r - ct)/r is advanced by one step to
  first order in dx;
* centre of mass: the fixed-point reduction agrees with numpy and does not
  depend on particle order.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from rrgalaxsim_amd import ics

C = 4574.337022617616


def _poisson_dirichlet(nx, dx, f, bval):
    """Solve sum_nbrs(phi) - 6 phi = dx^2 f on interior cells, phi = bval on faces."""
    idx = -np.ones((nx, nx, nx), dtype=np.int64)
    inner = np.zeros((nx, nx, nx), dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    idx[inner] = np.arange(inner.sum())
    f3 = f.reshape(nx, nx, nx)   # [z, y, x]
    b3 = bval.reshape(nx, nx, nx)
    rows, cols, vals = [], [], []
    rhs = np.zeros(inner.sum())
    for z, y, x in zip(*np.nonzero(inner)):
        r = idx[z, y, x]
        rows.append(r); cols.append(r); vals.append(-6.0)
        rhs[r] = dx * dx * f3[z, y, x]
        for dz, dy, dx_ in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            zz, yy, xx = z + dz, y + dy, x + dx_
            if inner[zz, yy, xx]:
                rows.append(r); cols.append(idx[zz, yy, xx]); vals.append(1.0)
            else:
                rhs[r] -= b3[zz, yy, xx]
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(rhs), len(rhs)))
    out = b3.copy()
    out[inner] = spla.spsolve(A, rhs)
    return out.reshape(-1)


def _relax(orc, nx, dx, boundary, mass, iters, **kw):
    dt = orc.wave_dt(dx, C)
    prev = np.zeros(nx ** 3)
    cur = np.zeros(nx ** 3)
    for _ in range(iters):
        nxt = orc.wave_iteration_bc(nx, dx, 0.785, C, dt, boundary, cur, prev, mass, **kw)
        prev, cur = cur, nxt
    return cur, prev


def _coords(nx, dx):
    lo = -0.5 * nx * dx
    i = np.arange(nx)
    return lo + i * dx  # coord(i) = domain_min + i dx (wave.cu:18-20)


def _source(nx, dx, mass):
    return 4 * 3.14159265358979 * mass / dx ** 3


def test_absorbing_layer_keeps_the_dirichlet_poisson_fixed_point(orc):
    nx, dx = 12, 1.0
    rng = np.random.default_rng(1)
    mass = np.zeros(nx ** 3)
    m3 = mass.reshape(nx, nx, nx)
    m3[3:9, 3:9, 3:9] = rng.uniform(0, 1e-3, (6, 6, 6))
    ref = _poisson_dirichlet(nx, dx, _source(nx, dx, mass), np.zeros(nx ** 3))
    plain, _ = _relax(orc, nx, dx, "zero_dirichlet", mass, 4000)
    layer, _ = _relax(orc, nx, dx, "zero_dirichlet", mass, 4000, bc_frac=0.4, diss_extra=2.0)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(plain - ref)) < 1e-9 * scale
    assert np.max(np.abs(layer - ref)) < 1e-9 * scale
    # ... but the transients differ, and the layer damps them faster
    p50, _ = _relax(orc, nx, dx, "zero_dirichlet", mass, 60)
    l50, _ = _relax(orc, nx, dx, "zero_dirichlet", mass, 60, bc_frac=0.4, diss_extra=2.0)
    assert np.max(np.abs(p50 - l50)) > 1e-6 * scale


def test_layer_dissipation_profile(orc):
    """One step from phi = 0, phi_old = 1, no source leaves K_OLD = (Q-1)/(Q+1)
    in every interior cell; inverting it gives the per-cell dissipation,
    which must be diss_base + diss_extra * sum over axes of
    max(0, |x| - sim_l/2)^2 with x = domain_min + i dx (wave.cu:18-33)."""
    nx, dx, bcf, base, extra = 16, 0.5, 0.375, 0.785, 1.7
    dt = orc.wave_dt(dx, C)
    out = orc.wave_iteration_bc(nx, dx, base, C, dt, "zero_dirichlet", np.zeros(nx ** 3), np.ones(nx ** 3),
                                np.zeros(nx ** 3), bc_frac=bcf, diss_extra=extra)
    q = (1 + out) / (1 - out)                 # Q from K_OLD = (Q - 1) / (1 + Q)
    diss = (q * 2 / (C * dt)).reshape(nx, nx, nx)[1:-1, 1:-1, 1:-1]
    c = _coords(nx, dx)
    x_in = 0.5 * (1 - bcf) * nx * dx
    lay = extra * np.maximum(0, np.abs(c) - x_in) ** 2
    Z, Y, X = np.meshgrid(lay, lay, lay, indexing="ij")
    ref = (base + X + Y + Z)[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(diss - ref)) < 1e-9 * np.max(ref)
    assert np.max(ref) > 2 * base  # the layer is active in this geometry


def test_layer_damps_a_free_perturbation(orc):
    nx, dx = 14, 1.0
    rng = np.random.default_rng(2)
    phi0 = rng.standard_normal(nx ** 3)
    zero = np.zeros(nx ** 3)
    dt = orc.wave_dt(dx, C)
    out = {}
    for extra in (0.0, 3.0):
        prev, cur = phi0.copy(), phi0.copy()
        for _ in range(40):
            nxt = orc.wave_iteration_bc(nx, dx, 0.3, C, dt, "zero_dirichlet", cur, prev, zero, bc_frac=0.5,
                                        diss_extra=extra)
            prev, cur = cur, nxt
        out[extra] = np.sum(cur ** 2)
    assert out[3.0] < 0.9 * out[0.0]


def test_monopole_dirichlet_fixed_point(orc):
    nx, dx = 12, 1.0
    com = np.array([0.3, -0.2, 0.1])
    rng = np.random.default_rng(3)
    mass = np.zeros(nx ** 3)
    mass.reshape(nx, nx, nx)[4:8, 4:8, 4:8] = rng.uniform(0, 1e-2, (4, 4, 4))
    c = _coords(nx, dx)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.maximum(np.sqrt((X - com[0]) ** 2 + (Y - com[1]) ** 2 + (Z - com[2]) ** 2), 0.5 * dx)
    bval = (-1.0 / r).reshape(-1)
    ref = _poisson_dirichlet(nx, dx, _source(nx, dx, mass), bval)
    got, _ = _relax(orc, nx, dx, "monopole_dirichlet", mass, 4000, com=com, bc_mass=1.0, bc_frac=0.3,
                    diss_extra=1.0)
    assert np.max(np.abs(got - ref)) < 1e-9 * np.max(np.abs(ref))


def test_radiative_boundary_keeps_the_monopole(orc):
    nx, dx = 10, 1.0
    com = np.array([0.21, -0.37, 0.05])
    c = _coords(nx, dx)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.maximum(np.sqrt((X - com[0]) ** 2 + (Y - com[1]) ** 2 + (Z - com[2]) ** 2), 0.5 * dx)
    mono = (-0.8 / r).reshape(-1)
    out = orc.wave_iteration_bc(nx, dx, 0.785, C, orc.wave_dt(dx, C), "monopole_radiative", mono, mono,
                                np.zeros(nx ** 3), com=com, bc_mass=0.8)
    face = np.ones((nx, nx, nx), dtype=bool)
    face[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(out[face.reshape(-1)], mono[face.reshape(-1)])


def test_radiative_boundary_advances_an_outgoing_wave(orc):
    """psi = f(r - c t)/r solves d_t psi + c (d_r psi + psi/r) = 0; one
    boundary update must follow it to first order in dx."""
    nx, dx = 40, 0.25
    com = np.zeros(3)
    dt = orc.wave_dt(dx, C)
    c = _coords(nx, dx)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    r0, w = 4.0, 2.0

    def psi(t):
        return (np.exp(-((r - C * t - r0) / w) ** 2) / np.maximum(r, dx)).reshape(-1)

    mono = (-1.0 / np.maximum(r, 0.5 * dx)).reshape(-1)
    face = np.ones((nx, nx, nx), dtype=bool)
    face[1:-1, 1:-1, 1:-1] = False
    assert np.min(r[face]) > 4 * dx  # psi is exact f(r - ct)/r wherever the update reads it
    out = orc.wave_iteration_bc(nx, dx, 0.785, C, dt, "monopole_radiative", psi(0.0) + mono, psi(0.0) + mono,
                                np.zeros(nx ** 3), com=com, bc_mass=1.0)
    fm = face.reshape(-1)
    got = (out - mono - psi(0.0))[fm]        # the scheme's increment
    exact = (psi(dt) - psi(0.0))[fm]         # the exact increment
    assert np.linalg.norm(got - exact) < 0.15 * np.linalg.norm(exact)
    # the upwind side matters: taking the downwind neighbour breaks this
    assert np.corrcoef(got, exact)[0, 1] > 0.98


def test_center_of_mass_fixed_point(orc):
    pos, _, m = ics.exp_disk(5000, seed=7)
    pos = pos + np.array([0.3, -1.2, 0.05])
    m = m * np.random.default_rng(7).uniform(0.5, 1.5, len(m))
    e = 61 - int(np.frexp(np.sum(np.abs(m)))[1])
    got = orc.center_of_mass(pos, m, e)
    ref = np.average(pos, axis=0, weights=m)
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(pos))
    perm = np.random.default_rng(8).permutation(len(m))
    assert np.array_equal(orc.center_of_mass(pos[perm], m[perm], e), got)
    assert np.array_equal(orc.center_of_mass(pos, 0 * m, e), np.zeros(3))


@pytest.mark.parametrize("boundary", ["monopole_dirichlet", "monopole_radiative"])
def test_pm_run_with_monopole_boundaries(orc, boundary):
    pos, vel, m = ics.exp_disk(3000, seed=1)
    o = orc.pm_run(pos, vel, m, 3, nx=16, dx=1.0, dt=5e-3, diss_base=0.785, wave_speed=C, setup_iterations=200,
                   iterations=30, boundary=boundary, bc_frac=0.25, diss_extra=1.0, fixed_e=60)
    assert np.all(np.isfinite(o["pos"])) and np.all(np.isfinite(o["phi"]))
    # an isolated disk: its centre of mass barely moves
    assert np.max(np.abs(np.average(o["pos"], axis=0, weights=m) - np.average(pos, axis=0, weights=m))) < 1e-3
