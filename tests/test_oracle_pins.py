"""Pins two oracle semantics that no other CPU test fixes (oracle/rrg_oracle.cpp
ngp_phi and solve_wave_model):

* The NGP potential projection (rm_mesh2p.cu:29-38 with the cell rule of
  rm_p2mesh.cu:34-53): particle phi = phi of the cell given by the truncated
  (x - domain_min) / dx per axis, clamped to [0, nx - 1].  The expected values
  come from numpy truncation on a field whose value encodes (ix, iy, iz), with
  particles on cell faces, exactly at domain_min and domain_max, and outside
  the box on both sides.
* The steady-state early exit of solve_wave_equation (wave.cu:253-287) with
  SteadyStateChecker::check (wave_steady_checker.cu:131-143): iterate, rotate
  the buffers, and only then, when iter > 0 and iter % check_step == 0, test
  err = max|phi_curr - phi_prev| / (max|phi_curr| + 1e-30) < tol; `hold`
  consecutive passes stop the loop, a failed check resets the count.
  - An exactly steady field (uniform phi, zero source) passes every check, so
    the loop stops after check_step * hold + 1 stencil applications (16 for
    check_step 5 and hold 3; 91 for the reference's defaults 30 and 3), or at
    max_iter.
  - The Nyquist mode phi = A (-1)^ix on a periodic grid is exact in floating
    point: every cell's 7-point sum is 2 phi, so the amplitude follows the
    scalar recurrence A' = 2A k_main + A k_actual + A_old k_old of the
    scheme's coefficients (wave.cu:164-171), and the error at a check is
    |A' - A| / |A'|.  The oscillating mode passes and fails checks in turn;
    the expected count follows from the stopping rule alone, and the cases are
    chosen so that a dropped `iter &&` guard, a check before the rotation or a
    missing hold reset each give a different count.
"""
import numpy as np
import pytest


def _cell_field(nx, dtype):
    iz, iy, ix = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    return (ix + 100 * iy + 10000 * iz).astype(dtype).ravel()  # AT(ix, iy, iz): x fastest


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ngp_phi_truncates_and_clamps(orc, dtype):
    nx, dx = 16, 0.5  # box [-4, 4]; faces at -4 + k/2 are exact in binary
    lo, hi = -4.0, 4.0
    rng = np.random.default_rng(11)
    inside = rng.uniform(lo, hi, (300, 3))
    faces = lo + dx * rng.integers(0, nx + 1, (200, 3))  # on cell faces, domain_max included
    special = np.array([
        [lo, lo, lo],                 # exactly domain_min: cell 0
        [hi, hi, hi],                 # exactly domain_max: (L/dx) = nx, clamped to nx-1
        [lo - 0.1, 0.3, hi + 0.2],    # just outside: trunc(-0.2) = 0; above: clamped
        [-50.0, 3.99, 0.0],           # far below: clamped to 0
        [0.0, 50.0, -4.25],           # far above / below
        [1.0, -2.5, 3.5],             # interior faces
    ])
    pos = np.concatenate([inside, faces, special]).astype(dtype)
    phi = _cell_field(nx, dtype)
    got = orc.ngp_phi(pos, nx, dx, phi, dtype=dtype)
    idx = np.clip(((pos - dtype(lo)) / dtype(dx)).astype(np.int64), 0, nx - 1)  # C (int) cast truncates
    want = (idx[:, 0] + 100 * idx[:, 1] + 10000 * idx[:, 2]).astype(dtype)
    assert np.array_equal(got, want)
    # the special rows, spelled out
    assert got[-6] == 0 and got[-5] == 15 + 100 * 15 + 10000 * 15
    assert got[-4] == 0 + 100 * 8 + 10000 * 15
    assert got[-3] == 0 + 100 * 15 + 10000 * 8
    assert got[-2] == 8 + 100 * 15 + 0
    assert got[-1] == 10 + 100 * 3 + 10000 * 15


def _solve(orc, nx, field_prev, field_curr, max_iter, tol, check_step, hold, diss=0.05):
    dx, c = 1.0, 1.0
    dt = 0.5 * dx / (c * np.sqrt(3.0))
    m = np.zeros(nx ** 3)
    return orc.solve_wave(nx, dx, diss, c, dt, 1, 0.0, field_prev, field_curr, m, max_iter, tol,
                          check_step=check_step, hold=hold)


@pytest.mark.parametrize("check_step,hold,max_iter,want", [
    (5, 3, 1000, 16), (30, 3, 1000, 91), (5, 3, 12, 12), (1, 1, 100, 2), (7, 2, 100, 15)])
def test_early_exit_on_a_steady_field(orc, check_step, hold, max_iter, want):
    nx = 6
    u = np.ones(nx ** 3)
    prev, curr, it = _solve(orc, nx, u, u, max_iter, 1e-6, check_step, hold)
    assert it == want == min(max_iter, check_step * hold + 1)
    assert np.max(np.abs(curr - 1.0)) < 1e-14


def _nyquist_errors(n, diss=0.05):
    """err after iteration iter (0-based) of the mode A (-1)^ix, A_0 = A_-1 = 1."""
    dx, c = 1.0, 1.0
    dt = 0.5 * dx / (c * np.sqrt(3.0))
    q = diss * c * dt * 0.5
    w = 1 + q
    k = c * c * dt * dt / (dx * dx)
    k_main, k_actual, k_old = k / w, 2 * (1 - 3 * k) / w, (q - 1) / w
    a_old, a = 1.0, 1.0
    amps, errs = [], []
    for _ in range(n):
        lap_sum = (-a) + (-a) + a + a + a + a  # x neighbours -A, y and z neighbours +A
        a_old, a = a, lap_sum * k_main + a * k_actual + a_old * k_old
        amps.append(a)
        errs.append(abs(a - a_old) / (abs(a) + 1e-30))
    return np.array(amps), np.array(errs)


def _stop_rule(errs, tol, check_step, hold, max_iter, guard=True, reset=True, before_rotation=False):
    hold_count = 0
    for it in range(max_iter):
        if (it or not guard) and it % check_step == 0:
            e = (errs[it - 1] if it else 0.0) if before_rotation else errs[it]
            if e < tol:
                hold_count += 1
            elif reset:
                hold_count = 0
            if hold_count >= hold:
                return it + 1
    return max_iter


@pytest.mark.parametrize("check_step,hold,tol", [(4, 2, 0.5), (3, 2, 1.0)])
def test_early_exit_cadence_and_hold_reset(orc, check_step, hold, tol):
    nx, max_iter = 4, 300
    amps, errs = _nyquist_errors(max_iter)
    want = _stop_rule(errs, tol, check_step, hold, max_iter)
    # the case discriminates the plausible slips
    assert want < max_iter
    assert _stop_rule(errs, tol, check_step, hold, max_iter, guard=False) != want
    assert _stop_rule(errs, tol, check_step, hold, max_iter, reset=False) != want
    assert _stop_rule(errs, tol, check_step, hold, max_iter, before_rotation=True) != want
    checks = errs[[i for i in range(1, want) if i % check_step == 0]]
    assert np.min(np.abs(checks - tol)) > 0.05  # no decision sits near the tolerance
    sign = np.tile(((-1.0) ** np.arange(nx))[None, None, :], (nx, nx, 1)).ravel()
    prev, curr, it = _solve(orc, nx, sign, sign, max_iter, tol, check_step, hold)
    assert it == want
    assert np.array_equal(curr, amps[it - 1] * sign)
    assert np.array_equal(prev, amps[it - 2] * sign)


def test_no_exit_at_zero_tolerance(orc):
    nx = 4
    sign = np.tile(((-1.0) ** np.arange(nx))[None, None, :], (nx, nx, 1)).ravel()
    amps, _ = _nyquist_errors(57)
    _, curr, it = _solve(orc, nx, sign, sign, 57, 0.0, 1, 1)
    assert it == 57 and np.array_equal(curr, amps[-1] * sign)
