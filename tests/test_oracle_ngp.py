"""Pins the oracle's NGP binning and cell projections (oracle/rrg_oracle.cpp
ngp_bin / ngp_project; reference rm_p2mesh.cu:25-168) against numpy: the cell
of each particle (truncated (x - min)/dx, clamped), counts by bincount, the
exclusive prefix by cumsum, the order by a stable argsort, and the cell
averages by their defining sums.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics


def _case(n=4000, seed=3):
    pos, _, m = ics.exp_disk(n, seed=seed)
    rng = np.random.default_rng(seed)
    pos[:50] = rng.uniform(-12, 12, (50, 3))  # outside the [-8, 8] box: clamped
    return pos


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ngp_bin_is_the_counting_sort(orc, dtype):
    nx, dx = 16, 1.0
    pos = _case().astype(dtype)
    cell_of, count, start, order = orc.ngp_bin(pos, nx, dx, dtype=dtype)
    idx = np.clip(((pos - dtype(-8.0)) / dtype(dx)).astype(np.int64), 0, nx - 1)  # C cast truncates
    ref_cell = idx[:, 0] + nx * (idx[:, 1] + nx * idx[:, 2])
    assert np.array_equal(cell_of, ref_cell)
    assert np.array_equal(count, np.bincount(ref_cell, minlength=nx ** 3))
    assert np.array_equal(start, np.concatenate([[0], np.cumsum(count)[:-1]]))
    assert np.array_equal(order, np.argsort(ref_cell, kind="stable"))


def test_ngp_project_averages(orc):
    nx, dx = 8, 2.0
    pos = _case(3000, 5)
    rng = np.random.default_rng(5)
    g = rng.standard_normal(len(pos))
    a = rng.standard_normal((len(pos), 3))
    cell_of, count, start, order = orc.ngp_bin(pos, nx, dx)
    phi, acc = orc.ngp_project(nx, count, start, order, g, a)
    assert np.all(phi[count == 0] == 0) and np.all(acc[count == 0] == 0)
    occ = count > 0
    ref_phi = np.bincount(cell_of, weights=g, minlength=nx ** 3)[occ] / count[occ]
    assert np.allclose(phi[occ], ref_phi, rtol=1e-12, atol=1e-14)
    for c in range(3):
        ref = np.bincount(cell_of, weights=a[:, c], minlength=nx ** 3)[occ] / count[occ]
        assert np.allclose(acc[occ, c], ref, rtol=1e-12, atol=1e-14)
    ones, _ = orc.ngp_project(nx, count, start, order, np.ones(len(pos)), a)
    assert np.allclose(ones[occ], 1.0, rtol=1e-14)
