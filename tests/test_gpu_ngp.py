"""NGP binning and the direct-sum model's grid projections on the GPU
(reference rm_p2mesh.cu:25-168, 398-485; rrgsim_integration_nbody.cuh:32-48)
against the oracle: the counting sort bitwise, the CIC cell mass bitwise, the
cell averages bitwise when fed the same particle potential and acceleration.
"""
import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu


def _case(n, seed):
    pos, vel, m = ics.exp_disk(n, seed=seed)
    rng = np.random.default_rng(seed)
    pos[:100] = rng.uniform(-12, 12, (100, 3))
    return pos, vel, m


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,nx", [(0, 4), (1, 2), (100_000, 32)])
def test_ngp_bin_matches_oracle(rrg, orc, dtype, n, nx):
    pos = _case(max(n, 200), 4)[0][:n].astype(dtype)
    dx = 16.0 / nx
    got = rrg.ngp_bin(pos, nx, dx, dtype=dtype)
    ref = orc.ngp_bin(pos, nx, dx, dtype=dtype)
    for a, b, name in zip(got, ref, ("cell_of", "count", "start", "order")):
        assert np.array_equal(a, b), name


def test_direct_sum_grid_projections(rrg, orc):
    pos, vel, m = _case(20_000, 6)
    s2 = np.full(len(m), 1e-3)
    p = rrg.make_params(dt=1e-3, grid=dict(nx=16, dx=1.0, bc_frac=0.0))
    assert not p.use_wave_model
    with rrg.Simulation(p, pos, vel, m, s2) as sim:
        sim.precalc()
        sim.calc_step(3)
        sim.sync_particles_and_grid()
        st = sim.state()
        gr = sim.grid()
        e = sim.stats()["fixed_exponent"]
    assert e in (60, 61)
    mfix = orc.cic_fixed(st["pos"], m, 16, 1.0, e)
    assert np.array_equal(gr["mass"], orc.fixed_to_real(mfix, e))
    _, count, start, order = orc.ngp_bin(st["pos"], 16, 1.0)
    phi, acc = orc.ngp_project(16, count, start, order, st["grav"], st["acc"])
    assert np.array_equal(gr["phi"], phi) and np.array_equal(gr["acc"], acc)
    assert np.isclose(np.sum(gr["mass"]), np.sum(m), rtol=1e-14)
