"""The process-level drop-in: `rrgalaxsim_amd/rrgsim <work_dir>` with a
reference-format params.json and ini file (rrgsim.cu:12-62,
rrgsim_integration.cu:19-145, io_particles_bin.cu:9-37, io_handler.cpp:71-105).

CPU tests: usage and error exits.  GPU tests: a PM run and a direct-sum run
whose dump files are compared with the oracle's run over the same number of
steps (bitwise for the PM model, 1e-12 for the FP64 direct sum), and whose
data tables and log follow the reference's layout; at every save of the PM
run data/sample_particle.csv and the grid axis cuts (data/a_grid/,
grav_grid/, mass_grid/; rrgsim_conservation.cu:334-424) hold, as text, the
oracle's particle 0 and grids at that save, and check_bounds
(rrgsim_conservation.cu:280-332) logs the particles' extent and the axes
that leave the domain.
"""
import glob
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RRGSIM = os.path.join(ROOT, "rrgalaxsim_amd", "rrgsim")
GOLDEN = os.path.join(ROOT, "tests", "golden")
C = 4574.337022617616


def _work_dir(tmp_path, params):
    (tmp_path / "ini").mkdir()
    shutil.copy(os.path.join(GOLDEN, "S_head256.txt"), tmp_path / "ini" / "S.txt")
    with open(tmp_path / "params.json", "w") as fh:
        json.dump(params, fh, indent=2)
    return tmp_path


def _params(wave=True, real="fp64", tolerance=0.0):
    p = {
        "ini_params": {"mass_star": 1.0, "mass_dark": 1.0, "soft_star": 0.05, "soft_dark": 0.05,
                       "ini_file_star": "ini\\S.txt", "ini_file_dark": None},
        "sim_params": {"time_max": 0.05, "dt_save": 0.02, "dt_dynamics": 0.01, "saves_without_dump": 10,
                       "acceleration_model": "Wave"},
        "mi355x": {"real": real},
    }
    if wave:
        p["grid_params"] = {"nx": 16, "dx": 1.0, "bc_frac": 0.2}
        p["wave_params"] = {"diss_base": 0.785, "diss_extra": 0.01, "wave_speed": C, "setup_iterations": 40,
                            "iterations": 10, "tolerance": tolerance, "print_step": 500, "steady_check_step": 5,
                            "steady_hold_steps": 2}
    return p


def _run(work):
    return subprocess.run([RRGSIM, str(work)], cwd=str(work), capture_output=True, text=True, timeout=120)


def _save_iterations(p, dtype=np.float64):
    """The reference's loop (rrgsim_integration.cu:79-92): time and next_save
    accumulate in the run's `real`.  Returns (save iterations, step count)."""
    s = p["sim_params"]
    r = np.dtype(dtype).type
    t = r(0.0)
    nxt = t + r(s["dt_save"])
    it, out = 0, []
    while t < r(s["time_max"]):
        it += 1
        t = r(t + r(s["dt_dynamics"]))
        if t >= nxt:
            nxt = r(nxt + r(s["dt_save"]))
            out.append((it, float(t)))
    return out, it


def _read_dump(path, dtype):
    with open(path, "rb") as fh:
        n = int(np.frombuffer(fh.read(8), dtype=np.uint64)[0])
        t = float(np.frombuffer(fh.read(8), dtype=np.float64)[0])
        pos = np.frombuffer(fh.read(3 * n * np.dtype(dtype).itemsize), dtype=dtype).reshape(n, 3)
        vel = np.frombuffer(fh.read(3 * n * np.dtype(dtype).itemsize), dtype=dtype).reshape(n, 3)
    return n, t, pos, vel


def _ini(rrg):
    ini = rrg._lib.IniParams(mass_star=1.0, mass_dark=1.0, soft_star=0.05, soft_dark=0.05,
                             ini_file_star=b"S_head256.txt", ini_file_dark=b"")
    return rrg.read_ini(ini, GOLDEN)


def test_cli_usage_exit_code():
    r = subprocess.run([RRGSIM], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def test_cli_bad_params_exits_2_and_logs(tmp_path):
    (tmp_path / "params.json").write_text("{ not json")
    r = _run(tmp_path)
    assert r.returncode == 2
    logs = glob.glob(str(tmp_path / "logs" / "rrgsim_*.log"))
    assert len(logs) == 1 and "rrg_params_parse_json" in open(logs[0]).read()


def test_cli_missing_ini_exits_2(tmp_path):
    p = _params()
    p["ini_params"]["ini_file_star"] = "ini/none.txt"
    (tmp_path / "ini").mkdir()
    (tmp_path / "params.json").write_text(json.dumps(p))
    r = _run(tmp_path)
    assert r.returncode == 2


@pytest.mark.gpu
@pytest.mark.parametrize("real,tol", [("fp64", 0.0), ("fp32", 0.0), ("fp64", 1e-6)])
def test_cli_pm_run_dumps_match_oracle(rrg, orc, tmp_path, real, tol):
    p = _params(real=real, tolerance=tol)
    work = _work_dir(tmp_path, p)
    r = _run(work)
    assert r.returncode == 0, r.stdout + r.stderr
    log = open(glob.glob(str(work / "logs" / "rrgsim_*.log"))[0]).read()
    dtype = np.float32 if real == "fp32" else np.float64
    saves, steps = _save_iterations(p, dtype)
    assert f"integration completed: {steps} steps" in log
    # data tables: header + one row per save
    for name in ("delta_energy", "delta_momentum", "delta_ang_z", "delta_ang_norm"):
        rows = open(work / "data" / f"{name}.csv").read().strip().splitlines()
        assert len(rows) == 1 + len(saves)
    assert open(work / "data" / "delta_energy.csv").readline().strip() == "t, E, E_rel_signed, E_rel, Ek, Ep"
    pos0, vel0, m, _ = _ini(rrg)
    e = 61 - math.frexp(float(np.sum(np.abs(m))))[1]
    g, w = p["grid_params"], p["wave_params"]
    sample_rows = []
    for it, t in saves:
        n, tt, pos, vel = _read_dump(work / "dump" / f"S_{it:5d}.bin", dtype)
        assert n == 256 and tt == pytest.approx(t, rel=1e-6)
        o = orc.pm_run(pos0, vel0, m, it, nx=g["nx"], dx=g["dx"], bc_frac=g["bc_frac"], dt=p["sim_params"]["dt_dynamics"],
                       diss_base=w["diss_base"], wave_speed=C, setup_iterations=w["setup_iterations"],
                       iterations=w["iterations"], tolerance=w["tolerance"], check_step=w["steady_check_step"],
                       hold=w["steady_hold_steps"], fixed_e=e, dtype=dtype)
        assert np.array_equal(pos, o["pos"]) and np.array_equal(vel, o["vel"]), (it, real)
        mass = orc.fixed_to_real(orc.cic_fixed(o["pos"], m.astype(dtype), g["nx"], g["dx"], e, bc_frac=g["bc_frac"],
                                               dtype=dtype), e, dtype=dtype)
        _check_cuts(work, g, dtype, t, o["phi"], o["grid_acc"], mass)
        sample_rows.append(_row([0.0, *o["pos"][0], *o["vel"][0], o["grav"][0], *o["acc"][0]]))
    got = open(work / "data" / "sample_particle.csv").read().splitlines()
    assert got[0] == "i, x, y, z, vx, vy, vz, grav, ax, ay, az" and got[1:] == sample_rows


def _row(vals):
    return ", ".join("%.8f" % float(v) for v in vals)


def _check_cuts(work, g, dtype, t, phi, gacc, mass):
    """The three axis cuts at save time t: coordinate domain_min + dx * i in
    the run's precision, values at AT(i, h, h) (and the y and z cuts)."""
    r = np.dtype(dtype).type
    nx, h = g["nx"], g["nx"] // 2
    dx = r(g["dx"])
    dmin = r(-0.5 * float(dx * r(nx)))
    coord = [float(r(dmin + dx * r(i))) for i in range(nx)]
    at = lambda x, y, z: x + nx * (y + nx * z)  # noqa: E731
    want = {
        ("a_grid", "a_grid_center_%.8f" % t, "x, ax, ay, az"):
            [_row([coord[i], *gacc[at(i, h, h)]]) for i in range(nx)],
        ("grav_grid", "grav_axis_center_%.10f" % t, "r, gx, gy, gz"):
            [_row([coord[i], phi[at(i, h, h)], phi[at(h, i, h)], phi[at(h, h, i)]]) for i in range(nx)],
        ("mass_grid", "mass_axis_center_%.10f" % t, "r, mx, my, mz"):
            [_row([coord[i], mass[at(i, h, h)], mass[at(h, i, h)], mass[at(h, h, i)]]) for i in range(nx)],
    }
    for (sub, name, head), rows in want.items():
        got = open(work / "data" / sub / f"{name}.csv").read().splitlines()
        assert got[0] == head and got[1:] == rows, (sub, name)


@pytest.mark.gpu
@pytest.mark.parametrize("real", ["fp64", "fp32"])
def test_cli_check_bounds_messages(rrg, tmp_path, real):
    """A grid narrower than the disk (domain [-0.04, 0.04]): at each save
    the log gives the particles' extent and one error line per axis and
    side that leaves the domain, compared in the run's precision."""
    p = _params(real=real)
    p["grid_params"] = {"nx": 4, "dx": 0.02, "bc_frac": 0.0}
    work = _work_dir(tmp_path, p)
    r = _run(work)
    assert r.returncode == 0, r.stdout + r.stderr
    log = open(glob.glob(str(work / "logs" / "rrgsim_*.log"))[0]).read()
    dtype = np.float32 if real == "fp32" else np.float64
    rep = (lambda v: str(np.float32(v))) if real == "fp32" else (lambda v: repr(float(v)))
    dmax = np.dtype(dtype).type(0.04)
    saves, _ = _save_iterations(p, dtype)
    want = []
    for it, _ in saves:
        _, _, pos, _ = _read_dump(work / "dump" / f"S_{it:5d}.bin", dtype)
        lo, hi = pos.min(axis=0), pos.max(axis=0)
        for c, ax in enumerate("XYZ"):
            assert f"\t {ax}: [{rep(lo[c])} .. {rep(hi[c])}]" in log, (it, ax)
        want += [f"{ax} min bounds violation! Particle {ax}: {rep(lo[c])}; domain_min: -0.04"
                 for c, ax in enumerate("XYZ") if lo[c] < -dmax]
        want += [f"{ax} max bounds violation! Particle {ax}: {rep(hi[c])}; domain_max: 0.04"
                 for c, ax in enumerate("XYZ") if hi[c] > dmax]
    got = [ln.split("] ", 2)[-1] for ln in log.splitlines() if "bounds violation" in ln]
    assert got == want and any(w.startswith("X min") for w in want)
    assert all("[error]" in ln for ln in log.splitlines() if "bounds violation" in ln)
    assert log.count("Check bounds") == len(saves)


@pytest.mark.gpu
def test_cli_direct_sum_run_matches_oracle(rrg, orc, tmp_path):
    p = _params(wave=False)
    p["mi355x"] = {"force": "fp64_exact"}
    work = _work_dir(tmp_path, p)
    r = _run(work)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "direct-sum model" in r.stdout
    pos0, vel0, m, s2 = _ini(rrg)
    for it, _ in _save_iterations(p)[0]:
        _, _, pos, vel = _read_dump(work / "dump" / f"S_{it:5d}.bin", np.float64)
        opos, ovel, _, _ = orc.nbody_run(pos0, vel0, m, s2, p["sim_params"]["dt_dynamics"], it)
        assert np.max(np.abs(pos - opos)) < 1e-12 and np.max(np.abs(vel - ovel)) < 1e-12
