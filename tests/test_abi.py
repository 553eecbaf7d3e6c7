"""C ABI surface of librrg.so, checked without a GPU: the library loads, it
exports every function include/rrg.h declares, the params.json / ini readers
follow the reference's semantics (io_json.cpp:13-242,
io_particles_csv.cpp:11-134), and errors come back as status codes."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    text = open(os.path.join(ROOT, "include", "rrg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rrg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(rrg):
    names = header_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(rrg._lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(rrg._lib.SIGNATURES), set(names) ^ set(rrg._lib.SIGNATURES)


def test_version_and_device_count(rrg):
    assert b"gfx950" in rrg.lib().rrg_version()
    assert rrg.lib().rrg_device_count() >= 0


def test_defaults_follow_io_json(rrg):
    p = rrg.default_params()
    assert p.wave.setup_iterations == 30000 and p.wave.iterations == 3000
    assert p.wave.tolerance == 1e-6 and p.wave.print_step == 500
    assert p.wave.steady_check_step == 30 and p.wave.steady_hold_steps == 3
    assert p.sim.saves_without_dump == 2 ** 64 - 1
    assert p.has_grid == 0 and p.has_wave == 0


def test_parse_sample_params(rrg):
    p, ini = rrg.parse_params_json(os.path.join(GOLDEN, "sample_params.json"))
    assert p.use_wave_model == 1
    assert p.grid.nx == 512 and p.grid.dx == 0.025 and p.grid.bc_frac == 0.2
    assert p.wave.diss_base == 0.1 and p.wave.wave_speed == 4574.337022617616
    # CFL step (io_json.cpp:174-178)
    assert p.wave.dt == 0.5 * 0.025 / (4574.337022617616 * np.sqrt(3.0))
    assert p.sim.dt_dynamics == 0.002 and p.sim.saves_without_dump == 10
    assert ini.mass_star == 1.0 and ini.soft_star == 0.004
    assert ini.ini_file_star == b"ini\\S.txt" and ini.ini_file_dark == b""


def test_parse_fp32_cfl_step(rrg, tmp_path):
    d = json.load(open(os.path.join(GOLDEN, "sample_params.json")))
    d["mi355x"] = {"real": "fp32", "force": "mixed", "boundary": "zero_dirichlet", "fixed_exponent": 50}
    f = tmp_path / "params.json"
    f.write_text(json.dumps(d))
    p, _ = rrg.parse_params_json(str(f))
    dx, c = np.float32(0.025), np.float32(4574.337022617616)
    expect = np.float32(0.5 * np.float64(dx) / np.float64(c * np.sqrt(np.float32(3))))
    assert p.wave.dt == float(expect)
    assert p.real_type == 1 and p.force_kind == 1 and p.boundary == 0 and p.fixed_exponent == 50


def test_parse_launch_options(rrg, tmp_path):
    d = json.load(open(os.path.join(GOLDEN, "sample_params.json")))
    p0, _ = rrg.parse_params_json(os.path.join(GOLDEN, "sample_params.json"))
    assert p0.use_graphs == 1 and p0.autotune == 0 and p0.wave_fused == 1 and p0.sort_interval == 0
    d["mi355x"] = {"graphs": False, "autotune": True, "wave_fused": False, "sort_interval": 12}
    f = tmp_path / "params.json"
    f.write_text(json.dumps(d))
    p, _ = rrg.parse_params_json(str(f))
    assert p.use_graphs == 0 and p.autotune == 1 and p.wave_fused == 0 and p.sort_interval == 12
    for bad in ({"wave_fused": 1}, {"sort_interval": -1}, {"sort_interval": 2.5}, {"sort_interval": "8"}):
        d["mi355x"] = bad
        f.write_text(json.dumps(d))
        with pytest.raises(rrg.RRGError):
            rrg.parse_params_json(str(f))


def test_nbody_model_when_wave_section_absent(rrg, tmp_path):
    d = json.load(open(os.path.join(GOLDEN, "sample_params.json")))
    del d["wave_params"]
    f = tmp_path / "params.json"
    f.write_text(json.dumps(d))
    p, _ = rrg.parse_params_json(str(f))
    assert p.use_wave_model == 0 and p.has_grid == 1


@pytest.mark.parametrize("drop", [("sim_params", None), ("ini_params", None), ("sim_params", "dt_save"),
                                  ("grid_params", "bc_frac"), ("wave_params", "wave_speed")])
def test_missing_required_key_is_io_error(rrg, tmp_path, drop):
    d = json.load(open(os.path.join(GOLDEN, "sample_params.json")))
    sec, key = drop
    if key is None:
        del d[sec]
    else:
        del d[sec][key]
    f = tmp_path / "params.json"
    f.write_text(json.dumps(d))
    with pytest.raises(rrg.RRGError) as ei:
        rrg.parse_params_json(str(f))
    assert ei.value.code == rrg._lib.RRG_ERR_IO


def test_malformed_and_missing_json(rrg, tmp_path):
    f = tmp_path / "params.json"
    f.write_text("{ \"sim_params\": ")
    with pytest.raises(rrg.RRGError):
        rrg.parse_params_json(str(f))
    with pytest.raises(rrg.RRGError) as ei:
        rrg.parse_params_json(str(tmp_path / "nope.json"))
    assert "Expected params json" in str(ei.value)


def test_ini_reader_masses_and_softening(rrg, tmp_path):
    """m = M_component / N_cumulative, soft2 = soft^2 (io_particles_csv.cpp:46-47)."""
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 6)); b = rng.standard_normal((3, 6))
    (tmp_path / "ini").mkdir()
    for name, rows in (("S.txt", a), ("D.txt", b)):
        with open(tmp_path / "ini" / name, "w") as fh:
            fh.write(f"{len(rows)} 0.000000\n")
            np.savetxt(fh, rows, fmt="%.17g")
    ini = rrg._lib.IniParams(mass_star=2.0, mass_dark=8.0, soft_star=0.5, soft_dark=0.25,
                             ini_file_star=b"ini\\S.txt", ini_file_dark=b"ini/D.txt")
    pos, vel, m, s2 = rrg.read_ini(ini, str(tmp_path))
    assert np.array_equal(pos, np.vstack([a[:, :3], b[:, :3]]))
    assert np.array_equal(vel, np.vstack([a[:, 3:], b[:, 3:]]))
    assert np.array_equal(m, [2.0 / 5] * 5 + [8.0 / 8] * 3)
    assert np.array_equal(s2, [0.25] * 5 + [0.0625] * 3)


def test_ini_reader_golden_sample_rows(rrg):
    """First 256 rows of the reference's sample/ini/S.txt (header rewritten to 256)."""
    ini = rrg._lib.IniParams(mass_star=1.0, mass_dark=1.0, soft_star=0.004, soft_dark=0.004,
                             ini_file_star=b"S_head256.txt", ini_file_dark=b"")
    pos, vel, m, s2 = rrg.read_ini(ini, GOLDEN)
    assert pos.shape == (256, 3)
    assert np.array_equal(pos[0], [0.000035, 0.000047, 0.001940])
    assert np.array_equal(vel[1], [0.709329, -0.256181, 0.276377])
    assert np.all(m == 1.0 / 256) and np.all(s2 == 0.004 ** 2)


def test_ini_row_count_mismatch(rrg, tmp_path):
    with open(tmp_path / "S.txt", "w") as fh:
        fh.write("4 0.0\n1 2 3 4 5 6\n")
    ini = rrg._lib.IniParams(mass_star=1.0, mass_dark=1.0, soft_star=0.1, soft_dark=0.1,
                             ini_file_star=b"S.txt", ini_file_dark=b"")
    with pytest.raises(rrg.RRGError) as ei:
        rrg.read_ini(ini, str(tmp_path))
    assert "mismatch" in str(ei.value)


def test_bad_arguments_are_rejected_before_device_work(rrg):
    p = rrg.make_params(dt=1e-3)
    with pytest.raises(rrg.RRGError) as ei:
        rrg.Simulation(p, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    assert ei.value.code == rrg._lib.RRG_ERR_INVALID
    with pytest.raises(rrg.RRGError):
        rrg.Simulation(p, np.full((2, 3), np.nan), np.zeros((2, 3)), np.ones(2), np.ones(2))
    assert rrg.lib().rrg_destroy(None) == 0


def test_ensemble_offsets_validated_before_device_work(rrg):
    p = rrg.make_params(dt=1e-3)
    pos = np.zeros((6, 3)); vel = np.zeros((6, 3)); m = np.ones(6); s2 = np.ones(6)
    for off in ([1, 3, 6], [0, 3, 3, 6], [0, 4, 2, 6]):
        with pytest.raises(rrg.RRGError) as ei:
            rrg.Simulation(p, pos, vel, m, s2, offsets=off)
        assert ei.value.code == rrg._lib.RRG_ERR_INVALID


@pytest.mark.parametrize("real,nx,boundary,iters,fused,expect", [
    ("fp32", 128, "periodic", 200, True, {1: 0, 2: 0, 3: 0, 4: 50}),
    ("fp64", 128, "periodic", 9, True, {1: 0, 2: 1, 3: 1, 4: 1}),
    ("fp32", 256, "periodic", 7, True, {1: 0, 2: 0, 3: 1, 4: 1}),
    ("fp32", 256, "periodic", 6, True, {1: 0, 2: 0, 3: 2, 4: 0}),
    ("fp32", 256, "periodic", 5, True, {1: 0, 2: 1, 3: 1, 4: 0}),
    ("fp64", 256, "periodic", 200, True, {1: 0, 2: 1, 3: 66, 4: 0}),
    ("fp32", 512, "periodic", 200, True, {1: 0, 2: 1, 3: 66, 4: 0}),
    ("fp64", 512, "periodic", 7, True, {1: 1, 2: 3, 3: 0, 4: 0}),
    ("fp64", 64, "periodic", 4, True, {1: 1, 2: 0, 3: 1, 4: 0}),
    ("fp32", 96, "periodic", 10, True, {1: 10, 2: 0, 3: 0, 4: 0}),
    ("fp32", 128, "zero_dirichlet", 10, True, {1: 10, 2: 0, 3: 0, 4: 0}),
    ("fp32", 128, "periodic", 10, False, {1: 10, 2: 0, 3: 0, 4: 0}),
    ("fp32", 128, "periodic", 0, True, {1: 0, 2: 0, 3: 0, 4: 0}),
])
def test_wave_schedule(rrg, real, nx, boundary, iters, fused, expect):
    """The library's split of unchecked stencil iterations into launches of
    four, three, two and one iterations (no device needed)."""
    got = rrg.wave_schedule(nx, iters, real=real, boundary=boundary, fused=fused)
    assert got == expect
    assert sum(k * v for k, v in got.items()) == iters
