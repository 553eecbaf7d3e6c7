"""The direct-sum autotuner's cache: shapes keyed by GPU architecture, force
kind, softening form and half-octave size buckets, reused within a process
and, through the cache file ($RRG_TUNE_CACHE), by later processes; a reused
shape gives the same launch and the same forces as the tuned one."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import rrgalaxsim_amd as rrg
from rrgalaxsim_amd import ics
sizes = [3000, 5000, 4100, 6000]
rng = np.random.default_rng(3)
parts = [ics.plummer(n, rng=rng) for n in sizes]
pos = np.concatenate([p[0] for p in parts]); vel = np.concatenate([p[1] for p in parts])
m = np.concatenate([p[2] for p in parts])
off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
p = rrg.make_params(dt=1e-3, force="mixed", autotune=True)
with rrg.Simulation(p, pos, vel, m, np.full(len(m), 1e-4), offsets=off) as sim:
    sim.precalc()
    sim.calc_step(2)
    st = sim.stats()
    acc = sim.state()["acc"]
print(json.dumps({"tuned": st["tuned"], "candidates": st["tune_candidates"], "ib": st["tile_ib"],
                  "len": st["slice_len"], "acc_sum": float(np.abs(acc).sum())}))
"""


def _child(cache):
    env = dict(os.environ, RRG_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tune_cache_file_is_reused_by_a_later_process(tmp_path):
    cache = tmp_path / "tune" / "autotune.txt"
    a = _child(cache)
    assert a["tuned"] == 1 and a["candidates"] == 12
    lines = cache.read_text().split("\n")
    assert len([ln for ln in lines if ln.strip()]) == 1 and lines[0].startswith("gfx")
    b = _child(cache)
    assert b["tuned"] == 1 and b["candidates"] == 0
    assert (b["ib"], b["len"]) == (a["ib"], a["len"]) and b["acc_sum"] == a["acc_sum"]


def test_tune_cache_within_a_process_by_size_bucket(rrg):
    """A second ensemble in the same half-octave buckets reuses the shape."""
    out = []
    for sizes in ([2000, 3000, 2500], [2050, 3010, 2480]):
        rng = np.random.default_rng(8)
        parts = [ics.plummer(n, rng=rng) for n in sizes]
        pos = np.concatenate([q[0] for q in parts]); vel = np.concatenate([q[1] for q in parts])
        m = np.concatenate([q[2] for q in parts])
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        p = rrg.make_params(dt=1e-3, force="fp64", autotune=True)
        with rrg.Simulation(p, pos, vel, m, np.full(len(m), 1e-4), offsets=off) as sim:
            sim.precalc()
            out.append(sim.stats())
    assert out[0]["tune_candidates"] == 8 and out[1]["tune_candidates"] == 0
    assert out[1]["tuned"] == 1 and out[1]["tile_ib"] == out[0]["tile_ib"] and out[1]["slice_len"] == out[0]["slice_len"]
