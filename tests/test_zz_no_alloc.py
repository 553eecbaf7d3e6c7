"""rrg_op_direct_accel_noalloc (the allocation-free variant of
rrg_op_direct_accel, scratch in a fixed device area) against
rrg_op_direct_accel, both through the C ABI: bitwise equal for every force
kind, on two calls in a row with the second one smaller (the scratch keeps the
first call's contents), with uniform and per-particle softening; sizes beyond
the scratch area and host pointers are rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

from rrgalaxsim_amd import _lib as L
from rrgalaxsim_amd import ics

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def _noalloc(kind, pos, m, s2):
    n = len(m)
    dp, dm, ds = _dev(pos), _dev(m), _dev(s2)
    dacc = torch.full((n, 3), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.lib().rrg_op_direct_accel_noalloc(L.FORCE_KINDS[kind], n, dp.data_ptr(), dm.data_ptr(), ds.data_ptr(),
                                             dacc.data_ptr(), ctypes.c_void_p(stream))
    L.check(rc)
    torch.cuda.current_stream().synchronize()
    return dacc.cpu().numpy()


@pytest.mark.parametrize("kind", ["fp64", "mixed", "fp32", "fp64_exact"])
@pytest.mark.parametrize("uniform", [True, False])
def test_noalloc_equals_direct_accel(rrg, kind, uniform):
    for n, seed in ((20_000, 1), (3_001, 2)):  # second call smaller, not a multiple of 256
        print(f"noalloc {kind} uniform={uniform} n={n}", flush=True)
        pos, _, m = ics.plummer(n, seed=seed)
        s2 = np.full(n, 1e-4) if uniform else np.random.default_rng(seed).uniform(0.5e-4, 2e-4, n)
        ref = rrg.direct_accel(pos, m, s2, force=kind)
        got = _noalloc(kind, pos, m, s2)
        assert np.array_equal(got, ref), (kind, n)


def test_noalloc_rejects_oversize_and_host_pointers(rrg):
    n = 300_000  # the slice partials alone exceed the fixed area
    dp = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dm = torch.zeros(n, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    f = L.lib().rrg_op_direct_accel_noalloc
    rc = f(L.FORCE_MIXED, n, dp.data_ptr(), dm.data_ptr(), dm.data_ptr(), dp.data_ptr(), ctypes.c_void_p(stream))
    assert rc == L.RRG_ERR_INVALID and b"bytes of scratch" in L.lib().rrg_last_error()
    hp = np.zeros((100, 3))
    dq = torch.zeros((100, 3), dtype=torch.float64, device="cuda")
    dmm = torch.zeros(100, dtype=torch.float64, device="cuda")
    rc = f(L.FORCE_F64, 100, L.ptr(hp), dmm.data_ptr(), dmm.data_ptr(), dq.data_ptr(), ctypes.c_void_p(stream))
    assert rc == L.RRG_ERR_INVALID and b"not device memory" in L.lib().rrg_last_error()
