"""svpf_run_sequence_device_noalloc: the device loop with its per-tick records in
the library's fixed scratch array (no device allocation in the call), against
svpf_run_sequence_device on identical filters, bit for bit: two calls in a row,
the second shorter than the first (the scratch keeps the first call's records),
on each kernel path; and the size limit."""
import ctypes

import numpy as np
import pytest

from svpf_amd import _lib, config, data
from svpf_amd.filter import Filter

pytestmark = pytest.mark.gpu


def pair(cfg, path):
    out = []
    for _ in range(2):
        f = Filter(cfg, path=path)
        f.initialize(seed=2024)
        out.append(f)
    return out


@pytest.mark.parametrize("n,path", [(128, "single"), (512, "coop"), (2048, "grid")])
def test_noalloc_device_loop_equals_the_allocating_one(torch_cuda, n, path):
    torch = torch_cuda
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(400, seed=n)
    a, b = pair(cfg, path)
    assert a.path() == path
    L = a.L
    for T0, T in ((0, 300), (300, 120)):  # the second call is shorter than the first
        print(f"noalloc n={n} path={path} ticks [{T0}, {T0 + T})", flush=True)
        dy = torch.tensor(y[T0:T0 + T], device="cuda")
        lla, vola = torch.zeros(T, device="cuda"), torch.zeros(T, device="cuda")
        llb, volb = torch.full((T,), float("nan"), device="cuda"), torch.full((T,), float("nan"), device="cuda")
        P = ctypes.c_void_p
        L.svpf_run_sequence_device(a.sp, P(dy.data_ptr()), T, ctypes.byref(a.params), P(lla.data_ptr()),
                                   P(vola.data_ptr()))
        _lib.check(a.sp, "svpf_run_sequence_device")
        L.svpf_run_sequence_device_noalloc(b.sp, P(dy.data_ptr()), T, ctypes.byref(b.params), P(llb.data_ptr()),
                                           P(volb.data_ptr()))
        _lib.check(b.sp, "svpf_run_sequence_device_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(lla, llb) and torch.equal(vola, volb)
        assert np.array_equal(a.particles(), b.particles())
        assert np.array_equal(a.last_outputs(), b.last_outputs())
        assert a.state.timestep == b.state.timestep == T0 + T
    a.close()
    b.close()


def test_noalloc_rejects_more_ticks_than_its_scratch(torch_cuda):
    torch = torch_cuda
    f = Filter(config.plain_config(n=64))
    f.initialize(seed=1)
    T = _lib.NOALLOC_MAX_TICKS + 1
    dy = torch.zeros(T, device="cuda")
    f.L.svpf_run_sequence_device_noalloc(f.sp, ctypes.c_void_p(dy.data_ptr()), T, ctypes.byref(f.params), None,
                                         None)
    assert f.L.svpf_last_error(f.sp) == _lib.SVPF_RC_TOO_LARGE
    assert b"scratch" in f.L.svpf_last_error_string(f.sp)
    assert f.state.timestep == 0
    f.close()
