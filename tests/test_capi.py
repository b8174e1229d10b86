"""C ABI checks that need no GPU: the library loads, exports every function
include/svpf.h, include/svpf_fully_fused_v2.h and include/svpf_diagnostics.h declare, and the ctypes
mirrors of SVPFState and svpf_v2_noise match the C layouts field by field."""
import ctypes
import os
import re
import subprocess

import pytest

from svpf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "svpf.h")
HEADERS = [os.path.join(ROOT, "include", f) for f in ("svpf.h", "svpf_fully_fused_v2.h", "svpf_diagnostics.h")]


def declared_functions():
    names = set()
    for hdr in HEADERS:
        src = open(hdr).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?!static)[A-Za-z_][\w\s\*]*?\b(svpf_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTED)
    assert L.svpf_version().startswith(b"svpf-mi355x")


def test_tile_table():
    L = _lib.load()
    caps = [L.svpf_tile_capacity_query(t, 256) for t in range(8)]
    assert caps == [64, 128, 256, 512, 1024, 2048, 256, 512]
    assert L.svpf_tile_capacity_query(99, 256) == 0
    assert [L.svpf_tile_capacity_query(t, 1024) for t in range(8)] == [256, 512, 1024, 2048, 4096, 0, 0, 0]
    assert L.svpf_tile_capacity_query(2, 128) == 0
    assert L.svpf_lds_bytes_query(2048) >= 2048 * 20


def test_state_layout_matches_header(tmp_path):
    prog = tmp_path / "layout.c"
    fields = [f for f, _ in _lib.SVPFState._fields_]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svpf.h"', "int main(void) {",
             'printf("size %zu\\n", sizeof(SVPFState));']
    for f in fields:
        lines.append(f'printf("{f} %zu\\n", offsetof(SVPFState, {f}));')
    lines += ['printf("params %zu\\n", sizeof(SVPFParams));', 'printf("result %zu\\n", sizeof(SVPFResult));',
              "return 0; }"]
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(prog), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    assert int(got["size"]) == ctypes.sizeof(_lib.SVPFState)
    for f in fields:
        assert int(got[f]) == getattr(_lib.SVPFState, f).offset, f
    assert int(got["params"]) == ctypes.sizeof(_lib.SVPFParams)
    assert int(got["result"]) == ctypes.sizeof(_lib.SVPFResult)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _lib.load(str(tmp_path / "nope.so"))


def test_record_layouts_match_c(tmp_path):
    """numpy record dtypes of the batched API against csrc/svpf_types.h."""
    from svpf_amd import batched, diff, twofactor

    csrc = os.path.join(ROOT, "newton-stein-variational-particle-filter_amd", "csrc")
    checks = [("svpf_fparams", batched.FPARAMS_DTYPE), ("svpf_fstate", batched.FSTATE_DTYPE),
              ("svpf_fdesc", batched.FDESC_DTYPE), ("svpf_dconfig", diff.DCONFIG_DTYPE),
              ("svpf_ddesc", diff.DDESC_DTYPE), ("svpf_d2config", twofactor.D2CONFIG_DTYPE),
              ("svpf_d2desc", twofactor.D2DESC_DTYPE)]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svpf_types.h"', "int main(void) {"]
    for cname, dt in checks:
        lines.append(f'printf("{cname}.size %zu\\n", sizeof({cname}));')
        for f in dt.names:
            if f.startswith("pad"):
                continue
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('printf("flags %u %u %u\\n", (unsigned)SVPF_F_EXACT_GRAD, (unsigned)SVPF_F_KSD_EARLY_STOP,'
                 ' (unsigned)SVPF_F_PAPER_SIGN);')
    lines.append("return 0; }")
    prog = tmp_path / "rec.c"
    prog.write_text("\n".join(lines))
    exe = tmp_path / "rec"
    subprocess.run(["gcc", "-I", csrc, str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.strip().split("\n")
    got = {}
    for line in out:
        k, *v = line.split()
        got[k] = v
    for cname, dt in checks:
        assert int(got[f"{cname}.size"][0]) == dt.itemsize
        for f in dt.names:
            if not f.startswith("pad"):
                assert int(got[f"{cname}.{f}"][0]) == dt.fields[f][1], (cname, f)
    assert [int(x) for x in got["flags"]] == [batched.F_EXACT_GRAD, batched.F_KSD_EARLY_STOP, batched.F_PAPER_SIGN]


def test_pysvpf_surface_matches_reference_module():
    """Method and property names of pysvpf.SVPF (py-torch/svpf_python.cpp:483-654)."""
    from svpf_amd import pysvpf

    names = ["initialize", "step", "run", "reset", "set_adaptive_mu", "set_adaptive_sigma", "set_adaptive_guide",
             "set_asymmetric_rho", "set_mim", "set_svld", "set_annealing", "set_newton", "set_guide", "set_params",
             "set_adaptive_beta", "set_full_newton", "set_exact_gradient", "set_stein_steps",
             "set_guided_innovation", "set_rejuvenation", "get_particles", "get_ess", "invalidate_graph",
             "is_graph_captured", "nu", "lik_offset", "n_particles", "n_stein_steps"]
    for k in names:
        assert hasattr(pysvpf.SVPF, k), k
    r = pysvpf.SVPFResult(1.0, 2.0, 0.0, 3.0, 4.0)
    assert (r.log_lik, r.vol_mean, r.h_mean, r.mu_estimate) == (1.0, 2.0, 3.0, 4.0)


def test_v2_noise_layout_matches_header(tmp_path):
    prog = tmp_path / "v2.c"
    fields = [f for f, _ in _lib.SVPFV2Noise._fields_]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svpf_fully_fused_v2.h"', "int main(void) {",
             'printf("size %zu\\n", sizeof(svpf_v2_noise));']
    lines += [f'printf("{f} %zu\\n", offsetof(svpf_v2_noise, {f}));' for f in fields]
    lines.append("return 0; }")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "v2"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line.strip())
    assert int(got["size"]) == ctypes.sizeof(_lib.SVPFV2Noise)
    for f in fields:
        assert int(got[f]) == getattr(_lib.SVPFV2Noise, f).offset, f
    L = _lib.load()
    assert L.svpf_fully_fused_v2_supported(1024) and not L.svpf_fully_fused_v2_supported(1025)
    assert not L.svpf_fully_fused_v2_supported(0)


def test_digamma_is_bounded_and_accurate():
    """The adjoint and score kernels' digamma (compiled for the host from the same
    source): accurate for x > 0, NaN -- not an endless recurrence -- for a
    non-positive or non-finite argument such as a diverged degrees of freedom."""
    import math

    import scipy.special

    L = _lib.load()
    for x in (1e-3, 0.5, 1.0, 2.5, 5.0, 6.0, 17.3, 1e4):
        assert L.svpf_digamma_f32(x) == pytest.approx(scipy.special.digamma(x), rel=3e-6, abs=3e-6), x
    for x in (0.0, -1.0, -0.5, -3e7, float("-inf"), float("inf"), float("nan")):
        assert math.isnan(L.svpf_digamma_f32(x)), x


def test_import_pysvpf_by_the_reference_module_name():
    import pysvpf
    from svpf_amd import pysvpf as impl

    assert pysvpf.SVPF is impl.SVPF and pysvpf.SVPFResult is impl.SVPFResult
    for name in ("initialize", "step", "run", "reset", "set_params", "set_mim", "set_newton", "get_particles",
                 "get_ess", "invalidate_graph", "is_graph_captured"):
        assert hasattr(pysvpf.SVPF, name), name


def test_destructors_skip_destroy_during_interpreter_shutdown(monkeypatch):
    """During shutdown the HIP runtime may already be finalising: destructors leave
    the state to the process exit instead of calling svpf_destroy; explicit close()
    and destructors before shutdown destroy as usual."""
    from svpf_amd import filter as flt

    calls = []

    class FakeLib:
        def svpf_destroy(self, sp):
            calls.append(sp)

    f = flt.Filter.__new__(flt.Filter)
    f.L, f.sp = FakeLib(), 1234
    monkeypatch.setattr(_lib, "_exiting", True)
    assert _lib.finalizing()
    f.__del__()
    assert calls == [] and f.sp == 1234
    monkeypatch.setattr(_lib, "_exiting", False)
    f.__del__()
    assert calls == [1234] and f.sp is None
