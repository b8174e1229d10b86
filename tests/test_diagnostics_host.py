"""Host side of the diagnostics C ABI, on the CPU (no device call): struct
layouts against the C headers, status names and classes, the input check,
JSON, the shock state machine against the oracle's restatement, and the
natural-gradient tuner's parameter maps."""
import ctypes
import json
import os
import subprocess

import numpy as np

from oracle import diagnostics as od
from svpf_amd import _lib
from svpf_amd import diagnostics as dg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_layout(tmp_path, header, struct, fields):
    prog = tmp_path / f"{struct}.c"
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{header}"', "int main(void) {",
             f'printf("size %zu\\n", sizeof({struct}));']
    lines += [f'printf("{f} %zu\\n", offsetof({struct}, {f}));' for f in fields]
    prog.write_text("\n".join(lines + ["return 0; }"]))
    exe = tmp_path / struct
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(prog), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    return {k: int(v) for k, v in (ln.split() for ln in out if ln)}


def test_struct_layouts_match_headers(tmp_path):
    for header, cls, name in (("svpf_diagnostics.h", dg.SVPFDiagnostics, "SVPFDiagnostics"),
                              ("svpf.h", dg.SVPFGradientDiagnostics, "SVPFGradientDiagnostics"),
                              ("svpf.h", dg.SVPFNaturalGradientTuner, "SVPFNaturalGradientTuner")):
        fields = [f for f, _ in cls._fields_]
        got = _c_layout(tmp_path, header, name, fields)
        assert got["size"] == ctypes.sizeof(cls), name
        for f in fields:
            assert got[f] == getattr(cls, f).offset, (name, f)


def test_both_headers_compile_together(tmp_path):
    prog = tmp_path / "both.c"
    prog.write_text('#include "svpf.h"\n#include "svpf_diagnostics.h"\n'
                    "int main(void) { return SVPF_OK + SVPF_RC_OK + (SVPF_ERR_NOT_INITIALIZED == 400 ? 0 : 1); }\n")
    exe = tmp_path / "both"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(prog), "-o", str(exe), "-lm"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_status_names_and_classes():
    L = dg.bind()
    names = {"SVPF_OK": "OK", "SVPF_ERR_CUDA_KERNEL": "CUDA_KERNEL_FAILED", "SVPF_ERR_NAN_PARTICLES":
             "NAN_IN_PARTICLES", "SVPF_ERR_INF_PARTICLES": "INF_IN_PARTICLES", "SVPF_ERR_WEIGHT_UNDERFLOW":
             "WEIGHT_UNDERFLOW", "SVPF_ERR_INPUT_EXTREME": "INPUT_EXTREME_VALUE", "SVPF_ERR_NOT_INITIALIZED":
             "NOT_INITIALIZED", "SVPF_WARN_ESS_LOW": "ESS_LOW", "SVPF_WARN_VOL_EXTREME": "VOL_EXTREME"}
    for k, v in names.items():
        assert L.svpf_status_name(dg.STATUS[k]).decode() == v
    seen = {L.svpf_status_name(c).decode() for c in dg.STATUS.values()}
    assert len(seen) == len(dg.STATUS) and "UNKNOWN_STATUS" not in seen
    assert L.svpf_status_name(7).decode() == "UNKNOWN_STATUS"
    for k, c in dg.STATUS.items():
        assert L.svpf_status_is_error(c) == ("_ERR_" in k)
        assert L.svpf_status_is_warning(c) == ("_WARN_" in k)


def test_init_input_check_json_and_null_state():
    h = dg.Health()
    d = h.d
    assert (d.status, d.ess_ratio, d.steps_processed) == (0, 1.0, 0)
    assert np.allclose([d.ess_warning_threshold, d.vol_max_threshold, d.vol_min_threshold, d.input_max_return],
                       [0.3, 10.0, 1e-4, 0.5])
    for y in (0.01, -0.5, float("nan"), float("inf"), -1e31, 0.51):
        assert h.check_input(y) == od.check_input(y), y
    assert d.error_count == 4 and d.status == dg.STATUS["SVPF_ERR_INPUT_EXTREME"]
    assert h.message() == "Input return extreme: 0.5100 (threshold: 0.5000)"
    js = json.loads(h.to_json())
    assert js["status"] == "INPUT_EXTREME_VALUE" and js["status_code"] == 302 and js["errors"] == 4
    assert js["message"] == h.message()
    h.L.svpf_diag_clear_error(ctypes.byref(d))
    assert d.status == 0 and h.message() == ""
    assert json.loads(h.to_json())["message"] == ""
    assert h.check() == "NOT_INITIALIZED" and h.message() == "SVPF state is NULL"
    assert d.steps_processed == 0 and d.error_count == 5


def test_shock_machine_matches_oracle():
    L = dg.bind()
    g = dg.SVPFGradientDiagnostics()
    g.shock_threshold, g.shock_duration, g.recovery_duration, g.recovery_exit_threshold = 9.0, 20, 50, 4.0
    g.initialized = True
    o = od.ShockState()
    rng = np.random.default_rng(4)
    z2 = rng.chisquare(1, 2000) * np.where(rng.random(2000) < 0.02, 12.0, 1.0)
    for z in z2.astype(np.float32):
        L.svpf_update_shock_state(ctypes.byref(g), float(z))
        assert L.svpf_get_shock_state(ctypes.byref(g)) == o.update(float(z))
        assert g.ticks_in_state == o.ticks
        assert L.svpf_should_learn(ctypes.byref(g)) == (o.state != 1)
        assert L.svpf_get_lr_multiplier(ctypes.byref(g), 2.0) == o.lr_multiplier(2.0)
    assert not L.svpf_should_learn(None) and L.svpf_get_lr_multiplier(None, 2.0) == 1.0


def test_tuner_maps_parameters():
    L = dg.bind()
    p = _lib.SVPFParams(0.97, 0.15, -1.0, 0.2)
    t = L.svpf_tuner_create(ctypes.byref(p), 7.0, 0.01, 0.001)
    assert t
    tc = t.contents
    assert abs(tc.theta.eta - np.arctanh(0.97)) < 1e-5 and abs(tc.theta.kappa - np.log(0.15)) < 1e-6
    assert abs(tc.theta.kappa_nu - np.log(5.0)) < 1e-6 and not tc.learning_enabled and tc.warmup_ticks == 100
    assert list(tc.F) == [1.0 if i % 5 == 0 else 0.0 for i in range(16)]
    q, nu = _lib.SVPFParams(), ctypes.c_float()
    L.svpf_tuner_update(t, None, ctypes.byref(q), ctypes.byref(nu))
    assert np.allclose([q.rho, q.sigma_z, q.mu, q.gamma, nu.value], [0.97, 0.15, -1.0, 0.0, 7.0], atol=1e-5)
    L.svpf_tuner_destroy(t)


def test_multiscale_layouts_exports_and_defaults(tmp_path):
    from svpf_amd import multiscale as ms

    for cls, name in ((ms.MSConfig, "MS_SVPF_Config"), (ms.MSOutput, "MS_SVPF_Output"), (ms.MSState, "MS_SVPF")):
        fields = [f for f, _ in cls._fields_]
        got = _c_layout(tmp_path, "svpf_multiscale.h", name, fields)
        assert got["size"] == ctypes.sizeof(cls), name
        for f in fields:
            assert got[f] == getattr(cls, f).offset, (name, f)
    L = ms.bind()
    for name in ms.EXPORTED:
        assert hasattr(L, name), name
    c = L.ms_svpf_default_config()
    assert (c.n_particles_reactive, c.n_particles_inertial, c.n_stein_reactive, c.n_stein_inertial) == (256, 1024, 3, 8)
    assert np.allclose([c.surprise_mild_threshold, c.surprise_high_threshold, c.boost_mild_factor,
                        c.boost_high_factor, c.weight_reactive, c.weight_inertial], [2.0, 3.5, 1.5, 2.5, 0.1, 0.9])
    o = ms.MSOutput(vol_reactive=0.03, vol_inertial=0.01)
    assert abs(L.ms_svpf_vol_for_execution(ctypes.byref(o)) - (0.7 * 0.03 + 0.3 * 0.01)) < 1e-8
    assert abs(L.ms_svpf_vol_for_position(ctypes.byref(o)) - (0.2 * 0.03 + 0.8 * 0.01)) < 1e-8
    assert abs(L.ms_svpf_vol_conservative(ctypes.byref(o)) - 0.03) < 1e-8
