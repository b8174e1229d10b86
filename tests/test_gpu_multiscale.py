"""The two-scale filter (include/svpf_multiscale.h) on the GPU.  Each scale's
outputs must be bit-identical to stepping that scale's filter directly with
the parameters the reference routes to it (include/svpf_multiscale.cuh:
408-532: the scout with its own rho / sigma_z, the production filter with the
caller's, both fed the previous observation), and the combined estimate and
the flags must follow from them as restated here."""
import ctypes

import numpy as np
import pytest

from svpf_amd import _lib, data
from svpf_amd import multiscale as msm

pytestmark = pytest.mark.gpu


def test_configuration_of_the_two_scales(torch_cuda):
    m = msm.TwoScale()
    s = m.state
    r, i = s.reactive.contents, s.inertial.contents
    assert (r.n_particles, i.n_particles) == (256, 1024)
    assert (r.use_mim, r.mim_jump_prob, r.use_guide, r.use_newton, r.use_annealing) == (1, pytest.approx(0.30), 0, 0, 0)
    assert (i.use_full_newton, i.use_adaptive_mu, i.use_adaptive_sigma, i.use_guided) == (1, 1, 1, 1)
    assert (i.stein_min_steps, i.stein_max_steps, r.stein_min_steps, r.stein_max_steps) == (8, 16, 2, 5)
    assert (s.rho_reactive, s.sigma_z_reactive) == (pytest.approx(0.92), pytest.approx(0.18))
    m.close()


def test_scales_match_direct_steps(torch_cuda):
    L = msm.bind()
    # calm series (vol about 1%), then a two-day shock: the scout's vol crosses 2% while the production
    # filter's adaptive mu stays clamped at -4 (vol above 2%), so the classification changes
    y, _ = data.sv_series(60, mu=-9.0, rho=0.98, sigma_z=0.2, nu=5.0, seed=31)
    y[30], y[31] = 0.15, -0.12
    base = _lib.SVPFParams(0.98, 0.10, -9.0, -0.5)
    a, b = msm.TwoScale(), msm.TwoScale()
    a.init(0.01, base, 7)
    b.init(0.01, base, 7)
    sb = b.state
    pr = _lib.SVPFParams(sb.rho_reactive, sb.sigma_z_reactive, base.mu, base.gamma)
    vol_prev, y_prev = 0.01, 0.0
    print("two-scale filter, 60 ticks", flush=True)
    flags = set()
    for t in range(len(y)):
        o = a.step(float(y[t]), base)
        f = ctypes.c_float
        lr, vr, hr, li, vi, hi = f(), f(), f(), f(), f(), f()
        b.L.svpf_step_adaptive(sb.reactive, float(y[t]), y_prev, ctypes.byref(pr), lr, vr, hr)
        b.L.svpf_step_adaptive(sb.inertial, float(y[t]), y_prev, ctypes.byref(base), li, vi, hi)
        _lib.check(sb.reactive, "reactive")
        _lib.check(sb.inertial, "inertial")
        assert (o.vol_reactive, o.h_reactive, o.loglik_reactive) == (vr.value, hr.value, lr.value), t
        assert (o.vol_inertial, o.h_inertial, o.loglik_inertial) == (vi.value, hi.value, li.value), t
        assert o.stein_steps_inertial == sb.inertial.contents.stein_steps_used
        assert o.ksd_reactive == sb.reactive.contents.ksd_prev
        # the combination and the classification
        v_r, v_i = np.float32(vr.value), np.float32(vi.value)
        assert o.vol_combined == v_i and o.h_combined == hi.value and o.loglik_combined == li.value
        assert o.surprise_reactive == pytest.approx(abs(float(y[t])) / max(vol_prev, 1e-6), rel=1e-6)
        assert o.transient_ratio == pytest.approx(float(v_r / max(v_i, np.float32(1e-8))), rel=1e-6)
        assert o.is_calm == int(v_r < 0.02 and v_i < 0.02)
        assert o.is_regime_shift == int(v_r > 0.02 and v_i > 0.02)
        assert o.is_transient == int(v_r > 0.02 and o.transient_ratio > 1.3 and v_i < 0.02)
        assert (o.cross_scale_boost, o.mim_prob_inertial_used) == (0, pytest.approx(0.25))
        flags.add((o.is_calm, o.is_regime_shift, o.is_transient))
        vol_prev, y_prev = o.vol_combined, float(y[t])
    assert a.state.tick_count == len(y) and a.state.y_prev == np.float32(y[-1])
    assert len(flags) >= 2  # the shock changes the classification
    a.close()
    b.close()
