"""Float64 restatement of the diagnostics beside the step.

TEST INFRASTRUCTURE ONLY: the checker the HIP diagnostics kernel and the C
host logic are compared against.  Only tests/ import it.

  health_sums / health_check   src/svpf_diagnostics.cu:126-221 (kernel sums),
                               :227-343 (decisions), :87-120 (input check)
  digamma                      src/tools/svpf_gradient_diagnostic.cu:31-45
  nu_score                     src/tools/svpf_gradient_diagnostic.cu:61-180
  sigma_score                  src/tools/svpf_gradient_diagnostic.cu:421-532
  ShockState                   src/tools/svpf_gradient_diagnostic.cu:336-392

Pins (tests/test_oracle_diagnostics.py): digamma against scipy.special.digamma;
the per-particle nu and sigma scores against central differences of
scipy.stats' Student-t and normal log-densities; ESS against its closed forms
(equal weights: N; one live weight: 1); the shock machine against a
hand-traced sequence.
"""
from __future__ import annotations

import math

import numpy as np


def health_sums(h, lw):
    """(nan_count, inf_count, sum_w, sum_w2, min_h, max_h, sum_h) with the
    weights exp(lw - max) over log-weights that are not NaN and < 1e10."""
    h = np.asarray(h, dtype=np.float64)
    lw = np.asarray(lw, dtype=np.float64)
    nan = float(np.isnan(h).sum() + np.isnan(lw).sum())
    inf = float(((h > 1e10) | (h < -1e10)).sum() + (lw > 1e10).sum())
    okw = ~np.isnan(lw) & (lw < 1e10)
    mx = max(-1e10, float(lw[okw].max())) if okw.any() else -1e10
    w = np.exp(lw[okw] - mx)
    okh = ~np.isnan(h) & (h > -1e10) & (h < 1e10)
    hv = h[okh]
    return (nan, inf, float(w.sum()), float((w * w).sum()), float(hv.min()) if hv.size else 1e10,
            float(hv.max()) if hv.size else -1e10, float(hv.sum()))


def health_check(h, lw, ess_warning=0.3, vol_max=10.0, vol_min=1e-4):
    """Status name and metrics of one svpf_diag_check after a step."""
    nan, inf, sw, sw2, _, _, sh = health_sums(h, lw)
    n = len(h)
    if nan > 0:
        return "NAN_IN_PARTICLES", {}
    if inf > 0:
        return "INF_IN_PARTICLES", {}
    if not sw > 1e-10:
        return "WEIGHT_UNDERFLOW", {}
    ess = sw * sw / sw2
    hm = sh / n
    m = {"ess": ess, "ess_ratio": ess / n, "h_mean": hm, "vol_estimate": math.exp(hm / 2)}
    st = "OK"
    if m["ess_ratio"] < ess_warning:
        st = "ESS_LOW"
    if m["vol_estimate"] > vol_max or m["vol_estimate"] < vol_min:
        st = "VOL_EXTREME"
    return st, m


def check_input(y, max_return=0.5):
    if y != y:
        return "INPUT_IS_NAN"
    if y > 1e30 or y < -1e30:
        return "INPUT_IS_INF"
    if abs(y) > max_return:
        return "INPUT_EXTREME_VALUE"
    return "OK"


def digamma(x):
    """psi(x), x > 0: recurrence up to x >= 6, then the asymptotic series
    ln x - 1/(2x) - 1/(12x^2) + 1/(120x^4) - 1/(252x^6)."""
    x = np.asarray(x, dtype=np.float64).copy()
    r = np.zeros_like(x)
    for _ in range(6):
        lo = x < 6.0
        r = np.where(lo, r - 1.0 / x, r)
        x = np.where(lo, x + 1.0, x)
    ix2 = 1.0 / (x * x)
    return r + np.log(x) - 0.5 / x + ix2 * (-1.0 / 12 + ix2 * (1.0 / 120 - ix2 / 252))


def _weights(lw):
    lw = np.asarray(lw, dtype=np.float64)
    fin = np.isfinite(lw)
    mx = max(-1e30, float(lw[fin].max())) if fin.any() else -1e30
    return np.exp(lw - mx)


def nu_score_particles(h, y, nu):
    """d/dnu of log p(y | h, nu) at each particle (Student-t observation, z = y / (exp(h/2) + 1e-8))."""
    h = np.asarray(h, dtype=np.float64)
    z2 = (y / (np.exp(0.5 * h) + 1e-8)) ** 2
    A = 1.0 + z2 / nu
    psi = float(digamma(0.5 * (nu + 1.0)) - digamma(0.5 * nu))
    return 0.5 * (psi - 1.0 / nu - np.log(A) + (nu + 1.0) * z2 / (nu * nu * A)), z2


def nu_score(h, lw, y, nu):
    """(weighted mean nu score, weighted mean z^2), weights exp(lw - max finite lw), / (sum w + 1e-8)."""
    w = _weights(lw)
    g, z2 = nu_score_particles(h, y, nu)
    sw = w.sum()
    return float((w * g).sum() / (sw + 1e-8)), float((w * z2).sum() / (sw + 1e-8))


def sigma_score_particles(h_pred, h_prev, mu, rho, sigma):
    """d/dsigma of log N(h_pred; mu + rho (h_prev - mu), sigma^2) at each particle, and eps^2 / sigma^2."""
    eps = np.asarray(h_pred, np.float64) - mu - rho * (np.asarray(h_prev, np.float64) - mu)
    e2 = eps * eps / (sigma * sigma)
    return (e2 - 1.0) / sigma, e2


def sigma_score(h_pred, h_prev, lw, mu, rho, sigma):
    w = _weights(lw)
    g, e2 = sigma_score_particles(h_pred, h_prev, mu, rho, sigma)
    sw = w.sum()
    return float((w * g).sum() / (sw + 1e-8)), float((w * e2).sum() / (sw + 1e-8))


class ShockState:
    """CALM (0) -> SHOCK (1) when z^2 > threshold; SHOCK -> RECOVERY (2) after
    `shock_duration` ticks; RECOVERY -> SHOCK on another surprise, -> CALM after
    `recovery_duration` ticks with z^2 < exit threshold.  The tick counter is
    advanced before the test and reset on every transition."""

    def __init__(self, threshold=9.0, shock_duration=20, recovery_duration=50, exit_threshold=4.0):
        self.state, self.ticks = 0, 0
        self.threshold, self.shock_duration = threshold, shock_duration
        self.recovery_duration, self.exit_threshold = recovery_duration, exit_threshold

    def update(self, z2):
        self.ticks += 1
        nxt = self.state
        if self.state == 0 and z2 > self.threshold:
            nxt = 1
        elif self.state == 1 and self.ticks >= self.shock_duration:
            nxt = 2
        elif self.state == 2:
            if z2 > self.threshold:
                nxt = 1
            elif self.ticks >= self.recovery_duration and z2 < self.exit_threshold:
                nxt = 0
        if nxt != self.state:
            self.state, self.ticks = nxt, 0
        return self.state

    def lr_multiplier(self, shock_mult):
        return {0: 1.0, 1: 0.0, 2: shock_mult}[self.state]
