// Two-scale filter (include/svpf_multiscale.h): host orchestration of two
// filters of this library, following include/svpf_multiscale.cuh of the
// reference (configuration at :173-365, step at :408-532).  Each scale runs
// on its own HIP stream and its own captured graph; a step reads each
// scale's three outputs back before the next, as svpf_step_adaptive does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/svpf_multiscale.h"

namespace {

// The scout: fast dynamics, aggressive MIM, no guide, no annealing, no
// Newton, a small KSD budget (include/svpf_multiscale.cuh:216-268).
void configure_reactive(MS_SVPF* ms) {
    SVPFState* f = ms->reactive;
    ms->rho_reactive = 0.92f;
    ms->sigma_z_reactive = 0.18f;
    f->use_asymmetric_rho = 1;
    f->rho_up = 0.94f;
    f->rho_down = 0.88f;
    f->use_mim = 1;
    f->mim_jump_prob = 0.30f;
    f->mim_jump_scale = 10.0f;
    f->use_guide = 0;
    f->use_guided = 0;
    f->use_adaptive_guide = 0;
    f->use_svld = 1;
    f->temperature = 0.55f;
    f->rmsprop_rho = 0.9f;
    f->rmsprop_eps = 1e-6f;
    f->use_annealing = 0;
    f->n_anneal_steps = 1;
    f->use_newton = 0;
    f->use_full_newton = 0;
    f->use_adaptive_mu = 0;
    f->use_adaptive_sigma = 0;
    f->use_local_params = 0;
    f->delta_rho = 0.0f;
    f->delta_sigma = 0.0f;
    f->use_exact_gradient = 1;
    f->lik_offset = 0.35f;
    f->stein_min_steps = 2;
    f->stein_max_steps = 5;
    f->ksd_improvement_threshold = 0.10f;
}

// The production filter: asymmetric persistence, MIM, the adaptive guide and
// guided predict, annealing, full Newton, adaptive mu and sigma, a generous
// KSD budget (include/svpf_multiscale.cuh:277-357).
void configure_inertial(MS_SVPF* ms) {
    SVPFState* f = ms->inertial;
    ms->rho_inertial = 0.98f;
    ms->sigma_z_inertial = 0.10f;
    f->use_asymmetric_rho = 1;
    f->rho_up = 0.98f;
    f->rho_down = 0.93f;
    f->use_mim = 1;
    f->mim_jump_prob = 0.25f;
    f->mim_jump_scale = 9.0f;
    ms->mim_jump_prob_inertial_base = 0.25f;
    f->use_guide = 1;
    f->use_guide_preserving = 1;
    f->guide_strength = 0.05f;
    f->guide_mean = 0.0f;
    f->guide_var = 0.0f;
    f->guide_K = 0.0f;
    f->guide_initialized = 0;
    f->use_adaptive_guide = 1;
    f->guide_strength_base = 0.05f;
    f->guide_strength_max = 0.30f;
    f->guide_innovation_threshold = 1.0f;
    f->use_guided = 1;
    f->guided_alpha_base = 0.0f;
    f->guided_alpha_shock = 0.40f;
    f->guided_innovation_threshold = 1.5f;
    f->use_svld = 1;
    f->temperature = 0.45f;
    f->rmsprop_rho = 0.9f;
    f->rmsprop_eps = 1e-6f;
    f->use_annealing = 1;
    f->n_anneal_steps = 3;
    f->use_newton = 1;
    f->use_full_newton = 1;
    f->use_adaptive_mu = 1;
    f->mu_state = -3.5f;
    f->mu_var = 1.0f;
    f->mu_process_var = 0.001f;
    f->mu_obs_var_scale = 11.0f;
    f->mu_min = -4.0f;
    f->mu_max = -1.0f;
    f->use_adaptive_sigma = 1;
    f->sigma_boost_threshold = 0.95f;
    f->sigma_boost_max = 3.2f;
    f->sigma_z_effective = 0.10f;
    f->use_local_params = 0;
    f->delta_rho = 0.0f;
    f->delta_sigma = 0.0f;
    f->use_exact_gradient = 1;
    f->lik_offset = 0.35f;
    f->stein_min_steps = 8;
    f->stein_max_steps = 16;
    f->ksd_improvement_threshold = 0.05f;
}

SVPFParams scaled(const SVPFParams* base, float rho, float sigma_z) {
    SVPFParams p = *base;
    p.rho = rho;
    p.sigma_z = sigma_z;
    return p;
}

}  // namespace

extern "C" {

MS_SVPF_Config ms_svpf_default_config(void) {
    MS_SVPF_Config c;
    c.n_particles_reactive = 256;
    c.n_particles_inertial = 1024;
    c.n_stein_reactive = 3;
    c.n_stein_inertial = 8;
    c.surprise_mild_threshold = 2.0f;
    c.surprise_high_threshold = 3.5f;
    c.boost_mild_factor = 1.5f;
    c.boost_high_factor = 2.5f;
    c.weight_reactive = 0.10f;
    c.weight_inertial = 0.90f;
    return c;
}

MS_SVPF* ms_svpf_create(const MS_SVPF_Config* config) {
    auto* ms = static_cast<MS_SVPF*>(std::calloc(1, sizeof(MS_SVPF)));
    if (!ms) return nullptr;
    ms->config = config ? *config : ms_svpf_default_config();
    bool ok = hipStreamCreateWithFlags(&ms->stream_reactive, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&ms->stream_inertial, hipStreamNonBlocking) == hipSuccess;
    if (ok) {
        ms->reactive = svpf_create(ms->config.n_particles_reactive, ms->config.n_stein_reactive, 5.0f,
                                   ms->stream_reactive);
        ms->inertial = svpf_create(ms->config.n_particles_inertial, ms->config.n_stein_inertial, 5.0f,
                                   ms->stream_inertial);
    }
    if (!ok || !ms->reactive || !ms->inertial) {
        ms_svpf_destroy(ms);
        return nullptr;
    }
    configure_reactive(ms);
    configure_inertial(ms);
    ms->vol_prev = 0.01f;
    return ms;
}

void ms_svpf_destroy(MS_SVPF* ms) {
    if (!ms) return;
    if (ms->reactive) svpf_destroy(ms->reactive);
    if (ms->inertial) svpf_destroy(ms->inertial);
    if (ms->stream_reactive) (void)hipStreamDestroy(ms->stream_reactive);
    if (ms->stream_inertial) (void)hipStreamDestroy(ms->stream_inertial);
    std::free(ms);
}

// Both scales start from their own dynamics with the same seed.
void ms_svpf_init(MS_SVPF* ms, float initial_vol, const SVPFParams* base, unsigned int seed) {
    if (!ms || !base) return;
    SVPFParams pr = scaled(base, ms->rho_reactive, ms->sigma_z_reactive);
    SVPFParams pi = scaled(base, ms->rho_inertial, ms->sigma_z_inertial);
    svpf_initialize(ms->reactive, &pr, seed);
    svpf_initialize(ms->inertial, &pi, seed);
    ms->vol_prev = initial_vol;
    ms->y_prev = 0.0f;
    ms->surprise_ema = 0.0f;
    ms->initialized = 1;
    ms->tick_count = 0;
}

// The scout steps with its own rho / sigma_z, the production filter with the
// caller's parameters; the combined estimate is the production filter's
// (weights 0 / 1, as there), and the scales' ratio classifies the tick.
void ms_svpf_step(MS_SVPF* ms, float y_t, const SVPFParams* base, MS_SVPF_Output* out) {
    if (!ms || !base || !out) return;
    SVPFParams pr = scaled(base, ms->rho_reactive, ms->sigma_z_reactive);
    float ll_r = 0.f, vol_r = 0.f, h_r = 0.f;
    svpf_step_adaptive(ms->reactive, y_t, ms->y_prev, &pr, &ll_r, &vol_r, &h_r);
    out->vol_reactive = vol_r;
    out->h_reactive = h_r;
    out->loglik_reactive = ll_r;
    out->ksd_reactive = ms->reactive->ksd_prev;
    out->stein_steps_reactive = ms->reactive->stein_steps_used;

    const float surprise = std::fabs(y_t) / std::fmax(ms->vol_prev, 1e-6f);
    ms->surprise_ema = 0.3f * surprise + 0.7f * ms->surprise_ema;
    out->surprise_reactive = surprise;
    out->cross_scale_boost = 0;
    out->mim_prob_inertial_used = ms->mim_jump_prob_inertial_base;

    float ll_i = 0.f, vol_i = 0.f, h_i = 0.f;
    svpf_step_adaptive(ms->inertial, y_t, ms->y_prev, base, &ll_i, &vol_i, &h_i);
    out->vol_inertial = vol_i;
    out->h_inertial = h_i;
    out->loglik_inertial = ll_i;
    out->ksd_inertial = ms->inertial->ksd_prev;
    out->stein_steps_inertial = ms->inertial->stein_steps_used;

    const float w_r = 0.0f, w_i = 1.0f;
    out->h_combined = w_r * h_r + w_i * h_i;
    out->vol_combined = std::pow(vol_r, w_r) * std::pow(vol_i, w_i);
    out->loglik_combined = w_r * ll_r + w_i * ll_i;
    out->transient_magnitude = vol_r - vol_i;
    out->transient_ratio = vol_r / std::fmax(vol_i, 1e-8f);
    const float elevated = 0.02f, ratio_transient = 1.3f;
    out->is_transient = vol_r > elevated && out->transient_ratio > ratio_transient && vol_i < elevated;
    out->is_regime_shift = vol_r > elevated && vol_i > elevated;
    out->is_calm = vol_r < elevated && vol_i < elevated;

    ms->vol_prev = out->vol_combined;
    ms->y_prev = y_t;
    ms->tick_count++;
}

float ms_svpf_vol_for_execution(const MS_SVPF_Output* o) { return 0.7f * o->vol_reactive + 0.3f * o->vol_inertial; }
float ms_svpf_vol_for_position(const MS_SVPF_Output* o) { return 0.2f * o->vol_reactive + 0.8f * o->vol_inertial; }
float ms_svpf_vol_conservative(const MS_SVPF_Output* o) { return std::fmax(o->vol_reactive, o->vol_inertial); }
int ms_svpf_is_noise(const MS_SVPF_Output* o) { return o->is_transient; }
int ms_svpf_regime_warning(const MS_SVPF_Output* o) { return o->is_regime_shift; }

void ms_svpf_print_state(const MS_SVPF* ms, const MS_SVPF_Output* o) {
    std::printf("tick %d\n", ms->tick_count);
    std::printf("  reactive  vol %.4f  h %.3f  ksd %.4f  stein %d\n", o->vol_reactive, o->h_reactive,
                o->ksd_reactive, o->stein_steps_reactive);
    std::printf("  inertial  vol %.4f  h %.3f  ksd %.4f  stein %d%s\n", o->vol_inertial, o->h_inertial,
                o->ksd_inertial, o->stein_steps_inertial, o->cross_scale_boost ? "  boosted" : "");
    std::printf("  transient %.4f (ratio %.2f)  %s\n", o->transient_magnitude, o->transient_ratio,
                o->is_transient ? "TRANSIENT" : o->is_regime_shift ? "REGIME SHIFT" : o->is_calm ? "CALM" : "MIXED");
}

void ms_svpf_print_config(const MS_SVPF* ms) {
    const float hl_r = std::log(0.5f) / std::log(ms->rho_reactive);
    const float hl_i = std::log(0.5f) / std::log(ms->rho_inertial);
    std::printf("two-scale filter\n");
    std::printf("  reactive  N=%d rho=%.3f sigma_z=%.3f half-life %.1f ticks, MIM %.0f%%, guide %s\n",
                ms->config.n_particles_reactive, ms->rho_reactive, ms->sigma_z_reactive, hl_r,
                ms->reactive->mim_jump_prob * 100, ms->reactive->use_guide ? "on" : "off");
    std::printf("  inertial  N=%d rho=%.3f sigma_z=%.3f half-life %.1f ticks, MIM %.0f%%, guide %s\n",
                ms->config.n_particles_inertial, ms->rho_inertial, ms->sigma_z_inertial, hl_i,
                ms->mim_jump_prob_inertial_base * 100, ms->inertial->use_guide ? "on" : "off");
    std::printf("  boost at %.1f (x%.1f) / %.1f (x%.1f); weights %.0f%% / %.0f%%\n",
                ms->config.surprise_mild_threshold, ms->config.boost_mild_factor,
                ms->config.surprise_high_threshold, ms->config.boost_high_factor, ms->config.weight_reactive * 100,
                ms->config.weight_inertial * 100);
}

void ms_svpf_set_reactive_dynamics(MS_SVPF* ms, float rho, float sigma_z, float mim_prob) {
    ms->rho_reactive = rho;
    ms->sigma_z_reactive = sigma_z;
    ms->reactive->rho_up = rho;
    ms->reactive->rho_down = rho - 0.02f;
    ms->reactive->mim_jump_prob = mim_prob;
}

void ms_svpf_set_inertial_dynamics(MS_SVPF* ms, float rho, float sigma_z, float mim_prob_base) {
    ms->rho_inertial = rho;
    ms->sigma_z_inertial = sigma_z;
    ms->inertial->rho_up = rho + 0.003f;
    ms->inertial->rho_down = rho - 0.005f;
    ms->inertial->mim_jump_prob = mim_prob_base;
    ms->mim_jump_prob_inertial_base = mim_prob_base;
}

void ms_svpf_set_weights(MS_SVPF* ms, float weight_reactive, float weight_inertial) {
    const float s = weight_reactive + weight_inertial;
    ms->config.weight_reactive = weight_reactive / s;
    ms->config.weight_inertial = weight_inertial / s;
}

void ms_svpf_set_boost_config(MS_SVPF* ms, float mild_threshold, float high_threshold, float mild_factor,
                              float high_factor) {
    ms->config.surprise_mild_threshold = mild_threshold;
    ms->config.surprise_high_threshold = high_threshold;
    ms->config.boost_mild_factor = mild_factor;
    ms->config.boost_high_factor = high_factor;
}

}  // extern "C"
