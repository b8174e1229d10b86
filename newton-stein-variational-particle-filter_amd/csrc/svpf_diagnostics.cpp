// Diagnostics beside the step: the reference's health check
// (include/svpf_diagnostics.h, src/svpf_diagnostics.cu) and score diagnostics
// (include/svpf.cuh:895-1256, src/tools/svpf_gradient_diagnostic.cu), over the
// reduction kernel of svpf_diag.hip.  The device work reads the state's
// arrays on its stream and returns a few floats through the state's pinned
// scratch; nothing is allocated per call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/svpf.h"
#include "../../include/svpf_diagnostics.h"
#include "svpf_internal.h"
#include "svpf_kernels.h"

namespace {

// Launch one reduction on the state's stream and read back `nout` floats from
// the device scratch (out0 = scratch, out1 = scratch + 1 for the score modes).
hipError_t reduce_on_state(const SVPFState* s, int mode, const float* h, const float* h2, float a, float b, float c,
                           int nout, float* host) {
    float* dev = svpf_internal_diag_dev(s);
    float* pin = svpf_internal_diag_host(s);
    if (!dev || !pin) return hipErrorInvalidValue;
    hipError_t e = svpf_launch_diag(mode, h, h2, s->log_weights, s->n_particles, a, b, c, dev, dev + 1, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pin, dev, nout * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        svpf_internal_set_error(s, SVPF_RC_HIP, e);
        return e;
    }
    std::memcpy(host, pin, nout * sizeof(float));
    return hipSuccess;
}

SVPFStatus record_error(SVPFDiagnostics* d, SVPFStatus st, bool stamp) {
    d->status = st;
    d->last_error = st;
    if (stamp) d->last_error_timestep = d->steps_processed;
    d->error_count++;
    return st;
}

void set_warning(SVPFDiagnostics* d, SVPFStatus st) {
    d->status = st;
    d->warning_count++;
}

}  // namespace

extern "C" {

// src/svpf_diagnostics.cu:17-64
const char* svpf_status_name(SVPFStatus status) {
    switch (status) {
        case SVPF_OK: return "OK";
        case SVPF_ERR_CUDA_INIT: return "CUDA_INIT_FAILED";
        case SVPF_ERR_CUDA_MALLOC: return "CUDA_MALLOC_FAILED";
        case SVPF_ERR_CUDA_MEMCPY: return "CUDA_MEMCPY_FAILED";
        case SVPF_ERR_CUDA_KERNEL: return "CUDA_KERNEL_FAILED";
        case SVPF_ERR_CUDA_GRAPH: return "CUDA_GRAPH_FAILED";
        case SVPF_ERR_CUDA_SYNC: return "CUDA_SYNC_FAILED";
        case SVPF_ERR_CUDA_DEVICE: return "CUDA_DEVICE_ERROR";
        case SVPF_ERR_NAN_PARTICLES: return "NAN_IN_PARTICLES";
        case SVPF_ERR_NAN_WEIGHTS: return "NAN_IN_WEIGHTS";
        case SVPF_ERR_NAN_OUTPUT: return "NAN_IN_OUTPUT";
        case SVPF_ERR_INF_PARTICLES: return "INF_IN_PARTICLES";
        case SVPF_ERR_INF_WEIGHTS: return "INF_IN_WEIGHTS";
        case SVPF_ERR_PARTICLE_COLLAPSE: return "PARTICLE_COLLAPSE";
        case SVPF_ERR_WEIGHT_UNDERFLOW: return "WEIGHT_UNDERFLOW";
        case SVPF_ERR_BANDWIDTH_ZERO: return "BANDWIDTH_ZERO";
        case SVPF_ERR_INPUT_NAN: return "INPUT_IS_NAN";
        case SVPF_ERR_INPUT_INF: return "INPUT_IS_INF";
        case SVPF_ERR_INPUT_EXTREME: return "INPUT_EXTREME_VALUE";
        case SVPF_ERR_PARAMS_INVALID: return "PARAMS_INVALID";
        case SVPF_ERR_NOT_INITIALIZED: return "NOT_INITIALIZED";
        case SVPF_ERR_ALREADY_DESTROYED: return "ALREADY_DESTROYED";
        case SVPF_ERR_STATE_CORRUPTED: return "STATE_CORRUPTED";
        case SVPF_ERR_OUT_OF_MEMORY: return "OUT_OF_MEMORY";
        case SVPF_ERR_SHARED_MEM_EXCEEDED: return "SHARED_MEM_EXCEEDED";
        case SVPF_WARN_ESS_LOW: return "ESS_LOW";
        case SVPF_WARN_VOL_EXTREME: return "VOL_EXTREME";
        case SVPF_WARN_PARTICLES_CLUSTERED: return "PARTICLES_CLUSTERED";
        case SVPF_WARN_GRAPH_RECAPTURE: return "GRAPH_RECAPTURE";
        default: return "UNKNOWN_STATUS";
    }
}

// src/svpf_diagnostics.cu:70-81
void svpf_diag_init(SVPFDiagnostics* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->status = SVPF_OK;
    d->ess_ratio = 1.0f;
    d->ess_warning_threshold = 0.3f;
    d->vol_max_threshold = 10.0f;
    d->vol_min_threshold = 0.0001f;
    d->input_max_return = 0.5f;
}

// src/svpf_diagnostics.cu:87-120: NaN, then infinite (|y| > 1e30), then
// beyond input_max_return; each is an error
SVPFStatus svpf_diag_check_input(float y_t, SVPFDiagnostics* d) {
    if (y_t != y_t) {
        std::snprintf(d->last_error_message, 256, "Input return is NaN");
        return record_error(d, SVPF_ERR_INPUT_NAN, false);
    }
    if (y_t > 1e30f || y_t < -1e30f) {
        std::snprintf(d->last_error_message, 256, "Input return is infinite: %.2e", y_t);
        return record_error(d, SVPF_ERR_INPUT_INF, false);
    }
    if (std::fabs(y_t) > d->input_max_return) {
        std::snprintf(d->last_error_message, 256, "Input return extreme: %.4f (threshold: %.4f)", y_t,
                      d->input_max_return);
        return record_error(d, SVPF_ERR_INPUT_EXTREME, false);
    }
    return SVPF_OK;
}

// src/svpf_diagnostics.cu:227-343.  Order of the checks: no state; a HIP
// error pending on the thread or recorded on the state; NaN in h or
// log-weights; Inf; sum of weights <= 1e-10 (underflow).  Then ESS =
// (sum w)^2 / sum w^2, h_mean = (sum of finite h) / N, vol = exp(h_mean / 2),
// and the warnings ESS low, vol high, vol low (the last one raised is the
// status).  steps_processed counts the checks that got that far.
SVPFStatus svpf_diag_check(void* state_ptr, SVPFDiagnostics* d) {
    SVPFState* s = static_cast<SVPFState*>(state_ptr);
    if (!s || !s->impl) {
        std::snprintf(d->last_error_message, 256, "SVPF state is NULL");
        return record_error(d, SVPF_ERR_NOT_INITIALIZED, false);
    }
    hipError_t pending = hipGetLastError();
    if (pending == hipSuccess && svpf_last_error(s) == SVPF_RC_HIP) pending = hipErrorUnknown;
    if (pending != hipSuccess) {
        std::snprintf(d->last_error_message, 256, "HIP error: %s",
                      svpf_last_error(s) == SVPF_RC_HIP ? svpf_last_error_string(s) : hipGetErrorString(pending));
        return record_error(d, SVPF_ERR_CUDA_KERNEL, false);
    }
    float r[7];
    hipError_t e = reduce_on_state(s, SVPF_DIAG_HEALTH, s->h, nullptr, 0.f, 0.f, 0.f, 7, r);
    if (e != hipSuccess) {
        std::snprintf(d->last_error_message, 256, "HIP error: %s", hipGetErrorString(e));
        return record_error(d, SVPF_ERR_CUDA_KERNEL, false);
    }
    const float nan_count = r[0], inf_count = r[1], sum_w = r[2], sum_w2 = r[3], sum_h = r[6];
    if (nan_count > 0) {
        std::snprintf(d->last_error_message, 256, "NaN detected in %.0f particles/weights", nan_count);
        return record_error(d, SVPF_ERR_NAN_PARTICLES, true);
    }
    if (inf_count > 0) {
        std::snprintf(d->last_error_message, 256, "Inf detected in %.0f particles/weights", inf_count);
        return record_error(d, SVPF_ERR_INF_PARTICLES, true);
    }
    if (!(sum_w > 1e-10f)) {
        std::snprintf(d->last_error_message, 256, "Weight underflow: sum_w = %.2e", sum_w);
        return record_error(d, SVPF_ERR_WEIGHT_UNDERFLOW, false);
    }
    d->ess = (sum_w * sum_w) / sum_w2;
    d->ess_ratio = d->ess / (float)s->n_particles;
    d->h_mean = sum_h / (float)s->n_particles;
    d->vol_estimate = std::exp(d->h_mean / 2.0f);
    d->status = SVPF_OK;
    if (d->ess_ratio < d->ess_warning_threshold) {
        std::snprintf(d->last_error_message, 256, "ESS low: %.1f%% (threshold: %.1f%%)", d->ess_ratio * 100.0f,
                      d->ess_warning_threshold * 100.0f);
        set_warning(d, SVPF_WARN_ESS_LOW);
    }
    if (d->vol_estimate > d->vol_max_threshold) {
        std::snprintf(d->last_error_message, 256, "Vol estimate extreme high: %.4f", d->vol_estimate);
        set_warning(d, SVPF_WARN_VOL_EXTREME);
    }
    if (d->vol_estimate < d->vol_min_threshold) {
        std::snprintf(d->last_error_message, 256, "Vol estimate extreme low: %.6f", d->vol_estimate);
        set_warning(d, SVPF_WARN_VOL_EXTREME);
    }
    d->steps_processed++;
    return d->status;
}

// src/svpf_diagnostics.cu:349-368
const char* svpf_diag_message(const SVPFDiagnostics* d) { return d->last_error_message; }
bool svpf_status_is_error(SVPFStatus st) { return st > 0 && st < 1000; }
bool svpf_status_is_warning(SVPFStatus st) { return st >= 1000; }
void svpf_diag_clear_error(SVPFDiagnostics* d) {
    d->status = SVPF_OK;
    d->last_error_message[0] = '\0';
}

void svpf_diag_print_summary(const SVPFDiagnostics* d) {
    std::printf("\nSVPF diagnostics\n");
    std::printf("  status           %s\n", svpf_status_name(d->status));
    std::printf("  steps processed  %llu\n", (unsigned long long)d->steps_processed);
    std::printf("  errors           %u\n", d->error_count);
    std::printf("  warnings         %u\n", d->warning_count);
    std::printf("  ESS ratio        %.1f%% (threshold %.0f%%)\n", d->ess_ratio * 100.0f,
                d->ess_warning_threshold * 100.0f);
    std::printf("  vol estimate     %.4f\n", d->vol_estimate);
    std::printf("  h_mean           %.4f\n", d->h_mean);
    if (d->status != SVPF_OK) std::printf("  last message     %s\n", d->last_error_message);
    std::printf("\n");
}

// src/svpf_diagnostics.cu:383-403: the same keys, order and number formats
int svpf_diag_to_json(const SVPFDiagnostics* d, char* buffer, int size) {
    return std::snprintf(buffer, size,
                         "{\"status\":\"%s\",\"status_code\":%d,\"steps\":%llu,\"errors\":%u,\"warnings\":%u,"
                         "\"ess_ratio\":%.4f,\"vol_estimate\":%.6f,\"h_mean\":%.4f,\"message\":\"%s\"}",
                         svpf_status_name(d->status), (int)d->status, (unsigned long long)d->steps_processed,
                         d->error_count, d->warning_count, d->ess_ratio, d->vol_estimate, d->h_mean,
                         d->status != SVPF_OK ? d->last_error_message : "");
}

// ---- score diagnostics ----

// src/tools/svpf_gradient_diagnostic.cu:182-222 (defaults), 224-241
SVPFGradientDiagnostics* svpf_gradient_diagnostic_create(bool enable_logging, const char* log_path) {
    auto* g = static_cast<SVPFGradientDiagnostics*>(std::calloc(1, sizeof(SVPFGradientDiagnostics)));
    if (!g) return nullptr;
    float* buf = nullptr;
    if (hipMalloc(&buf, 2 * sizeof(float)) != hipSuccess) {
        std::free(g);
        return nullptr;
    }
    g->d_nu_grad = buf;
    g->d_z_sq_mean = buf + 1;
    g->z_sq_ema = 1.0f;  // z^2 has mean 1 under the right model
    g->shock_threshold = 9.0f;
    g->shock_duration = 20;
    g->recovery_duration = 50;
    g->recovery_exit_threshold = 4.0f;
    g->enable_logging = enable_logging;
    if (enable_logging && log_path) {
        g->log_file = std::fopen(log_path, "w");
        if (g->log_file) std::fprintf(g->log_file, "t,y_t,vol,z_sq,nu_grad,nu_grad_ema,z_sq_ema,shock_state\n");
    }
    g->initialized = true;
    return g;
}

void svpf_gradient_diagnostic_destroy(SVPFGradientDiagnostics* g) {
    if (!g) return;
    (void)hipFree(g->d_nu_grad);  // d_z_sq_mean shares its allocation
    if (g->log_file) std::fclose(g->log_file);
    std::free(g);
}

// src/tools/svpf_gradient_diagnostic.cu:247-294: the kernel writes diag's own
// device buffers; EMA with alpha 0.1 on the host; one CSV line per call
void svpf_compute_nu_diagnostic(SVPFState* s, SVPFGradientDiagnostics* g, float y_t, int timestep,
                                float* nu_grad_out, float* z_sq_mean_out) {
    if (!g || !g->initialized || !s || !s->impl) return;
    hipError_t e = svpf_launch_diag(SVPF_DIAG_NU, s->h, nullptr, s->log_weights, s->n_particles, y_t, s->nu, 0.f,
                                    g->d_nu_grad, g->d_z_sq_mean, s->stream);
    float* pin = svpf_internal_diag_host(s);
    if (e == hipSuccess) e = hipMemcpyAsync(pin, g->d_nu_grad, 2 * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        svpf_internal_set_error(s, SVPF_RC_HIP, e);
        return;
    }
    const float nu_grad = pin[0], z_sq = pin[1];
    const float alpha = 0.1f;
    g->nu_gradient_ema = (1.0f - alpha) * g->nu_gradient_ema + alpha * nu_grad;
    g->z_sq_ema = (1.0f - alpha) * g->z_sq_ema + alpha * z_sq;
    if (g->enable_logging && g->log_file)
        std::fprintf(g->log_file, "%d,%f,%f,%f,%f,%f,%f,%d\n", timestep, y_t, s->vol_prev, z_sq, nu_grad,
                     g->nu_gradient_ema, g->z_sq_ema, g->shock_state);
    if (nu_grad_out) *nu_grad_out = nu_grad;
    if (z_sq_mean_out) *z_sq_mean_out = z_sq;
}

// src/tools/svpf_gradient_diagnostic.cu:296-334
void svpf_compute_nu_diagnostic_simple(SVPFState* s, float y_t, float* nu_grad_out, float* z_sq_mean_out) {
    if (!s || !s->impl) return;
    float r[2];
    if (reduce_on_state(s, SVPF_DIAG_NU, s->h, nullptr, y_t, s->nu, 0.f, 2, r) != hipSuccess) return;
    if (nu_grad_out) *nu_grad_out = r[0];
    if (z_sq_mean_out) *z_sq_mean_out = r[1];
}

// src/tools/svpf_gradient_diagnostic.cu:537-582
void svpf_compute_sigma_diagnostic_simple(SVPFState* s, const SVPFParams* p, float* sigma_grad_out,
                                          float* eps_sq_norm_out) {
    if (!s || !s->impl || !p) return;
    float sigma = s->sigma_z_effective;
    if (sigma < 1e-6f) sigma = p->sigma_z;
    float r[2];
    if (reduce_on_state(s, SVPF_DIAG_SIGMA, s->h_pred, s->h_prev, p->mu, p->rho, sigma, 2, r) != hipSuccess) return;
    if (sigma_grad_out) *sigma_grad_out = r[0];
    if (eps_sq_norm_out) *eps_sq_norm_out = r[1];
}

// src/tools/svpf_gradient_diagnostic.cu:588-598
void svpf_snapshot_particles(SVPFState* s, float* d_h_buffer) {
    if (!s || !s->impl || !d_h_buffer) return;
    hipError_t e = hipMemcpyAsync(d_h_buffer, s->h, (size_t)s->n_particles * sizeof(float), hipMemcpyDeviceToDevice,
                                  s->stream);
    if (e != hipSuccess) svpf_internal_set_error(s, SVPF_RC_HIP, e);
}

// src/tools/svpf_gradient_diagnostic.cu:336-392: CALM -> SHOCK when z^2 >
// shock_threshold; SHOCK -> RECOVERY after shock_duration ticks; RECOVERY ->
// SHOCK on another surprise, -> CALM after recovery_duration ticks with z^2
// below recovery_exit_threshold.  ticks_in_state counts this call first.
void svpf_update_shock_state(SVPFGradientDiagnostics* g, float z_sq) {
    if (!g) return;
    g->ticks_in_state++;
    int next = g->shock_state;
    if (g->shock_state == 0) {
        if (z_sq > g->shock_threshold) next = 1;
    } else if (g->shock_state == 1) {
        if (g->ticks_in_state >= g->shock_duration) next = 2;
    } else if (g->shock_state == 2) {
        if (z_sq > g->shock_threshold)
            next = 1;
        else if (g->ticks_in_state >= g->recovery_duration && z_sq < g->recovery_exit_threshold)
            next = 0;
    }
    if (next != g->shock_state) {
        g->shock_state = next;
        g->ticks_in_state = 0;
    }
}

int svpf_get_shock_state(const SVPFGradientDiagnostics* g) { return g ? g->shock_state : 0; }
bool svpf_should_learn(const SVPFGradientDiagnostics* g) { return g ? g->shock_state != 1 : false; }
float svpf_get_lr_multiplier(const SVPFGradientDiagnostics* g, float lr_shock_mult) {
    if (!g) return 1.0f;
    if (g->shock_state == 1) return 0.0f;
    if (g->shock_state == 2) return lr_shock_mult;
    return 1.0f;
}

// src/tools/svpf_gradient_diagnostic.cu:605-679
SVPFNaturalGradientTuner* svpf_tuner_create(const SVPFParams* p, float nu, float base_lr, float prior_weight) {
    if (!p) return nullptr;
    auto* t = static_cast<SVPFNaturalGradientTuner*>(std::calloc(1, sizeof(SVPFNaturalGradientTuner)));
    if (!t) return nullptr;
    t->theta.mu = p->mu;
    t->theta.eta = svpf_unconstrain_rho(p->rho);
    t->theta.kappa = svpf_unconstrain_sigma(p->sigma_z);
    t->theta.kappa_nu = svpf_unconstrain_nu(nu);
    for (int i = 0; i < 4; ++i) t->F[i][i] = 1.0f;
    t->F_ema_decay = 0.95f;
    t->F_reg = 1e-4f;
    t->base_lr = base_lr;
    t->lr_shock_mult = 2.0f;
    t->grad_clip = 1.0f;
    t->prior_weight = prior_weight;
    t->theta_prior = t->theta;
    t->warmup_ticks = 100;
    t->learning_enabled = false;
    return t;
}

void svpf_tuner_destroy(SVPFNaturalGradientTuner* t) { std::free(t); }

void svpf_tuner_get_params(const SVPFNaturalGradientTuner* t, SVPFParams* p, float* nu_out) {
    if (!t) return;
    if (p) {
        p->mu = t->theta.mu;
        p->rho = svpf_constrain_rho(t->theta.eta);
        p->sigma_z = svpf_constrain_sigma(t->theta.kappa);
        p->gamma = 0.0f;
    }
    if (nu_out) *nu_out = svpf_constrain_nu(t->theta.kappa_nu);
}

void svpf_tuner_update(SVPFNaturalGradientTuner* t, const SVPFGradientDiagnostics* g, SVPFParams* p, float* nu_out) {
    (void)g;
    svpf_tuner_get_params(t, p, nu_out);
}

}  // extern "C"
