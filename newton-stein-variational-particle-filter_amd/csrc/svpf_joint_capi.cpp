// svpf_joint_capi.cpp -- host side of the joint state-parameter SVPF
// (include/svpf_joint.h; the reference's svpf_joint/src/svpf_joint.cu).
//
// Device memory is one allocation: the particles [h | h_prev | mu~ | rho~ | sigma~],
// the last scores and log weights, the reference's scratch arrays, and a summary
// block (parameter means, stds, unconstrained stds, collapse flags, then the last
// step's record) that one copy brings back when diagnostics are requested.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "../../include/svpf_joint.h"
#include "svpf_kernels.h"

static_assert(SVPF_JOINT_RECW == SVPF_JOINT_RECW_DEV, "record width");
static_assert(sizeof(SVPFJointDiagnostics) == 16 * sizeof(float), "diagnostics layout");

namespace {

constexpr int kSummWords = 16 + SVPF_JOINT_RECW;

struct joint_impl {
    float* base = nullptr;  // device allocation
    float* summ = nullptr;  // device summary block
    float* host = nullptr;  // pinned copy of the summary block
};

joint_impl* impl_of(const SVPFJointState* s) { return static_cast<joint_impl*>(s->impl); }

svpf_jparams pack(const SVPFJointState* s) {
    const SVPFJointConfig& c = s->config;
    svpf_jparams P;
    memset(&P, 0, sizeof(P));
    P.n = c.n_particles;
    P.n_stein = c.n_stein_steps;
    P.step_h = c.step_h;
    P.step_mu = c.step_mu;
    P.step_rho = c.step_rho;
    P.step_sigma = c.step_sigma;
    P.diff_mu = c.diffusion_mu;
    P.diff_rho = c.diffusion_rho;
    P.diff_sigma = c.diffusion_sigma;
    P.thr_mu = c.collapse_thresh_mu;
    P.thr_rho = c.collapse_thresh_rho;
    P.thr_sigma = c.collapse_thresh_sigma;
    P.mu_pm = c.mu_prior_mean;
    P.mu_pv = c.mu_prior_var;
    P.rho_pm = c.rho_prior_mean;
    P.rho_pv = c.rho_prior_var;
    P.sig_pm = c.sigma_prior_mean;
    P.sig_pv = c.sigma_prior_var;
    P.nu = c.nu;
    P.st_const = c.student_t_const;
    P.lik_offset = c.lik_offset;
    P.prior_weight = c.prior_weight;
    P.k0 = (uint32_t)s->seed;
    P.k1 = (uint32_t)(s->seed >> 32);
    P.filter = 0;
    return P;
}

bool fail(SVPFJointState* s, hipError_t e, const char* what) {
    if (e == hipSuccess) return false;
    s->last_error = (int)e;
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return true;
}

// Copy the summary block to pinned memory and wait.
bool fetch_summary(const SVPFJointState* s) {
    joint_impl* m = impl_of(s);
    hipError_t e = hipMemcpyAsync(m->host, m->summ, kSummWords * sizeof(float), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        const_cast<SVPFJointState*>(s)->last_error = (int)e;
        fprintf(stderr, "svpf_joint: summary copy failed: %s\n", hipGetErrorString(e));
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

// svpf_joint/src/svpf_joint.cu:21-72 (the same defaults; the Student-t constant and
// the implied offset computed the same way).
SVPFJointConfig svpf_joint_default_config(void) {
    SVPFJointConfig c;
    memset(&c, 0, sizeof(c));
    c.n_particles = SVPF_JOINT_DEFAULT_PARTICLES;
    c.n_stein_steps = SVPF_JOINT_DEFAULT_STEIN_STEPS;
    c.step_h = 0.15f;
    c.step_mu = 0.005f;
    c.step_rho = 0.002f;
    c.step_sigma = 0.03f;
    c.diffusion_mu = 0.005f;
    c.diffusion_rho = 0.0005f;
    c.diffusion_sigma = 0.002f;
    c.collapse_thresh_mu = 0.05f;
    c.collapse_thresh_rho = 0.02f;
    c.collapse_thresh_sigma = 0.02f;
    c.mu_prior_mean = -3.5f;
    c.mu_prior_var = 10.0f;
    c.rho_prior_mean = 2.0f;
    c.rho_prior_var = 5.0f;
    c.sigma_prior_mean = -2.0f;
    c.sigma_prior_var = 5.0f;
    c.nu = 30.0f;
    const float nu = c.nu;
    c.student_t_const = lgammaf(0.5f * (nu + 1.0f)) - lgammaf(0.5f * nu) - 0.5f * logf(3.14159265358979323846f * nu);
    // -E[log t_nu^2] = -(log nu + psi(1/2) - psi(nu/2)), psi(nu/2) by its two-term expansion
    const float half = 0.5f * nu;
    const float psi_half_nu = logf(half) - 1.0f / (2.0f * half) - 1.0f / (12.0f * half * half);
    c.student_t_implied_offset = -(logf(nu) + (-1.9635100260214235f) - psi_half_nu);
    c.lik_offset = 0.10f;
    c.prior_weight = 0.05f;
    return c;
}

SVPFJointState* svpf_joint_create(const SVPFJointConfig* config, hipStream_t stream) {
    const SVPFJointConfig cfg = config ? *config : svpf_joint_default_config();
    const int n = cfg.n_particles;
    if (n < 1 || n > SVPF_JOINT_MAX_PARTICLES || cfg.n_stein_steps < 0) {
        fprintf(stderr, "svpf_joint_create: n_particles must lie in [1, %d] and n_stein_steps be >= 0\n",
                SVPF_JOINT_MAX_PARTICLES);
        return nullptr;
    }
    SVPFJointState* s = static_cast<SVPFJointState*>(calloc(1, sizeof(SVPFJointState)));
    joint_impl* m = new (std::nothrow) joint_impl;
    if (!s || !m) {
        free(s);
        delete m;
        return nullptr;
    }
    s->config = cfg;
    s->stream = stream;
    s->impl = m;
    const size_t words = (size_t)11 * n + 8 + kSummWords;
    if (hipMalloc((void**)&m->base, words * sizeof(float)) != hipSuccess ||
        hipHostMalloc((void**)&m->host, kSummWords * sizeof(float), hipHostMallocDefault) != hipSuccess ||
        hipMemsetAsync(m->base, 0, words * sizeof(float), stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        svpf_joint_destroy(s);
        return nullptr;
    }
    float* b = m->base;
    s->d_h = b;
    s->d_h_prev = b + n;
    s->d_mu_tilde = b + 2 * n;
    s->d_rho_tilde = b + 3 * n;
    s->d_sigma_tilde = b + 4 * n;
    s->d_grad_h = b + 5 * n;
    s->d_grad_mu = b + 6 * n;
    s->d_grad_rho = b + 7 * n;
    s->d_grad_sigma = b + 8 * n;
    s->d_log_w = b + 9 * n;
    s->d_temp = b + 10 * n;
    s->d_reduce_out = b + 11 * n;
    m->summ = b + 11 * n + 8;
    s->d_param_mean = m->summ;
    s->d_param_std = m->summ + 4;
    s->d_std_unconstrained = m->summ + 7;
    s->d_collapse_flags = reinterpret_cast<int*>(m->summ + 10);
    s->d_rng = nullptr;
    // svpf_joint/src/svpf_joint.cu:124-128
    s->bw_h = 1.0f;
    s->bw_mu = 0.5f;
    s->bw_rho = 0.2f;
    s->bw_sigma = 0.2f;
    s->timestep = 0;
    s->y_prev = 0.0f;
    return s;
}

void svpf_joint_initialize(SVPFJointState* s, unsigned long long seed) {
    if (!s) return;
    s->seed = seed;
    const svpf_jparams P = pack(s);
    if (fail(s, svpf_launch_joint(&P, s->d_h, s->d_grad_h, impl_of(s)->summ, nullptr, 0.0f, nullptr, 0, 1, 1, s->stream),
             "svpf_joint_initialize"))
        return;
    fail(s, hipStreamSynchronize(s->stream), "svpf_joint_initialize");
    s->timestep = 0;
}

void svpf_joint_step(SVPFJointState* s, float y_t, SVPFJointDiagnostics* diag) {
    if (!s) return;
    const svpf_jparams P = pack(s);
    joint_impl* m = impl_of(s);
    if (fail(s, svpf_launch_joint(&P, s->d_h, s->d_grad_h, m->summ, nullptr, y_t, nullptr, s->timestep, 1, 0, s->stream),
             "svpf_joint_step"))
        return;
    if (diag && fetch_summary(s)) {
        const float* rec = m->host + 16;
        memcpy(diag, rec, sizeof(SVPFJointDiagnostics));
        s->bw_h = rec[16];
        s->bw_mu = rec[17];
        s->bw_rho = rec[18];
        s->bw_sigma = rec[19];
        // svpf_joint/src/svpf_joint.cu:364-368
        if (diag->mu_collapsed || diag->rho_collapsed || diag->sigma_collapsed)
            printf("WARNING [t=%d]: Diversity collapsed - mu:%d rho:%d sigma:%d\n", s->timestep, diag->mu_collapsed,
                   diag->rho_collapsed, diag->sigma_collapsed);
    }
    s->y_prev = y_t;
    s->timestep++;
}

float svpf_joint_get_vol(const SVPFJointState* s) {
    if (!s || !fetch_summary(s)) return NAN;
    return expf(impl_of(s)->host[3] * 0.5f);
}

void svpf_joint_get_params(const SVPFJointState* s, float* mu, float* rho, float* sigma) {
    if (!s || !fetch_summary(s)) return;
    const float* h = impl_of(s)->host;
    if (mu) *mu = h[0];
    if (rho) *rho = h[1];
    if (sigma) *sigma = h[2];
}

void svpf_joint_destroy(SVPFJointState* s) {
    if (!s) return;
    joint_impl* m = impl_of(s);
    if (m) {
        if (m->base) {
            (void)hipStreamSynchronize(s->stream);
            (void)hipFree(m->base);
        }
        if (m->host) (void)hipHostFree(m->host);
        delete m;
    }
    free(s);
}

int svpf_joint_run_sequence_device(SVPFJointState* s, const float* d_y, int T, float* d_records) {
    if (!s || !d_y || T < 1) return (int)hipErrorInvalidValue;
    const svpf_jparams P = pack(s);
    const hipError_t e =
        svpf_launch_joint(&P, s->d_h, s->d_grad_h, impl_of(s)->summ, d_y, 0.0f, d_records, s->timestep, T, 0, s->stream);
    if (fail(s, e, "svpf_joint_run_sequence_device")) return (int)e;
    s->timestep += T;
    return 0;
}

int svpf_joint_get_particles(const SVPFJointState* s, float* h, float* h_prev, float* mu_tilde, float* rho_tilde,
                             float* sigma_tilde) {
    if (!s) return (int)hipErrorInvalidValue;
    const size_t bytes = (size_t)s->config.n_particles * sizeof(float);
    float* dst[5] = {h, h_prev, mu_tilde, rho_tilde, sigma_tilde};
    const float* src[5] = {s->d_h, s->d_h_prev, s->d_mu_tilde, s->d_rho_tilde, s->d_sigma_tilde};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 5 && e == hipSuccess; ++k)
        if (dst[k]) e = hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return (int)e;
}

int svpf_joint_set_particles(SVPFJointState* s, const float* h, const float* h_prev, const float* mu_tilde,
                             const float* rho_tilde, const float* sigma_tilde) {
    if (!s) return (int)hipErrorInvalidValue;
    const size_t bytes = (size_t)s->config.n_particles * sizeof(float);
    const float* src[5] = {h, h_prev, mu_tilde, rho_tilde, sigma_tilde};
    float* dst[5] = {s->d_h, s->d_h_prev, s->d_mu_tilde, s->d_rho_tilde, s->d_sigma_tilde};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 5 && e == hipSuccess; ++k)
        if (src[k]) e = hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return (int)e;
}

int svpf_joint_lds_bytes(int n) {
    if (n < 1 || n > SVPF_JOINT_MAX_PARTICLES) return -1;
    return (int)svpf_joint_lds_bytes_host(n);
}

}  // extern "C"
