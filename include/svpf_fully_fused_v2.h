/*
 * svpf_fully_fused_v2.h -- the single-workgroup fused SVPF step, drop-in for
 * include/svpf_fully_fused_v2.cuh of Tugbars/Newton-Stein-Variational-Particle-Filter.
 *
 * One call advances one filter by one tick in one kernel launch; all filter
 * state lives in caller-owned device arrays and device scalars, as in the
 * reference.  Differences at the boundary:
 *   - hipStream_t / hipError_t in place of cudaStream_t / cudaError_t;
 *   - the per-particle cuRAND state array is replaced by a host-side noise
 *     descriptor (svpf_v2_noise): a counter-based generator keyed by
 *     (seed, stream_id, timestep, particle), or caller pre-drawn noise.
 * The step's own semantics (Silverman bandwidth, its beta ramp, raw-score KSD,
 * the EKF guide on the adaptive mu, ...) are those of
 * src/svpf_fully_fused_v2.cu:90-712, not those of svpf_step_graph.
 */
#ifndef SVPF_FULLY_FUSED_V2_H
#define SVPF_FULLY_FUSED_V2_H

#include <hip/hip_runtime_api.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Noise of one call.  z_state == NULL: every draw comes from the counter
 * generator keyed by (seed, stream_id) at (particle i, timestep, purpose):
 * predict normal and MIM uniform (purpose 0), the Langevin normal of Stein
 * iteration k (purpose 16 + k), the rejuvenation uniform and normal (purpose
 * 1).  Otherwise the pre-drawn device arrays are used: z_state[n]; u_jump[n]
 * (NULL: no jump); z_diffusion[diffusion_rows][n], row k feeding iteration k
 * (NULL or k >= diffusion_rows: 0); u_rejuv[n] (NULL: no rejuvenation);
 * z_rejuv[n] (NULL: 0). */
typedef struct svpf_v2_noise {
    unsigned long long seed;
    unsigned int stream_id;
    int diffusion_rows;
    const float* z_state;
    const float* u_jump;
    const float* z_diffusion;
    const float* u_rejuv;
    const float* z_rejuv;
} svpf_v2_noise;

/* include/svpf_fully_fused_v2.cuh:24 -- true for n_particles in [1, 1024]. */
bool svpf_fully_fused_v2_supported(int n_particles);

/* include/svpf_fully_fused_v2.cuh:43-116.  Arrays [n] and the scalar outputs
 * are device pointers; d_guide_mean, d_guide_var and d_mu_state are read and
 * written (guide_var < 1e-6 re-initialises the guide).  d_inv_hessian receives
 * the last iteration's curvature when use_newton or use_full_newton is set.
 * The launch is asynchronous on `stream`; the return value is the launch
 * status (hipErrorInvalidValue for an unsupported n or a NULL pointer). */
hipError_t svpf_fully_fused_step_v2(
    float* h, float* h_prev, float* grad_log_p, float* log_weights, float* d_grad_v, float* d_inv_hessian,
    const svpf_v2_noise* noise,
    float y_t, float y_prev, float h_mean_prev, float vol_prev, float ksd_prev,
    float* d_h_mean, float* d_vol, float* d_loglik, float* d_bandwidth, float* d_ksd,
    float* d_guide_mean, float* d_guide_var,
    float rho, float sigma_z, float mu, float nu, float lik_offset, float student_t_const,
    float student_t_implied_offset, float gamma,
    float delta_rho, float delta_sigma,
    float mim_jump_prob, float mim_jump_scale,
    float guide_strength_base, float guide_strength_max, float guide_innovation_threshold,
    float guided_alpha_base, float guided_alpha_shock, float guided_innovation_thresh_predict,
    int use_guide, int use_guided_predict, int use_guide_preserving,
    int use_newton, int use_full_newton,
    int use_rejuvenation, float rejuv_ksd_threshold, float rejuv_prob, float rejuv_blend,
    int use_adaptive_mu, float* d_mu_state, float mu_ema_alpha,
    int use_adaptive_sigma, float sigma_boost_threshold, float sigma_boost_max,
    float step_size, float temperature, float rmsprop_rho, float rmsprop_eps,
    int n_stein_steps, int n_anneal_steps, int stein_sign_mode,
    int n, int timestep, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif
