// svpf_v2.hip -- the single-workgroup fused step (svpf_fully_fused_step_v2)
// and its C ABI (include/svpf_fully_fused_v2.h).  One launch of one workgroup
// per call; the tile shape is chosen by n.
#include <hip/hip_runtime.h>

#include "../../include/svpf_fully_fused_v2.h"
#include "svpf_v2.hpp"

using namespace svpf;

namespace {

constexpr int kV2MaxN = 1024;

template <int BS, int R, int CS>
__global__ __launch_bounds__(BS) void svpf_v2_kernel(svpf_v2_kargs A) {
    extern __shared__ __align__(16) char lds[];
    V2Step<BS, R, CS> e;
    e.bind(lds, A.n, A.n);
    e.run(A);
}

template <int BS, int R, int CS>
hipError_t launch_v2(const svpf_v2_kargs& A, hipStream_t stream) {
    svpf_v2_kernel<BS, R, CS><<<1, BS, LdsLayout::bytes(A.n), stream>>>(A);
    return hipGetLastError();
}

}  // namespace

extern "C" bool svpf_fully_fused_v2_supported(int n_particles) { return n_particles >= 1 && n_particles <= kV2MaxN; }

// include/svpf_fully_fused_v2.cuh:43-116, src/svpf_fully_fused_v2.cu:720-784
extern "C" hipError_t svpf_fully_fused_step_v2(
    float* h, float* h_prev, float* grad_log_p, float* log_weights, float* d_grad_v, float* d_inv_hessian,
    const svpf_v2_noise* noise, float y_t, float y_prev, float h_mean_prev, float vol_prev, float ksd_prev,
    float* d_h_mean, float* d_vol, float* d_loglik, float* d_bandwidth, float* d_ksd, float* d_guide_mean,
    float* d_guide_var, float rho, float sigma_z, float mu, float nu, float lik_offset, float student_t_const,
    float student_t_implied_offset, float gamma, float delta_rho, float delta_sigma, float mim_jump_prob,
    float mim_jump_scale, float guide_strength_base, float guide_strength_max, float guide_innovation_threshold,
    float guided_alpha_base, float guided_alpha_shock, float guided_innovation_thresh_predict, int use_guide,
    int use_guided_predict, int use_guide_preserving, int use_newton, int use_full_newton, int use_rejuvenation,
    float rejuv_ksd_threshold, float rejuv_prob, float rejuv_blend, int use_adaptive_mu, float* d_mu_state,
    float mu_ema_alpha, int use_adaptive_sigma, float sigma_boost_threshold, float sigma_boost_max, float step_size,
    float temperature, float rmsprop_rho, float rmsprop_eps, int n_stein_steps, int n_anneal_steps,
    int stein_sign_mode, int n, int timestep, hipStream_t stream) {
    if (!svpf_fully_fused_v2_supported(n) || !noise) return hipErrorInvalidValue;
    if (!h || !h_prev || !grad_log_p || !log_weights || !d_grad_v || !d_h_mean || !d_vol || !d_loglik ||
        !d_bandwidth || !d_ksd || !d_guide_mean || !d_guide_var)
        return hipErrorInvalidValue;
    if ((use_newton || use_full_newton) && !d_inv_hessian) return hipErrorInvalidValue;
    if (use_adaptive_mu && !d_mu_state) return hipErrorInvalidValue;
    svpf_v2_kargs A;
    A.h = h;
    A.h_prev = h_prev;
    A.grad = grad_log_p;
    A.logw = log_weights;
    A.v = d_grad_v;
    A.inv_hess = d_inv_hessian;
    A.d_h_mean = d_h_mean;
    A.d_vol = d_vol;
    A.d_loglik = d_loglik;
    A.d_bw = d_bandwidth;
    A.d_ksd = d_ksd;
    A.d_guide_mean = d_guide_mean;
    A.d_guide_var = d_guide_var;
    A.d_mu_state = d_mu_state;
    A.z_state = noise->z_state;
    A.u_jump = noise->u_jump;
    A.z_diff = noise->z_diffusion;
    A.u_rej = noise->u_rejuv;
    A.z_rej = noise->z_rejuv;
    A.seed = noise->seed;
    A.stream_id = noise->stream_id;
    A.diff_rows = noise->diffusion_rows;
    A.y_t = y_t;
    A.y_prev = y_prev;
    A.h_mean_prev = h_mean_prev;
    A.vol_prev = vol_prev;
    A.ksd_prev = ksd_prev;
    A.rho = rho;
    A.sigma_z = sigma_z;
    A.mu = mu;
    A.nu = nu;
    A.lik_offset = lik_offset;
    A.st_const = student_t_const;
    A.implied_offset = student_t_implied_offset;
    A.gamma = gamma;
    A.delta_rho = delta_rho;
    A.delta_sigma = delta_sigma;
    A.mim_p = mim_jump_prob;
    A.mim_scale = mim_jump_scale;
    A.gs_base = guide_strength_base;
    A.gs_max = guide_strength_max;
    A.gs_thr = guide_innovation_threshold;
    A.ga_base = guided_alpha_base;
    A.ga_shock = guided_alpha_shock;
    A.ga_thr = guided_innovation_thresh_predict;
    A.rej_thr = rejuv_ksd_threshold;
    A.rej_p = rejuv_prob;
    A.rej_blend = rejuv_blend;
    A.mu_alpha = mu_ema_alpha;
    A.sb_thr = sigma_boost_threshold;
    A.sb_max = sigma_boost_max;
    A.step_size = step_size;
    A.temperature = temperature;
    A.rms_rho = rmsprop_rho;
    A.rms_eps = rmsprop_eps;
    A.use_guide = use_guide;
    A.use_guided = use_guided_predict;
    A.use_guide_preserving = use_guide_preserving;
    A.use_newton = use_newton;
    A.use_full_newton = use_full_newton;
    A.use_rejuv = use_rejuvenation;
    A.use_adaptive_mu = use_adaptive_mu;
    A.use_adaptive_sigma = use_adaptive_sigma;
    A.n_stein = n_stein_steps;
    A.n_anneal = n_anneal_steps;
    A.sign_mode = stein_sign_mode;
    A.n = n;
    A.timestep = timestep;
    // one workgroup of 1024 lanes (four waves per SIMD: the pair loops are latency-bound at
    // one), with 2 or 4 lanes splitting each row's column loop below 1024 particles
    if (n <= 256) return launch_v2<1024, 1, 4>(A, stream);
    if (n <= 512) return launch_v2<1024, 1, 2>(A, stream);
    return launch_v2<1024, 1, 1>(A, stream);
}
