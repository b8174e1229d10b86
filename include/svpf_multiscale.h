/*
 * Two-scale filter over the C ABI: a fast REACTIVE scout and the INERTIAL
 * production filter stepped on the same observations, with the combined
 * estimate and the transient / regime-shift flags.  Same types, names,
 * defaults and call order as the reference's include/svpf_multiscale.cuh
 * (there static inline in the header; here exported by libsvpf_amd.so from
 * csrc/svpf_multiscale.cpp, so a caller includes this header and links the
 * library as for svpf.h).  Streams are HIP streams.
 */
#ifndef SVPF_MULTISCALE_H
#define SVPF_MULTISCALE_H

#include "svpf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/svpf_multiscale.cuh:43-61 */
typedef struct {
    int n_particles_reactive;
    int n_particles_inertial;
    int n_stein_reactive;
    int n_stein_inertial;
    float surprise_mild_threshold;
    float surprise_high_threshold;
    float boost_mild_factor;
    float boost_high_factor;
    float weight_reactive;
    float weight_inertial;
} MS_SVPF_Config;

/* include/svpf_multiscale.cuh:67-104 */
typedef struct {
    float vol_reactive;
    float vol_inertial;
    float h_reactive;
    float h_inertial;
    float loglik_reactive;
    float loglik_inertial;
    float vol_combined;
    float h_combined;
    float loglik_combined;
    float transient_magnitude; /* vol_reactive - vol_inertial */
    float transient_ratio;     /* vol_reactive / max(vol_inertial, 1e-8) */
    float ksd_reactive;
    float ksd_inertial;
    int stein_steps_reactive;
    int stein_steps_inertial;
    float surprise_reactive; /* |y_t| / max(previous combined vol, 1e-6) */
    int cross_scale_boost;   /* always 0: the boost is disabled there */
    float mim_prob_inertial_used;
    int is_transient;
    int is_regime_shift;
    int is_calm;
} MS_SVPF_Output;

/* include/svpf_multiscale.cuh:110-139 */
typedef struct {
    SVPFState* reactive;
    SVPFState* inertial;
    MS_SVPF_Config config;
    float rho_reactive;
    float sigma_z_reactive;
    float rho_inertial;
    float sigma_z_inertial;
    float vol_prev;
    float y_prev;
    float surprise_ema;
    float mim_jump_prob_inertial_base;
    hipStream_t stream_reactive;
    hipStream_t stream_inertial;
    int initialized;
    int tick_count;
} MS_SVPF;

/* include/svpf_multiscale.cuh:145-167 */
MS_SVPF_Config ms_svpf_default_config(void);
/* include/svpf_multiscale.cuh:173-365: two filters on their own streams,
 * configured as the reference's scout and production filters; NULL on failure */
MS_SVPF* ms_svpf_create(const MS_SVPF_Config* config);
/* include/svpf_multiscale.cuh:367-376 */
void ms_svpf_destroy(MS_SVPF* ms);
/* include/svpf_multiscale.cuh:383-402 */
void ms_svpf_init(MS_SVPF* ms, float initial_vol, const SVPFParams* base_params, unsigned int seed);
/* include/svpf_multiscale.cuh:408-532 */
void ms_svpf_step(MS_SVPF* ms, float y_t, const SVPFParams* base_params, MS_SVPF_Output* out);
/* include/svpf_multiscale.cuh:539-561 */
float ms_svpf_vol_for_execution(const MS_SVPF_Output* out);
float ms_svpf_vol_for_position(const MS_SVPF_Output* out);
float ms_svpf_vol_conservative(const MS_SVPF_Output* out);
int ms_svpf_is_noise(const MS_SVPF_Output* out);
int ms_svpf_regime_warning(const MS_SVPF_Output* out);
/* include/svpf_multiscale.cuh:567-620 */
void ms_svpf_print_state(const MS_SVPF* ms, const MS_SVPF_Output* out);
void ms_svpf_print_config(const MS_SVPF* ms);
/* include/svpf_multiscale.cuh:626-680 */
void ms_svpf_set_reactive_dynamics(MS_SVPF* ms, float rho, float sigma_z, float mim_prob);
void ms_svpf_set_inertial_dynamics(MS_SVPF* ms, float rho, float sigma_z, float mim_prob_base);
void ms_svpf_set_weights(MS_SVPF* ms, float weight_reactive, float weight_inertial);
void ms_svpf_set_boost_config(MS_SVPF* ms, float mild_threshold, float high_threshold, float mild_factor,
                              float high_factor);

#ifdef __cplusplus
}
#endif

#endif /* SVPF_MULTISCALE_H */
