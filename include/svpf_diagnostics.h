/*
 * Health checks beside the per-tick step: the C ABI of the reference's
 * include/svpf_diagnostics.h (same names, values, struct layout and call
 * order), implemented in newton-stein-variational-particle-filter_amd/csrc/
 * svpf_diagnostics.cpp over one HIP reduction kernel (csrc/svpf_diag.hip).
 *
 *   SVPFDiagnostics diag;
 *   svpf_diag_init(&diag);
 *   svpf_step_graph(state, y_t, y_prev, &params, &loglik, &vol, &h_mean);
 *   if (svpf_diag_check(state, &diag) != SVPF_OK) puts(svpf_diag_message(&diag));
 *
 * The library's own call results use the SVPF_RC_* codes of svpf.h; the names
 * below are the reference's and do not collide with them.
 */
#ifndef SVPF_DIAGNOSTICS_H
#define SVPF_DIAGNOSTICS_H

#include <hip/hip_runtime_api.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/svpf_diagnostics.h:39-83 */
typedef enum {
    SVPF_OK = 0,

    SVPF_ERR_CUDA_INIT = 100,
    SVPF_ERR_CUDA_MALLOC = 101,
    SVPF_ERR_CUDA_MEMCPY = 102,
    SVPF_ERR_CUDA_KERNEL = 103, /* here: a HIP launch or copy failed */
    SVPF_ERR_CUDA_GRAPH = 104,
    SVPF_ERR_CUDA_SYNC = 105,
    SVPF_ERR_CUDA_DEVICE = 106,

    SVPF_ERR_NAN_PARTICLES = 200,
    SVPF_ERR_NAN_WEIGHTS = 201,
    SVPF_ERR_NAN_OUTPUT = 202,
    SVPF_ERR_INF_PARTICLES = 203,
    SVPF_ERR_INF_WEIGHTS = 204,
    SVPF_ERR_PARTICLE_COLLAPSE = 210,
    SVPF_ERR_WEIGHT_UNDERFLOW = 211,
    SVPF_ERR_BANDWIDTH_ZERO = 212,

    SVPF_ERR_INPUT_NAN = 300,
    SVPF_ERR_INPUT_INF = 301,
    SVPF_ERR_INPUT_EXTREME = 302,
    SVPF_ERR_PARAMS_INVALID = 310,

    SVPF_ERR_NOT_INITIALIZED = 400,
    SVPF_ERR_ALREADY_DESTROYED = 401,
    SVPF_ERR_STATE_CORRUPTED = 402,

    SVPF_ERR_OUT_OF_MEMORY = 500,
    SVPF_ERR_SHARED_MEM_EXCEEDED = 501,

    SVPF_WARN_ESS_LOW = 1000,
    SVPF_WARN_VOL_EXTREME = 1001,
    SVPF_WARN_PARTICLES_CLUSTERED = 1002,
    SVPF_WARN_GRAPH_RECAPTURE = 1003,
} SVPFStatus;

/* include/svpf_diagnostics.h:92-128.  svpf_diag_check fills ess, ess_ratio,
 * vol_estimate and h_mean (the other metric fields stay as svpf_diag_init left
 * them, as in the reference). */
typedef struct {
    SVPFStatus status;
    uint32_t error_count;
    uint32_t warning_count;

    SVPFStatus last_error;
    uint64_t last_error_timestep;
    char last_error_message[256];

    float ess;
    float ess_ratio;
    float particle_spread;
    float weight_entropy;
    float max_weight_ratio;

    float vol_estimate;
    float h_mean;
    float bandwidth;
    float loglik;

    uint64_t steps_processed;
    uint64_t graph_recaptures;
    float avg_step_time_us;
    float max_step_time_us;

    float ess_warning_threshold; /* default 0.3 */
    float vol_max_threshold;     /* default 10.0 */
    float vol_min_threshold;     /* default 0.0001 */
    float input_max_return;      /* default 0.5 */
} SVPFDiagnostics;

/* include/svpf_diagnostics.h:137 */
void svpf_diag_init(SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:152.  `state` is an SVPFState*.  One reduction
 * kernel over h and log_weights on the state's stream, then a synchronous
 * read-back of 7 floats (no allocation per call). */
SVPFStatus svpf_diag_check(void* state, SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:161 */
SVPFStatus svpf_diag_check_input(float y_t, SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:166 */
const char* svpf_diag_message(const SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:171 */
const char* svpf_status_name(SVPFStatus status);
/* include/svpf_diagnostics.h:176 */
bool svpf_status_is_error(SVPFStatus status);
/* include/svpf_diagnostics.h:181 */
bool svpf_status_is_warning(SVPFStatus status);
/* include/svpf_diagnostics.h:186 */
void svpf_diag_clear_error(SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:191 */
void svpf_diag_print_summary(const SVPFDiagnostics* diag);
/* include/svpf_diagnostics.h:201 */
int svpf_diag_to_json(const SVPFDiagnostics* diag, char* buffer, int size);

/* include/svpf_diagnostics.h:207-239: with SVPF_ENABLE_CUDA_CHECKS defined the
 * macros record a failed HIP call in `diag`; otherwise they evaluate the call. */
#ifdef SVPF_ENABLE_CUDA_CHECKS
#define SVPF_CUDA_CHECK(call, diag)                                                                   \
    do {                                                                                              \
        hipError_t svpf_e_ = (call);                                                                  \
        if (svpf_e_ != hipSuccess) {                                                                  \
            (diag)->status = SVPF_ERR_CUDA_KERNEL;                                                    \
            snprintf((diag)->last_error_message, 256, "HIP error at %s:%d: %s", __FILE__, __LINE__, \
                     hipGetErrorString(svpf_e_));                                                     \
            (diag)->last_error = (diag)->status;                                                      \
            (diag)->error_count++;                                                                    \
        }                                                                                             \
    } while (0)
#define SVPF_CUDA_CHECK_LAST(diag) SVPF_CUDA_CHECK(hipGetLastError(), diag)
#else
#define SVPF_CUDA_CHECK(call, diag) (call)
#define SVPF_CUDA_CHECK_LAST(diag) ((void)0)
#endif

/* include/svpf_diagnostics.h:247-251 */
static inline bool svpf_is_healthy_float(float x) { return (x == x) && (x > -1e30f) && (x < 1e30f); }

#ifdef __cplusplus
}
#endif

#endif /* SVPF_DIAGNOSTICS_H */
