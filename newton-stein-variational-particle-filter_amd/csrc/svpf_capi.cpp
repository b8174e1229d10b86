// svpf_capi.cpp -- host side of the C ABI declared in include/svpf.h.
//
// One SVPFState owns one device block: the filter's parameter record, scalar
// state, descriptor, observation slot and output record (uploaded / read back
// once per tick through a pinned mirror), plus the particle pool (h, h_prev,
// RMSProp v, log weights).  Per tick: H2D of the records, one kernel launch,
// D2H of state + outputs, stream sync -- captured once as a HIP graph by
// svpf_step_graph and replayed afterwards.
//
// The per-tick host logic of the reference (src/svpf_optimized_graph.cu:
// 915-1389) runs on the device; what stays on the host is packing the
// configuration fields, the RTS-style smoothing of the reported h_mean
// (src/svpf_optimized_graph.cu:619-737) and the result unpacking.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/svpf.h"
#include "svpf_internal.h"
#include "svpf_kernels.h"
#include "svpf_special.h"
#include "svpf_types.h"

namespace {

constexpr int kBlockParams = 0;
constexpr int kBlockFstate = 256;
constexpr int kBlockY = 384;      // 4 floats
constexpr int kBlockOut = 400;    // 8 floats
constexpr int kBlockDesc = 448;   // svpf_fdesc
constexpr int kBlockDiag = 960;   // diagnostics scratch (16 floats)
constexpr int kBlockBytes = 1024;
constexpr int kUploadBytes = kBlockDesc + (int)sizeof(svpf_fdesc);
static_assert(sizeof(svpf_fparams) == 256, "fparams layout");
static_assert(sizeof(svpf_fstate) == 128, "fstate layout");
static_assert(SVPF_TILE_EXT == SVPF_TILE_EXTENDED, "extended tile bit");
static_assert(kBlockDesc + sizeof(svpf_fdesc) <= (size_t)kBlockDiag, "block layout");
static_assert(kBlockDiag + kDiagScratchFloats * sizeof(float) <= (size_t)kBlockBytes, "block layout");

}  // namespace

struct svpf_impl {
    int n = 0;
    int tile = -1, bs = 256;
    char* dblock = nullptr;   // device records
    char* hblock = nullptr;   // pinned mirror
    float* dpool = nullptr;   // h | hp | v | log_w  (4n floats)
    int* dorder = nullptr;
    float* dseq_out = nullptr;  // per-tick records of a device run
    int dseq_cap = 0;
    float* dseq_y = nullptr;
    int dseq_y_cap = 0;
    // device scalars the reference keeps on the device
    float h_mean_prev = -3.5f, bw_sq_ema = 0.f, return_ema = 0.f, return_var = 0.f;
    // pre-drawn noise
    const float* z_pred = nullptr;
    const float* u_pred = nullptr;
    const float* xi = nullptr;
    const float* u_rej = nullptr;
    const float* z_rej = nullptr;
    int xi_rows = 0, t_base = 0;
    uint64_t seed = 42;
    bool initialized = false;
    int last_error = SVPF_RC_OK;
    hipError_t last_hip = hipSuccess;  // HIP status behind the last SVPF_RC_HIP
    int last_status = 0;               // kernel-reported status behind the last kernel error
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool captured = false;
    float out_last[8] = {0};
    // cooperative multi-workgroup path
    bool coop_ok = false;
    int coop_cs = 8, coop_G = 1, coop_stride = 8;
    uint64_t* xbuf = nullptr;  // tagged exchange words
    uint32_t epoch = 0;
    float med_r2 = 0.0f;  // last median pairwise r^2 (a search seed, not a result)
    bool captured_coop = false;
    bool coop_graph_failed = false;
    // The cooperative kernel runs as a plain launch: its workgroups are co-resident on an
    // otherwise free GPU, and a cooperative launch costs about 20 us more per tick (100 us more
    // once other streams exist).  An exchange that times out (workgroups not all resident)
    // writes nothing back; the tick is then repeated with a cooperative launch, which this
    // filter keeps from then on.
    bool coop_cooperative = false;
    bool captured_cooperative = false;
    int coop_spin_limit = 0;
    int coop_fault_wg = -1;  // failure-path testing (svpf_set_coop_fault)
    // median bandwidth: a workgroup beside the row workgroups selects the median while they run the
    // first gradient (svpf_set_coop_median_workgroup)
    bool coop_bwg = true;
    // grid path (workspace allocated for 2 <= n <= SVPF_GRID_MAX_PARTICLES)
    int grid_mode = SVPF_GRID_AUTO;
    float* dgrid_ws = nullptr;
    int captured_path = -1, captured_smax = -1;
    unsigned captured_ext = 0;
    // coordinate-wise two-factor step: [6][n] f | s | f_prev | s_prev | v_f | v_s
    // (allocated for filters within the single-workgroup capacity) and the factors'
    // particle means after the last tick
    float* d2f = nullptr;
    float mean_fast_prev = 0.0f, mean_slow_prev = 0.0f;
    bool captured_2f = false;

    svpf_fparams* hp() { return reinterpret_cast<svpf_fparams*>(hblock + kBlockParams); }
    svpf_fstate* hs() { return reinterpret_cast<svpf_fstate*>(hblock + kBlockFstate); }
    float* hy() { return reinterpret_cast<float*>(hblock + kBlockY); }
    float* hout() { return reinterpret_cast<float*>(hblock + kBlockOut); }
    svpf_fdesc* hd() { return reinterpret_cast<svpf_fdesc*>(hblock + kBlockDesc); }
    svpf_fparams* dp() { return reinterpret_cast<svpf_fparams*>(dblock + kBlockParams); }
    svpf_fstate* ds() { return reinterpret_cast<svpf_fstate*>(dblock + kBlockFstate); }
    float* dy() { return reinterpret_cast<float*>(dblock + kBlockY); }
    float* dout() { return reinterpret_cast<float*>(dblock + kBlockOut); }
    svpf_fdesc* dd() { return reinterpret_cast<svpf_fdesc*>(dblock + kBlockDesc); }
    float* d_h() { return dpool; }
    float* d_hp() { return dpool + n; }
    float* d_v() { return dpool + 2 * (size_t)n; }
    float* d_lw() { return dpool + 3 * (size_t)n; }
    float* d_hpred() { return dpool + 4 * (size_t)n; }
};

namespace {

// Worker streams (default priority) and events for the batched fork/join,
// created once per process on the current device.  Calls that overlap from
// several host threads share them and are not supported.
constexpr int kMaxForkStreams = 8;
struct ForkJoin {
    hipStream_t streams[kMaxForkStreams];
    hipEvent_t fork;
    hipEvent_t join[kMaxForkStreams];
};

ForkJoin* fork_join() {
    static ForkJoin fj;
    static int state = 0;  // 0 untried, 1 ready, -1 failed
    if (state == 0) {
        bool ok = hipEventCreateWithFlags(&fj.fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; ok && i < kMaxForkStreams; ++i) {
            ok = hipStreamCreateWithFlags(&fj.streams[i], hipStreamNonBlocking) == hipSuccess;
            ok = ok && hipEventCreateWithFlags(&fj.join[i], hipEventDisableTiming) == hipSuccess;
        }
        state = ok ? 1 : -1;
    }
    return state == 1 ? &fj : nullptr;
}

// Cooperative shape for n particles: rows per workgroup = 256 / cs, G = ceil(n / rows).
// Default: about 16 workgroups (rows between 8 and 256), except that up to
// n = 1024 a workgroup takes at most 32 rows (8 lanes each), up to 32 workgroups
// (at n = 640 .. 1000 this measured 15-19% faster per tick than 16 workgroups of
// 4 lanes per row); all on one XCD (stride 8).
bool choose_coop_shape(svpf_impl* I, int cs, int stride) {
    const int n = I->n;
    if (cs <= 0) {
        int rows = 8;
        while (rows < 256 && rows * 16 < n) rows <<= 1;
        if (n <= 1024 && rows > 32) rows = 32;
        cs = 256 / rows;
    }
    if (cs != 1 && cs != 2 && cs != 4 && cs != 8 && cs != 16 && cs != 32) return false;
    if (stride <= 0) stride = 8;
    if (stride != 1 && stride != 8) return false;
    const int rows = 256 / cs;
    const int G = (n + rows - 1) / rows;
    // partial sums of all workgroups (G x 8 words) are gathered into an n-float LDS array
    // (G + 1: the median workgroup)
    const bool fits =
        svpf_coop_lds_bytes(n) <= 159 * 1024 && G * 8 <= n && G <= n / 8 + 2 && (G + 1) * stride <= 2048;
    if (!fits) return false;
    I->coop_cs = cs;
    I->coop_G = G;
    I->coop_stride = stride;
    I->coop_ok = true;
    return true;
}

// Exchange epochs to skip after a timed-out cooperative launch of T ticks: more than
// the launch could have used (per tick one exchange per Stein iteration and the final
// one, plus the commit), so the retry meets no word the failed launch tagged.
uint32_t epochs_after_timeout(const SVPFState* s, int T) {
    return (uint32_t)(1u << 16) + (uint32_t)T * (uint32_t)(std::max(s->stein_max_steps, 0) + 4);
}

int set_err(SVPFState* s, int code) {
    if (s && s->impl) s->impl->last_error = code;
    return code;
}

// Smallest tile (block 256 first, then 512) holding n particles.
bool choose_tile(int n, int* tile, int* bs) {
    static const int order[] = {SVPF_TILE_R1_C4, SVPF_TILE_R1_C2, SVPF_TILE_R1_C1,
                                SVPF_TILE_R2_C1, SVPF_TILE_R4_C1, SVPF_TILE_R8_C1};
    for (int b : {256, 512}) {
        for (int t : order) {
            if (svpf_tile_capacity(t, b) >= n) {
                *tile = t;
                *bs = b;
                return true;
            }
        }
    }
    return false;
}

float student_t_const(float nu) {
    return lgammaf((nu + 1.0f) / 2.0f) - lgammaf(nu / 2.0f) - 0.5f * logf(3.14159265358979f * nu);
}

// src/svpf_optimized_graph.cu:127-138
float implied_offset(float nu) {
    const float psi_half = -1.9635100260214235f;
    float nh = nu / 2.0f, psi;
    if (nh >= 1.0f) psi = logf(nh) - 1.0f / (2.0f * nh) - 1.0f / (12.0f * nh * nh);
    else psi = -0.5772156649f - 1.0f / nh;
    return -(logf(nu) + psi_half - psi);
}

void pack_params(const SVPFState* s, const SVPFParams* p, svpf_fparams* o) {
    std::memset(o, 0, sizeof(*o));
    uint32_t f = 0;
    if (s->use_exact_gradient) f |= SVPF_F_EXACT_GRAD;
    if (s->use_newton) f |= SVPF_F_NEWTON;
    if (s->use_newton && s->use_full_newton) f |= SVPF_F_FULL_NEWTON;
    if (s->use_svld) f |= SVPF_F_SVLD;
    if (s->use_annealing) f |= SVPF_F_ANNEAL;
    if (s->use_adaptive_beta) f |= SVPF_F_ADAPT_BETA;
    if (s->use_linear_beta) f |= SVPF_F_LINEAR_BETA;
    if (s->use_mim) f |= SVPF_F_MIM;
    if (s->use_asymmetric_rho) f |= SVPF_F_ASYM_RHO;
    if (s->use_local_params) f |= SVPF_F_LOCAL_PARAMS;
    if (s->use_guided) f |= SVPF_F_GUIDED;
    if (s->use_guide) f |= SVPF_F_GUIDE;
    if (s->use_guide_preserving) f |= SVPF_F_GUIDE_PRESERVING;
    if (s->use_adaptive_guide) f |= SVPF_F_ADAPT_GUIDE;
    if (s->use_rejuvenation) f |= SVPF_F_REJUV;
    if (s->use_fan_mode) f |= SVPF_F_FAN;
    if (s->use_adaptive_mu) f |= SVPF_F_ADAPT_MU;
    if (s->use_adaptive_sigma) f |= SVPF_F_ADAPT_SIGMA;
    if (s->use_ksd_early_stop) f |= SVPF_F_KSD_EARLY_STOP;
    if (s->stein_repulsive_sign == SVPF_STEIN_SIGN_PAPER) f |= SVPF_F_PAPER_SIGN;
    if (s->use_heun) f |= SVPF_F_HEUN;
    if (s->use_student_t_state) f |= SVPF_F_STUDENT_STATE;
    if (s->use_two_factor) f |= SVPF_F_TWO_FACTOR;
    o->nu_state = s->nu_state;
    o->coop_spin_limit = s->impl->coop_spin_limit;
    o->coop_fault_wg = s->impl->coop_fault_wg;
    o->rho_fast = s->rho_fast;
    o->sigma_fast = s->sigma_fast;
    o->rho_slow = s->rho_slow;
    o->sigma_slow = s->sigma_slow;
    o->bw_floor_fast = s->bw_floor_fast;
    o->bw_floor_slow = s->bw_floor_slow;
    o->guide_strength = s->guide_strength;
    o->n_stein = s->n_stein_steps;
    o->n = s->n_particles;
    o->flags = f;
    o->n_anneal = s->n_anneal_steps;
    o->stein_min = s->stein_min_steps;
    o->stein_max = s->stein_max_steps;
    o->bandwidth_mode = s->bandwidth_mode;
    o->mu = p->mu;
    o->rho = p->rho;
    o->sigma_z = p->sigma_z;
    o->gamma = p->gamma;
    o->nu = s->nu;
    o->lik_offset = s->lik_offset;
    o->temperature = s->temperature;
    o->rms_rho = s->rmsprop_rho;
    o->rms_eps = s->rmsprop_eps;
    o->mim_p = s->mim_jump_prob;
    o->mim_scale = s->mim_jump_scale;
    o->rho_up = s->rho_up;
    o->rho_down = s->rho_down;
    o->delta_rho = s->delta_rho;
    o->delta_sigma = s->delta_sigma;
    o->ga_base = s->guided_alpha_base;
    o->ga_shock = s->guided_alpha_shock;
    o->ga_thr = s->guided_innovation_threshold;
    o->gs_base = s->guide_strength_base;
    o->gs_max = s->guide_strength_max;
    o->gs_thr = s->guide_innovation_threshold;
    o->rej_thr = s->rejuv_ksd_threshold;
    o->rej_p = s->rejuv_prob;
    o->rej_blend = s->rejuv_blend;
    o->mu_q = s->mu_process_var;
    o->mu_rscale = s->mu_obs_var_scale;
    o->mu_min = s->mu_min;
    o->mu_max = s->mu_max;
    o->sb_thr = s->sigma_boost_threshold;
    o->sb_max = s->sigma_boost_max;
    o->ksd_thr = s->ksd_improvement_threshold;
    o->st_const = student_t_const(s->nu);  // recomputed per tick (src/svpf_optimized_graph.cu:950-952)
    o->implied_offset = s->student_t_implied_offset;
    o->base_step = SVPF_STEIN_STEP_SIZE * (s->use_guide ? 0.5f : 1.0f);
}

void pack_fstate(const SVPFState* s, float y_prev, svpf_fstate* o) {
    const svpf_impl* I = s->impl;
    std::memset(o, 0, sizeof(*o));
    o->timestep = s->timestep;
    o->guide_initialized = s->guide_initialized;
    o->stein_steps_used = s->stein_steps_used;
    o->ksd_prev = s->ksd_prev;
    o->vol_prev = s->vol_prev;
    o->h_mean_prev = I->h_mean_prev;
    o->guide_mean = s->guide_mean;
    o->guide_var = s->guide_var;
    o->bw_sq_ema = I->bw_sq_ema;
    o->return_ema = I->return_ema;
    o->return_var = I->return_var;
    o->mu_state = s->mu_state;
    o->mu_var = s->mu_var;
    o->sigma_z_eff = s->sigma_z_effective;
    o->y_prev = y_prev;
    o->epoch = I->epoch;
    o->med_r2 = I->med_r2;
    o->g2f_mean_fast = s->guide_mean_fast;
    o->g2f_mean_slow = s->guide_mean_slow;
    o->g2f_var_fast = s->guide_var_fast;
    o->g2f_var_slow = s->guide_var_slow;
    o->g2f_initialized = s->guide_2f_initialized;
    o->mean_fast_prev = I->mean_fast_prev;
    o->mean_slow_prev = I->mean_slow_prev;
}

void unpack_fstate(SVPFState* s, const svpf_fstate* o) {
    svpf_impl* I = s->impl;
    s->timestep = o->timestep;
    s->guide_initialized = o->guide_initialized;
    s->stein_steps_used = o->stein_steps_used;
    s->ksd_prev = o->ksd_prev;
    s->vol_prev = o->vol_prev;
    s->guide_mean = o->guide_mean;
    s->guide_var = o->guide_var;
    s->mu_state = o->mu_state;
    s->mu_var = o->mu_var;
    if (s->use_adaptive_sigma) s->sigma_z_effective = o->sigma_z_eff;
    I->h_mean_prev = o->h_mean_prev;
    I->bw_sq_ema = o->bw_sq_ema;
    I->return_ema = o->return_ema;
    I->return_var = o->return_var;
    I->epoch = o->epoch;
    I->med_r2 = o->med_r2;
    s->guide_mean_fast = o->g2f_mean_fast;
    s->guide_mean_slow = o->g2f_mean_slow;
    s->guide_var_fast = o->g2f_var_fast;
    s->guide_var_slow = o->g2f_var_slow;
    s->guide_2f_initialized = o->g2f_initialized;
    I->mean_fast_prev = o->mean_fast_prev;
    I->mean_slow_prev = o->mean_slow_prev;
}

void fill_desc(SVPFState* s, svpf_fdesc* d, const float* y, float* out, int tick0) {
    svpf_impl* I = s->impl;
    std::memset(d, 0, sizeof(*d));
    d->off = 0;
    d->y = y;
    d->out = out;
    const int n = I->n;
    if (I->z_pred) {
        long k = (long)tick0 - I->t_base;
        if (k < 0) k = 0;
        d->z_pred = I->z_pred + (size_t)k * n;
        d->u_pred = I->u_pred ? I->u_pred + (size_t)k * n : nullptr;
        d->xi = I->xi ? I->xi + (size_t)k * I->xi_rows * n : nullptr;
        d->u_rej = I->u_rej ? I->u_rej + (size_t)k * n : nullptr;
        d->z_rej = I->z_rej ? I->z_rej + (size_t)k * n : nullptr;
        d->xi_rows = I->xi ? I->xi_rows : 0;
    }
    d->log_w = I->d_lw();
    d->h_pred = I->d_hpred();
    d->seed = I->seed;
    d->param_idx = 0;
    d->filter_id = 0;
    d->tile = I->tile;
}

int check_supported(SVPFState* s) {
    if (s->use_two_factor) {
        // one workgroup holds the filter; its noise comes from the counter generator
        const svpf_impl* I = s->impl;
        if (I->tile < 0 || !I->d2f || I->z_pred) return SVPF_RC_UNSUPPORTED;
        if (s->n_stein_steps < 1 || (s->use_annealing && s->n_anneal_steps < 1)) return SVPF_RC_INVALID_ARGUMENT;
    }
    if (s->n_particles != s->impl->n) return SVPF_RC_INVALID_ARGUMENT;
    if (s->bandwidth_mode < 0 || s->bandwidth_mode > 2) return SVPF_RC_INVALID_ARGUMENT;
    if (s->n_anneal_steps < 1 && s->use_annealing) return SVPF_RC_INVALID_ARGUMENT;
    if (s->stein_max_steps < 0 || s->stein_min_steps < 0 || s->stein_min_steps > s->stein_max_steps)
        return SVPF_RC_INVALID_ARGUMENT;
    return SVPF_RC_OK;
}

// The cooperative multi-workgroup kernel serves one filter when the state asks
// for it (use_persistent_kernel) and the features it implements suffice.
bool coop_applies(const SVPFState* s) {
    const svpf_impl* I = s->impl;
    return !s->use_two_factor && s->use_persistent_kernel && I->coop_ok && !s->use_ksd_early_stop && !s->use_heun &&
           !s->use_student_t_state;
}

// Path of one filter's step.  Forced by svpf_set_grid_path, or by a filter above
// the single-workgroup capacity (tile < 0).  Automatic choice, from per-tick
// timings on MI355X (plain and full-feature filters): the cooperative kernel for
// 256 <= n < 1024 where it applies, the grid path from n = 1024 up and for
// 256 <= n < 1024 where the cooperative kernel does not apply, one workgroup
// below n = 256.
constexpr int kGridAutoMinN = 1024;
constexpr int kCoopMinN = 256;
bool use_grid(const SVPFState* s) {
    const svpf_impl* I = s->impl;
    if (!I->dgrid_ws || s->use_two_factor) return false;
    if (I->tile < 0 || I->grid_mode == SVPF_GRID_ON) return true;
    if (I->grid_mode == SVPF_GRID_OFF) return false;
    if (I->n >= kGridAutoMinN) return true;
    return I->n >= kCoopMinN && !coop_applies(s);
}

bool use_coop(const SVPFState* s) { return !use_grid(s) && coop_applies(s); }

// The median workgroup runs beside at most 32 row workgroups: the plain launch relies on
// every workgroup being resident at once, and 33 fit one XCD (32 CUs) as measured; beside 64
// row workgroups the launch exceeded the XCD's capacity.
bool coop_median_wg(const SVPFState* s) {
    const svpf_impl* I = s->impl;
    return I->coop_bwg && s->bandwidth_mode == SVPF_BW_MEDIAN && I->coop_G <= 32;
}

// launch-sequence variants of the grid path
unsigned grid_ext(const SVPFState* s) {
    return (s->use_heun ? SVPF_GRID_EXT_HEUN : 0u) | (s->use_student_t_state ? SVPF_GRID_EXT_STUDENT : 0u);
}

int step_path(const SVPFState* s) {
    return use_grid(s) ? SVPF_PATH_GRID : (use_coop(s) ? SVPF_PATH_COOP : SVPF_PATH_SINGLE);
}

hipError_t launch_step_kernel(SVPFState* s, int t0, int t1, hipStream_t cs) {
    svpf_impl* I = s->impl;
    if (s->use_two_factor)
        return svpf_launch_2f(I->bs, I->tile, I->dp(), I->ds(), I->dd(), I->d2f, I->d_h(), I->n, t0, t1, cs);
    if (use_grid(s))
        return svpf_launch_grid(I->dp(), I->ds(), I->dd(), I->d_h(), I->d_hp(), I->d_v(), I->d_lw(), I->dgrid_ws,
                                I->n, s->stein_max_steps, grid_ext(s), t0, t1, cs);
    if (use_coop(s))
        return svpf_launch_coop(I->coop_cs, I->coop_G, I->coop_stride, I->dp(), I->ds(), I->dd(), I->d_h(),
                                I->d_hp(), I->d_v(), I->xbuf, I->n, t0, t1, I->coop_cooperative ? 1 : 0,
                                coop_median_wg(s) ? 1 : 0, cs);
    const int tile = I->tile | ((s->use_heun || s->use_student_t_state) ? SVPF_TILE_EXT : 0);
    return svpf_launch_batched(I->bs, tile, 1, I->dp(), I->ds(), I->dd(), I->d_h(), I->d_hp(), I->d_v(), I->dorder,
                               1, nullptr, nullptr, I->n, t0, t1, cs);
}

hipError_t enqueue_tick(SVPFState* s, hipStream_t cs) {
    svpf_impl* I = s->impl;
    hipError_t e = hipMemcpyAsync(I->dblock, I->hblock, kUploadBytes, hipMemcpyHostToDevice, cs);
    if (e != hipSuccess) return e;
    e = launch_step_kernel(s, 0, 1, cs);
    if (e != hipSuccess) return e;
    // fstate (128 B) and the output record (32 B) are adjacent: one copy back.
    return hipMemcpyAsync(I->hblock + kBlockFstate, I->dblock + kBlockFstate, kBlockOut + 32 - kBlockFstate,
                          hipMemcpyDeviceToHost, cs);
}

// RTS-style smoothing of summary statistics (src/svpf_optimized_graph.cu:619-699).
// mu_state: the adaptive mean before this tick's update, which the reference applies after the
// smoothing (src/svpf_optimized_graph.cu:1366-1386).
void smooth_backward(SVPFState* st, float h_mean_new, float h_var_new, float y_t, const SVPFParams* p,
                     float mu_state) {
    if (!st->use_smoothing) return;
    int k = std::min(std::max(st->smooth_lag, 1), SVPF_SMOOTH_MAX_LAG);
    int head = st->smooth_head;
    st->smooth_h_mean[head] = h_mean_new;
    st->smooth_h_var[head] = h_var_new;
    st->smooth_y[head] = y_t;
    st->smooth_head = (head + 1) % k;
    if (st->timestep < k) return;
    float rho = p->rho;
    float mu = st->use_adaptive_mu ? mu_state : p->mu;
    float s2 = p->sigma_z * p->sigma_z;
    for (int lag = 1; lag < k; ++lag) {
        int ic = (st->smooth_head - lag - 1 + k) % k;
        int in = (st->smooth_head - lag + k) % k;
        float hc = st->smooth_h_mean[ic], hn = st->smooth_h_mean[in], vc = st->smooth_h_var[ic];
        float hpred = mu + rho * (hc - mu);
        float pv = rho * rho * vc + s2;
        float J = rho * vc / (pv + 1e-8f);
        st->smooth_h_mean[ic] = hc + J * (hn - hpred);
        st->smooth_h_var[ic] = vc * (1.0f - J * rho);
    }
}

// src/svpf_optimized_graph.cu:701-717
float smoothed_output(SVPFState* st, float raw) {
    if (!st->use_smoothing) return raw;
    int k = std::min(st->smooth_lag, SVPF_SMOOTH_MAX_LAG);
    if (st->timestep < k) return raw;
    int lag = st->smooth_output_lag;
    if (lag <= 0) return raw;
    if (lag >= k) lag = k - 1;
    int idx = (st->smooth_head - lag - 1 + k) % k;
    return st->smooth_h_mean[idx];
}

// Host tail of one tick after the record came back.
void finish_tick(SVPFState* s, const SVPFParams* p, float y_t, const float* rec, float* ll, float* vol, float* hm,
                 float mu_before) {
    svpf_impl* I = s->impl;
    std::memcpy(I->out_last, rec, sizeof(I->out_last));
    float h_mean = rec[SVPF_O_HMEAN], bw = rec[SVPF_O_BW];
    // smoothing reads the pre-increment timestep, as the reference does
    // the two-factor branch returns before the smoothing (src/svpf_optimized_graph.cu:924-928)
    int ts = s->timestep;
    float out_h = h_mean;
    if (!s->use_two_factor) {
        s->timestep = ts - 1;
        smooth_backward(s, h_mean, bw * bw, y_t, p, mu_before);
        out_h = smoothed_output(s, h_mean);
        s->timestep = ts;
    }
    if (ll) *ll = rec[SVPF_O_LOGLIK];
    if (vol) *vol = rec[SVPF_O_VOL];
    if (hm) *hm = out_h;
}

void step_impl(SVPFState* s, float y_t, float y_prev, const SVPFParams* p, float* ll, float* vol, float* hm,
               bool use_graph, uint64_t seed_override, bool seeded) {
    if (!s || !s->impl) return;
    svpf_impl* I = s->impl;
    if (!p) {
        set_err(s, SVPF_RC_INVALID_ARGUMENT);
        return;
    }
    if (!I->initialized) {
        set_err(s, SVPF_RC_NOT_INITIALIZED);
        return;
    }
    int rc = check_supported(s);
    if (rc) {
        set_err(s, rc);
        return;
    }
    pack_params(s, p, I->hp());
    pack_fstate(s, y_prev, I->hs());
    I->hy()[0] = y_t;
    uint64_t saved = I->seed;
    if (seeded) I->seed = seed_override;
    fill_desc(s, I->hd(), I->dy(), I->dout(), s->timestep);
    I->seed = saved;
    hipStream_t cs = s->stream;
    hipError_t e = hipSuccess;
    const bool coop = use_coop(s);
    const int path = step_path(s);
    // the captured launch sequence depends on the path and, on the grid path, on stein_max_steps
    if (I->captured && (I->captured_path != path || I->captured_2f != (s->use_two_factor != 0) ||
                        (path == SVPF_PATH_GRID && (I->captured_smax != s->stein_max_steps ||
                                                    I->captured_ext != grid_ext(s))) ||
                        (path == SVPF_PATH_COOP && I->captured_cooperative != I->coop_cooperative)))
        svpf_graph_invalidate(s);
    if (use_graph && coop && I->coop_graph_failed) use_graph = false;
    if (use_graph) {
        if (!I->captured) {
            hipStream_t cap = nullptr;
            e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e1 = enqueue_tick(s, cap);
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(cap, &g);
                if (e1 != hipSuccess) e = e1;
                if (e == hipSuccess) e = hipGraphInstantiate(&I->exec, g, nullptr, nullptr, 0);
                if (e == hipSuccess) {
                    I->graph = g;
                    I->captured = true;
                    I->captured_coop = coop;
                    I->captured_path = path;
                    I->captured_smax = s->stein_max_steps;
                    I->captured_cooperative = I->coop_cooperative;
                    I->captured_ext = grid_ext(s);
                    I->captured_2f = s->use_two_factor != 0;
                } else if (g) {
                    hipGraphDestroy(g);
                }
            }
            if (cap) hipStreamDestroy(cap);
            if (e != hipSuccess && coop) {
                // cooperative launches that cannot be captured run as plain stream work
                (void)hipGetLastError();
                I->coop_graph_failed = true;
                use_graph = false;
                e = hipSuccess;
            }
        }
    }
    if (e == hipSuccess) e = use_graph ? hipGraphLaunch(I->exec, cs) : enqueue_tick(s, cs);
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
    if (e == hipSuccess && I->hs()->status == 2 && coop && !I->coop_cooperative) {
        // a plain launch's exchange timed out; repeat the tick with a cooperative launch, on
        // fresh exchange epochs
        I->coop_cooperative = true;
        I->epoch += epochs_after_timeout(s, 1);
        pack_fstate(s, y_prev, I->hs());
        svpf_graph_invalidate(s);
        e = enqueue_tick(s, cs);
        if (e == hipSuccess) e = hipStreamSynchronize(cs);
    }
    if (e != hipSuccess) {
        I->last_hip = e;
        set_err(s, SVPF_RC_HIP);
        return;
    }
    if (I->hs()->status) {
        I->last_status = I->hs()->status;
        set_err(s, I->hs()->status == 2 ? SVPF_RC_TIMEOUT : SVPF_RC_UNSUPPORTED);
        return;
    }
    const float mu_before = s->mu_state;
    unpack_fstate(s, I->hs());
    finish_tick(s, p, y_t, I->hout(), ll, vol, hm, mu_before);
    set_err(s, SVPF_RC_OK);
}

// T ticks in one launch; records land in records[T][8] (null: the state's own
// buffer I->dseq_out, grown with hipMalloc when T exceeds it).
int run_device(SVPFState* s, const float* d_obs, int T, const SVPFParams* p, float y_prev0,
               float* records = nullptr) {
    svpf_impl* I = s->impl;
    if (!I->initialized) return SVPF_RC_NOT_INITIALIZED;
    int rc = check_supported(s);
    if (rc) return rc;
    if (T <= 0) return SVPF_RC_OK;
    if (!records) {
        if (I->dseq_cap < T) {
            if (I->dseq_out) hipFree(I->dseq_out);
            I->dseq_out = nullptr;
            if (hipMalloc(&I->dseq_out, (size_t)T * SVPF_OUTW * sizeof(float)) != hipSuccess)
                return SVPF_RC_OUT_OF_MEMORY;
            I->dseq_cap = T;
        }
        records = I->dseq_out;
    }
    pack_params(s, p, I->hp());
    pack_fstate(s, y_prev0, I->hs());
    fill_desc(s, I->hd(), d_obs, records, s->timestep);
    hipStream_t cs = s->stream;
    const bool coop = use_coop(s);
    hipError_t e = hipSuccess;
    for (int attempt = 0; attempt < 2 && e == hipSuccess; ++attempt) {
        e = hipMemcpyAsync(I->dblock, I->hblock, kUploadBytes, hipMemcpyHostToDevice, cs);
        if (e == hipSuccess) e = launch_step_kernel(s, 0, T, cs);
        if (e == hipSuccess)
            e = hipMemcpyAsync(I->hblock + kBlockFstate, I->dblock + kBlockFstate, sizeof(svpf_fstate),
                               hipMemcpyDeviceToHost, cs);
        if (e == hipSuccess) e = hipStreamSynchronize(cs);
        // a plain launch's exchange timed out (nothing was written back): repeat the run with a
        // cooperative launch, on fresh exchange epochs
        if (e != hipSuccess || I->hs()->status != 2 || !coop || I->coop_cooperative) break;
        I->coop_cooperative = true;
        I->epoch += epochs_after_timeout(s, T);
        pack_fstate(s, y_prev0, I->hs());
        svpf_graph_invalidate(s);
    }
    if (e != hipSuccess) {
        I->last_hip = e;
        return SVPF_RC_HIP;
    }
    if (I->hs()->status) {
        I->last_status = I->hs()->status;
        return I->hs()->status == 2 ? SVPF_RC_TIMEOUT : SVPF_RC_UNSUPPORTED;
    }
    unpack_fstate(s, I->hs());
    return SVPF_RC_OK;
}

}  // namespace

extern "C" {

SVPFState* svpf_create(int n_particles, int n_stein_steps, float nu, hipStream_t stream) {
    if (n_particles < 1 || nu <= 0.0f) return nullptr;
    int tile = -1, bs = 256;
    // one workgroup per filter while its LDS holds the filter, else the grid path only
    const bool grid_only = !choose_tile(n_particles, &tile, &bs);
    if (n_particles > SVPF_GRID_MAX_PARTICLES) return nullptr;
    const bool grid = n_particles >= 2;
    if (grid && svpf_grid_init() != hipSuccess) return nullptr;
    if (svpf_kernels_init() != hipSuccess) return nullptr;
    SVPFState* s = static_cast<SVPFState*>(std::calloc(1, sizeof(SVPFState)));
    if (!s) return nullptr;
    svpf_impl* I = new (std::nothrow) svpf_impl();
    if (!I) {
        std::free(s);
        return nullptr;
    }
    s->impl = I;
    I->n = n_particles;
    I->tile = tile;
    I->bs = bs;
    bool ok = hipMalloc(&I->dblock, kBlockBytes) == hipSuccess;
    if (grid) ok = ok && hipMalloc(&I->dgrid_ws, svpf_grid_workspace_floats(n_particles) * sizeof(float)) == hipSuccess;
    ok = ok && hipHostMalloc(&I->hblock, kBlockBytes, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMalloc(&I->dpool, 5 * (size_t)n_particles * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&I->dorder, sizeof(int)) == hipSuccess;
    if (!grid_only) ok = ok && hipMalloc(&I->d2f, 6 * (size_t)n_particles * sizeof(float)) == hipSuccess;
    // exchange words of the cooperative path: [2][4][n] records, [2][n] final positions,
    // [2][G][8] partials with G <= n / 8 + 1
    const size_t xwords = svpf_coop_xbuf_words_host(n_particles);
    ok = ok && hipMalloc(&I->xbuf, xwords * sizeof(uint64_t)) == hipSuccess;
    if (ok) {
        std::memset(I->hblock, 0, kBlockBytes);
        int zero = 0;
        ok = hipMemcpy(I->dorder, &zero, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemset(I->dpool, 0, 5 * (size_t)n_particles * sizeof(float)) == hipSuccess;
        ok = ok && hipMemset(I->dblock, 0, kBlockBytes) == hipSuccess;
        ok = ok && hipMemset(I->xbuf, 0, xwords * sizeof(uint64_t)) == hipSuccess;
        if (I->d2f) ok = ok && hipMemset(I->d2f, 0, 6 * (size_t)n_particles * sizeof(float)) == hipSuccess;
    }
    if (ok && !grid_only) choose_coop_shape(I, -1, -1);
    if (!ok) {
        svpf_destroy(s);
        return nullptr;
    }
    s->h = I->d_h();
    s->h_prev = I->d_hp();
    s->h_pred = I->d_hpred();
    s->grad_log_p = nullptr;
    s->log_weights = I->d_lw();
    s->d_grad_v = I->d_v();

    // src/svpf_optimized_graph.cu:110-302
    s->n_particles = n_particles;
    s->n_stein_steps = n_stein_steps;
    s->nu = nu;
    s->stream = stream;
    s->timestep = 0;
    s->y_prev = 0.0f;
    s->student_t_const = student_t_const(nu);
    s->student_t_implied_offset = implied_offset(nu);
    s->lik_offset = 0.45f;
    s->use_exact_gradient = 0;
    s->use_svld = 1;
    s->use_annealing = 1;
    s->use_adaptive_beta = 1;
    s->n_anneal_steps = 3;
    s->temperature = 1.0f;
    s->rmsprop_rho = 0.9f;
    s->rmsprop_eps = 1e-6f;
    s->use_mim = 1;
    s->mim_jump_prob = 0.05f;
    s->mim_jump_scale = 5.0f;
    s->use_asymmetric_rho = 1;
    s->rho_up = 0.98f;
    s->rho_down = 0.93f;
    s->use_guide = 1;
    s->use_guide_preserving = 1;
    s->guide_strength = 0.05f;
    s->use_adaptive_guide = 0;
    s->guide_strength_base = 0.05f;
    s->guide_strength_max = 0.30f;
    s->guide_innovation_threshold = 1.0f;
    s->vol_prev = 0.05f;
    s->use_rejuvenation = 1;
    s->rejuv_ksd_threshold = 0.30f;
    s->rejuv_prob = 0.30f;
    s->rejuv_blend = 0.30f;
    s->use_newton = 0;
    s->use_full_newton = 0;
    s->use_guided = 0;
    s->guided_alpha_base = 0.0f;
    s->guided_alpha_shock = 0.5f;
    s->guided_innovation_threshold = 1.5f;
    s->use_local_params = 0;
    s->delta_rho = 0.02f;
    s->delta_sigma = 0.1f;
    s->use_adaptive_mu = 0;
    s->mu_state = -3.5f;
    s->mu_var = 1.0f;
    s->mu_process_var = 0.001f;
    s->mu_obs_var_scale = 10.0f;
    s->mu_min = -6.0f;
    s->mu_max = -1.0f;
    s->use_adaptive_sigma = 0;
    s->sigma_boost_threshold = 1.0f;
    s->sigma_boost_max = 3.0f;
    s->sigma_z_effective = 0.10f;
    s->stein_repulsive_sign = SVPF_STEIN_SIGN_DEFAULT;
    s->use_fan_mode = 0;
    s->use_student_t_state = 0;
    s->nu_state = 5.0f;
    s->stein_min_steps = 4;
    s->stein_max_steps = 12;
    s->ksd_improvement_threshold = 0.05f;
    s->ksd_prev = 1e10f;
    s->stein_steps_used = n_stein_steps;
    s->use_heun = 0;
    // src/svpf_optimized_graph.cu:260-272
    s->use_two_factor = 0;
    s->rho_fast = 0.90f;
    s->sigma_fast = 0.15f;
    s->rho_slow = 0.99f;
    s->sigma_slow = 0.05f;
    s->bw_floor_fast = 0.01f;
    s->bw_floor_slow = 0.05f;
    s->guide_mean_fast = 0.0f;
    s->guide_mean_slow = 0.0f;
    s->guide_var_fast = 0.0f;
    s->guide_var_slow = 0.0f;
    s->guide_2f_initialized = 0;
    s->use_smoothing = 0;
    s->smooth_lag = 3;
    s->smooth_output_lag = 1;
    for (int i = 0; i < SVPF_SMOOTH_MAX_LAG; ++i) {
        s->smooth_h_mean[i] = 0.0f;
        s->smooth_h_var[i] = 1.0f;
        s->smooth_y[i] = 0.0f;
    }
    s->smooth_head = 0;
    s->use_persistent_kernel = (n_particles >= 256 && I->coop_ok) ? 1 : 0;
    s->persistent_kernel_supported = 1;
    s->bandwidth_mode = SVPF_BANDWIDTH_REFERENCE_EMA;
    s->use_linear_beta = 0;
    s->use_ksd_early_stop = 0;
    return s;
}

void svpf_destroy(SVPFState* s) {
    if (!s) return;
    svpf_impl* I = s->impl;
    if (I) {
        if (I->exec) hipGraphExecDestroy(I->exec);
        if (I->graph) hipGraphDestroy(I->graph);
        if (I->dblock) hipFree(I->dblock);
        if (I->hblock) hipHostFree(I->hblock);
        if (I->dpool) hipFree(I->dpool);
        if (I->dorder) hipFree(I->dorder);
        if (I->xbuf) hipFree(I->xbuf);
        if (I->dgrid_ws) hipFree(I->dgrid_ws);
        if (I->d2f) hipFree(I->d2f);
        if (I->dseq_out) hipFree(I->dseq_out);
        if (I->dseq_y) hipFree(I->dseq_y);
        delete I;
    }
    std::free(s);
}

// src/svpf_optimized_graph.cu:349-451
void svpf_initialize(SVPFState* s, const SVPFParams* p, unsigned long long seed) {
    if (!s || !s->impl) return;
    svpf_impl* I = s->impl;
    if (!p) {
        set_err(s, SVPF_RC_INVALID_ARGUMENT);
        return;
    }
    s->timestep = 0;
    s->y_prev = 0.0f;
    I->seed = seed;
    // nu_state >= 2.5 keeps the Student-t state variance finite; its stationary
    // variance is nu / (nu - 2) times the Gaussian one (src/svpf_optimized_graph.cu:366-383)
    s->nu_state = fmaxf(s->nu_state, 2.5f);
    float base_var = p->sigma_z * p->sigma_z / (1.0f - p->rho * p->rho + 1e-6f);
    if (s->use_student_t_state) base_var *= s->nu_state / (s->nu_state - 2.0f);
    float sd = sqrtf(base_var);
    hipError_t e = svpf_launch_init(I->d_h(), I->d_hp(), I->d_v(), I->n, p->mu, sd, seed, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        set_err(s, SVPF_RC_HIP);
        return;
    }
    s->guide_initialized = 0;
    s->guide_mean = p->mu;
    s->guide_var = base_var;
    if (s->use_adaptive_mu) {
        s->mu_state = p->mu;
        s->mu_var = 1.0f;
    }
    if (s->use_adaptive_sigma) s->sigma_z_effective = p->sigma_z;
    s->ksd_prev = 1e10f;
    s->stein_steps_used = s->n_stein_steps;
    I->med_r2 = 0.0f;
    if (s->use_two_factor && I->d2f) {
        // src/svpf_optimized_graph.cu:414-448
        e = svpf_launch_2f_init(I->d2f, I->n, s->rho_fast, s->sigma_fast, s->rho_slow, s->sigma_slow, seed, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            set_err(s, SVPF_RC_HIP);
            return;
        }
        s->guide_2f_initialized = 0;
        s->guide_mean_fast = 0.0f;
        s->guide_mean_slow = 0.0f;
        s->guide_var_fast = s->sigma_fast * s->sigma_fast / (1.0f - s->rho_fast * s->rho_fast + 1e-6f);
        s->guide_var_slow = s->sigma_slow * s->sigma_slow / (1.0f - s->rho_slow * s->rho_slow + 1e-6f);
        I->mean_fast_prev = 0.0f;
        I->mean_slow_prev = 0.0f;
    }
    svpf_graph_invalidate(s);
    I->initialized = true;
    set_err(s, SVPF_RC_OK);
}

void svpf_step_graph(SVPFState* s, float y_t, float y_prev, const SVPFParams* p, float* ll, float* vol, float* hm) {
    step_impl(s, y_t, y_prev, p, ll, vol, hm, true, 0, false);
}

void svpf_step_adaptive(SVPFState* s, float y_t, float y_prev, const SVPFParams* p, float* ll, float* vol,
                        float* hm) {
    step_impl(s, y_t, y_prev, p, ll, vol, hm, true, 0, false);
}

void svpf_step(SVPFState* s, float y_t, const SVPFParams* p, SVPFResult* r) {
    if (!s) return;
    float ll = 0.f, vol = 0.f, hm = 0.f;
    step_impl(s, y_t, s->y_prev, p, &ll, &vol, &hm, true, 0, false);
    if (r) {
        r->log_lik_increment = ll;
        r->vol_mean = vol;
        r->h_mean = hm;
        r->vol_std = 0.0f;
        r->mu_estimate = (s->use_adaptive_mu || !p) ? s->mu_state : p->mu;
    }
    s->y_prev = y_t;
}

void svpf_step_seeded(SVPFState* s, float y_t, const SVPFParams* p, unsigned long long seed, SVPFResult* r) {
    if (!s) return;
    float ll = 0.f, vol = 0.f, hm = 0.f;
    step_impl(s, y_t, s->y_prev, p, &ll, &vol, &hm, false, seed, true);
    if (r) {
        r->log_lik_increment = ll;
        r->vol_mean = vol;
        r->h_mean = hm;
        r->vol_std = 0.0f;
        r->mu_estimate = (s->use_adaptive_mu || !p) ? s->mu_state : p->mu;
    }
    s->y_prev = y_t;
}

void svpf_run_sequence(SVPFState* s, const float* obs, int T, const SVPFParams* p, float* ll, float* vol) {
    if (!s || !s->impl) return;
    svpf_impl* I = s->impl;
    if (!obs || !p || T < 0) {
        set_err(s, SVPF_RC_INVALID_ARGUMENT);
        return;
    }
    if (T == 0) {
        set_err(s, SVPF_RC_OK);
        return;
    }
    if (s->use_smoothing && s->use_adaptive_mu && !s->use_two_factor) {
        // the smoothing of each tick reads that tick's adaptive mean, which one device run
        // does not report: step tick by tick
        float yp = 0.0f;  // src/svpf_optimized_graph.cu:1414-1424
        for (int t = 0; t < T; ++t) {
            step_impl(s, obs[t], yp, p, ll ? &ll[t] : nullptr, vol ? &vol[t] : nullptr, nullptr, true, 0, false);
            if (I->last_error != SVPF_RC_OK) return;
            yp = obs[t];
        }
        return;
    }
    if (I->dseq_y_cap < T) {
        if (I->dseq_y) hipFree(I->dseq_y);
        I->dseq_y = nullptr;
        if (hipMalloc(&I->dseq_y, (size_t)T * sizeof(float)) != hipSuccess) {
            set_err(s, SVPF_RC_OUT_OF_MEMORY);
            return;
        }
        I->dseq_y_cap = T;
    }
    if (hipMemcpy(I->dseq_y, obs, (size_t)T * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        set_err(s, SVPF_RC_HIP);
        return;
    }
    int t0 = s->timestep;
    int rc = run_device(s, I->dseq_y, T, p, 0.0f);
    if (rc) {
        set_err(s, rc);
        return;
    }
    std::vector<float> rec((size_t)T * SVPF_OUTW);
    if (hipMemcpy(rec.data(), I->dseq_out, rec.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        set_err(s, SVPF_RC_HIP);
        return;
    }
    int tend = s->timestep;
    for (int t = 0; t < T; ++t) {
        const float* r = &rec[(size_t)t * SVPF_OUTW];
        s->timestep = t0 + t + 1;
        finish_tick(s, p, obs[t], r, ll ? &ll[t] : nullptr, vol ? &vol[t] : nullptr, nullptr, s->mu_state);
    }
    s->timestep = tend;
    set_err(s, SVPF_RC_OK);
}

}  // extern "C"

// svpf_run_sequence_device with the per-tick records in `records` (null: the
// state's own buffer).  Also the body of the allocation-free variant.
void svpf_internal_run_sequence_device(SVPFState* s, const float* d_obs, int T, const SVPFParams* p, float* d_ll,
                                       float* d_vol, float* records) {
    if (!s || !s->impl) return;
    svpf_impl* I = s->impl;
    if (!d_obs || !p || T < 0) {
        set_err(s, SVPF_RC_INVALID_ARGUMENT);
        return;
    }
    int rc = run_device(s, d_obs, T, p, 0.0f, records);
    if (rc) {
        set_err(s, rc);
        return;
    }
    if (!records) records = I->dseq_out;
    hipError_t e = hipSuccess;
    if (d_ll && T > 0)
        e = hipMemcpy2DAsync(d_ll, sizeof(float), records + SVPF_O_LOGLIK, SVPF_OUTW * sizeof(float), sizeof(float), T,
                             hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess && d_vol && T > 0)
        e = hipMemcpy2DAsync(d_vol, sizeof(float), records + SVPF_O_VOL, SVPF_OUTW * sizeof(float), sizeof(float), T,
                             hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && T > 0) {
        float last[SVPF_OUTW];
        e = hipMemcpy(last, records + (size_t)(T - 1) * SVPF_OUTW, sizeof(last), hipMemcpyDeviceToHost);
        if (e == hipSuccess) std::memcpy(I->out_last, last, sizeof(last));
    }
    set_err(s, e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP);
}

extern "C" {

void svpf_run_sequence_device(SVPFState* s, const float* d_obs, int T, const SVPFParams* p, float* d_ll,
                              float* d_vol) {
    svpf_internal_run_sequence_device(s, d_obs, T, p, d_ll, d_vol, nullptr);
}

bool svpf_graph_is_captured(SVPFState* s) { return s && s->impl && s->impl->captured; }

void svpf_graph_invalidate(SVPFState* s) {
    if (!s || !s->impl) return;
    svpf_impl* I = s->impl;
    if (I->exec) hipGraphExecDestroy(I->exec);
    if (I->graph) hipGraphDestroy(I->graph);
    I->exec = nullptr;
    I->graph = nullptr;
    I->captured = false;
}

void svpf_get_particles(const SVPFState* s, float* h_out) {
    if (!s || !s->impl || !h_out) return;
    hipStreamSynchronize(s->stream);
    hipError_t e = hipMemcpy(h_out, s->impl->d_h(), (size_t)s->impl->n * sizeof(float), hipMemcpyDeviceToHost);
    s->impl->last_error = e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

void svpf_get_stats(const SVPFState* s, float* h_mean, float* h_std) {
    if (!s || !s->impl) return;
    int n = s->impl->n;
    std::vector<float> h(n);
    svpf_get_particles(s, h.data());
    float sum = 0.f;
    for (int i = 0; i < n; ++i) sum += h[i];
    float m = sum / (float)n;
    float sq = 0.f;
    for (int i = 0; i < n; ++i) sq += (h[i] - m) * (h[i] - m);
    if (h_mean) *h_mean = m;
    if (h_std) *h_std = sqrtf(sq / (float)n);
}

// src/svpf_optimized_graph.cu:1470-1484
float svpf_get_ess(const SVPFState* s) {
    if (!s || !s->impl) return 0.0f;
    int n = s->impl->n;
    std::vector<float> lw(n);
    hipStreamSynchronize(s->stream);
    if (hipMemcpy(lw.data(), s->impl->d_lw(), (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        s->impl->last_error = SVPF_RC_HIP;
        return 0.0f;
    }
    float mx = lw[0];
    for (int i = 1; i < n; ++i) mx = std::max(mx, lw[i]);
    float sw = 0.f, sw2 = 0.f;
    for (int i = 0; i < n; ++i) {
        float w = expf(lw[i] - mx);
        sw += w;
        sw2 += w * w;
    }
    return (sw * sw) / (sw2 + 1e-10f);
}

void svpf_get_ksd_stats(const SVPFState* s, float* ksd_out, int* steps_used_out) {
    if (!s) return;
    if (ksd_out) *ksd_out = s->ksd_prev;
    if (steps_used_out) *steps_used_out = s->stein_steps_used;
}

int svpf_last_error(const SVPFState* s) { return (s && s->impl) ? s->impl->last_error : SVPF_RC_INVALID_ARGUMENT; }

const char* svpf_last_error_string(const SVPFState* s) {
    if (!s || !s->impl) return "invalid state";
    const svpf_impl* I = s->impl;
    switch (I->last_error) {
        case SVPF_RC_OK: return "ok";
        case SVPF_RC_INVALID_ARGUMENT: return "invalid argument";
        case SVPF_RC_UNSUPPORTED: return "unsupported feature";
        case SVPF_RC_HIP: return hipGetErrorString(I->last_hip);
        case SVPF_RC_NOT_INITIALIZED: return "filter not initialized";
        case SVPF_RC_OUT_OF_MEMORY: return "out of memory";
        case SVPF_RC_TIMEOUT: return "cooperative exchange timed out";
        case SVPF_RC_TOO_LARGE: return "request exceeds the allocation-free entry point's fixed scratch";
        default: return "unknown";
    }
}

int svpf_set_particles(SVPFState* s, const float* h, const float* h_prev, const float* v) {
    if (!s || !s->impl || !h) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    svpf_impl* I = s->impl;
    size_t nb = (size_t)I->n * sizeof(float);
    hipStreamSynchronize(s->stream);
    hipError_t e = hipMemcpy(I->d_h(), h, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(I->d_hp(), h_prev ? h_prev : h, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (v) e = hipMemcpy(I->d_v(), v, nb, hipMemcpyHostToDevice);
        else e = hipMemset(I->d_v(), 0, nb);
    }
    if (e != hipSuccess) return set_err(s, SVPF_RC_HIP);
    I->initialized = true;
    return set_err(s, SVPF_RC_OK);
}

int svpf_get_rmsprop_state(const SVPFState* s, float* v_out) {
    if (!s || !s->impl || !v_out) return SVPF_RC_INVALID_ARGUMENT;
    hipStreamSynchronize(s->stream);
    return hipMemcpy(v_out, s->impl->d_v(), (size_t)s->impl->n * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
               ? SVPF_RC_OK
               : SVPF_RC_HIP;
}

int svpf_set_predrawn_noise(SVPFState* s, const float* z_pred, const float* u_pred, const float* xi, int xi_rows,
                            const float* u_rej, const float* z_rej, int t_base) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    svpf_impl* I = s->impl;
    if (z_pred && (!xi || xi_rows < 0)) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    I->z_pred = z_pred;
    I->u_pred = z_pred ? u_pred : nullptr;
    I->xi = z_pred ? xi : nullptr;
    I->xi_rows = z_pred ? xi_rows : 0;
    I->u_rej = z_pred ? u_rej : nullptr;
    I->z_rej = z_pred ? z_rej : nullptr;
    I->t_base = t_base;
    svpf_graph_invalidate(s);
    return set_err(s, SVPF_RC_OK);
}

int svpf_set_scalar_state(SVPFState* s, float h_mean_prev, float bw_sq_ema, float return_ema, float return_var) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    s->impl->h_mean_prev = h_mean_prev;
    s->impl->bw_sq_ema = bw_sq_ema;
    s->impl->return_ema = return_ema;
    s->impl->return_var = return_var;
    return set_err(s, SVPF_RC_OK);
}

int svpf_get_two_factor_particles(const SVPFState* s, float* h_fast, float* h_slow) {
    if (!s || !s->impl || !s->impl->d2f) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_impl* I = s->impl;
    const size_t nb = (size_t)I->n * sizeof(float);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && h_fast) e = hipMemcpy(h_fast, I->d2f, nb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && h_slow) e = hipMemcpy(h_slow, I->d2f + I->n, nb, hipMemcpyDeviceToHost);
    return e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

int svpf_get_last_outputs(const SVPFState* s, float out8[8]) {
    if (!s || !s->impl || !out8) return SVPF_RC_INVALID_ARGUMENT;
    std::memcpy(out8, s->impl->out_last, 8 * sizeof(float));
    return SVPF_RC_OK;
}

int svpf_batched_run(const void* d_params, void* d_fstate, const void* d_desc, float* d_pool_h, float* d_pool_hp,
                     float* d_pool_v, const int* d_order, const int* group_start, const int* group_tile,
                     const int* group_bs, const int* group_grid, const int* group_nmax, int n_groups, int* d_queues,
                     int t0, int t1, hipStream_t stream) {
    if (!d_params || !d_fstate || !d_desc || !d_pool_h || !d_pool_hp || !d_pool_v || !d_order || !group_start ||
        !group_tile || !group_bs || !group_grid || !group_nmax || n_groups < 0 || t1 < t0)
        return SVPF_RC_INVALID_ARGUMENT;
    if (svpf_kernels_init() != hipSuccess) return SVPF_RC_HIP;
    int used = 0;
    for (int g = 0; g < n_groups; ++g) {
        int nf = group_start[g + 1] - group_start[g];
        if (nf <= 0) continue;
        int cap = svpf_tile_capacity(group_tile[g], group_bs[g]);
        if (cap <= 0 || group_nmax[g] > cap || group_grid[g] <= 0) return SVPF_RC_INVALID_ARGUMENT;
        if (svpf_lds_bytes(group_nmax[g]) > 159 * 1024) return SVPF_RC_UNSUPPORTED;
        ++used;
    }
    if (used > kMaxForkStreams) return SVPF_RC_INVALID_ARGUMENT;
    // Tile groups run concurrently: fork from `stream` onto worker streams and
    // join back, so the call stays one ordered operation on `stream` (and is
    // capturable into a graph).
    ForkJoin* fj = used > 1 ? fork_join() : nullptr;
    if (used > 1 && !fj) return SVPF_RC_HIP;
    if (fj && hipEventRecord(fj->fork, stream) != hipSuccess) return SVPF_RC_HIP;
    int k = 0;
    for (int g = 0; g < n_groups; ++g) {
        int nf = group_start[g + 1] - group_start[g];
        if (nf <= 0) continue;
        hipStream_t s = stream;
        if (fj) {
            s = fj->streams[k];
            if (hipStreamWaitEvent(s, fj->fork, 0) != hipSuccess) return SVPF_RC_HIP;
        }
        hipError_t e = svpf_launch_batched(group_bs[g], group_tile[g], group_grid[g],
                                           static_cast<const svpf_fparams*>(d_params),
                                           static_cast<svpf_fstate*>(d_fstate), static_cast<const svpf_fdesc*>(d_desc),
                                           d_pool_h, d_pool_hp, d_pool_v, d_order + group_start[g], nf,
                                           d_queues ? d_queues + g : nullptr,
                                           d_queues ? d_queues + n_groups + group_start[g] : nullptr, group_nmax[g],
                                           t0, t1, s);
        if (e != hipSuccess) return SVPF_RC_HIP;
        if (fj) {
            if (hipEventRecord(fj->join[k], s) != hipSuccess) return SVPF_RC_HIP;
            if (hipStreamWaitEvent(stream, fj->join[k], 0) != hipSuccess) return SVPF_RC_HIP;
        }
        ++k;
    }
    return SVPF_RC_OK;
}

int svpf_set_coop_shape(SVPFState* s, int lanes_per_row, int xcd_stride) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    svpf_impl* I = s->impl;
    if (I->tile < 0) return set_err(s, SVPF_RC_UNSUPPORTED);
    svpf_impl trial = {};
    trial.n = I->n;
    if (!choose_coop_shape(&trial, lanes_per_row, xcd_stride)) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    I->coop_cs = trial.coop_cs;
    I->coop_G = trial.coop_G;
    I->coop_stride = trial.coop_stride;
    I->coop_ok = true;
    svpf_graph_invalidate(s);
    return set_err(s, SVPF_RC_OK);
}

int svpf_set_coop_median_workgroup(SVPFState* s, int on) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    if (on != 0 && on != 1) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    s->impl->coop_bwg = on != 0;
    svpf_graph_invalidate(s);
    return set_err(s, SVPF_RC_OK);
}

int svpf_coop_launch_is_cooperative(const SVPFState* s) {
    if (!s || !s->impl) return -1;
    return s->impl->coop_cooperative ? 1 : 0;
}

int svpf_set_coop_spin_limit(SVPFState* s, int polls) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    if (polls < -(1 << 20) || polls > (1 << 24)) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    s->impl->coop_spin_limit = polls;
    return set_err(s, SVPF_RC_OK);
}

int svpf_set_coop_fault(SVPFState* s, int workgroup) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    if (workgroup < -1 || workgroup >= s->impl->coop_G) return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    s->impl->coop_fault_wg = workgroup;
    return set_err(s, SVPF_RC_OK);
}

int svpf_set_grid_path(SVPFState* s, int mode) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    svpf_impl* I = s->impl;
    if (mode != SVPF_GRID_AUTO && mode != SVPF_GRID_OFF && mode != SVPF_GRID_ON)
        return set_err(s, SVPF_RC_INVALID_ARGUMENT);
    if ((mode == SVPF_GRID_OFF && I->tile < 0) || (mode == SVPF_GRID_ON && !I->dgrid_ws))
        return set_err(s, SVPF_RC_UNSUPPORTED);
    I->grid_mode = mode;
    return set_err(s, SVPF_RC_OK);
}

int svpf_get_path(const SVPFState* s) {
    if (!s || !s->impl) return -1;
    return step_path(s);
}

int svpf_get_coop_shape(const SVPFState* s, int* lanes_per_row, int* workgroups, int* xcd_stride) {
    if (!s || !s->impl) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_impl* I = s->impl;
    if (lanes_per_row) *lanes_per_row = I->coop_ok ? I->coop_cs : 0;
    if (workgroups) *workgroups = I->coop_ok ? I->coop_G : 0;
    if (xcd_stride) *xcd_stride = I->coop_ok ? I->coop_stride : 0;
    return SVPF_RC_OK;
}

size_t svpf_diff_tape_floats(int n_particles, int n_stein, int checkpoint_every) {
    if (n_particles < 1 || n_stein < 1 || checkpoint_every < 1) return 0;
    return svpf_diff_slot_floats_dev(n_particles, n_stein) * (size_t)checkpoint_every;
}

int svpf_diff_forward(const void* config, const void* d_desc, int B, hipStream_t stream) {
    if (!config || !d_desc || B < 1) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_dconfig* c = static_cast<const svpf_dconfig*>(config);
    if (c->n < 2 || c->n > 2048 || c->S < 1 || c->T < 1 || c->C < 1) return SVPF_RC_INVALID_ARGUMENT;
    hipError_t e = svpf_launch_diff(0, c, static_cast<const svpf_ddesc*>(d_desc), B, stream);
    return e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

int svpf_diff_backward(const void* config, const void* d_desc, int B, hipStream_t stream) {
    if (!config || !d_desc || B < 1) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_dconfig* c = static_cast<const svpf_dconfig*>(config);
    if (c->n < 2 || c->n > 2048 || c->S < 1 || c->T < 1 || c->C < 1) return SVPF_RC_INVALID_ARGUMENT;
    hipError_t e = svpf_launch_diff(1, c, static_cast<const svpf_ddesc*>(d_desc), B, stream);
    return e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

size_t svpf_diff2d_tape_floats(int n_particles, int n_stein, int checkpoint_every) {
    if (n_particles < 1 || n_stein < 1 || checkpoint_every < 1) return 0;
    return svpf_diff2d_slot_floats_dev(n_particles, n_stein) * (size_t)checkpoint_every;
}

int svpf_diff2d_forward(const void* config, const void* d_desc, int B, hipStream_t stream) {
    if (!config || !d_desc || B < 1) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_d2config* c = static_cast<const svpf_d2config*>(config);
    if (c->n < 2 || c->n > 2048 || c->S < 1 || c->T < 1 || c->C < 1) return SVPF_RC_INVALID_ARGUMENT;
    hipError_t e = svpf_launch_diff2d(0, c, static_cast<const svpf_d2desc*>(d_desc), B, stream);
    return e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

int svpf_diff2d_backward(const void* config, const void* d_desc, int B, hipStream_t stream) {
    if (!config || !d_desc || B < 1) return SVPF_RC_INVALID_ARGUMENT;
    const svpf_d2config* c = static_cast<const svpf_d2config*>(config);
    if (c->n < 2 || c->n > 2048 || c->S < 1 || c->T < 1 || c->C < 1) return SVPF_RC_INVALID_ARGUMENT;
    hipError_t e = svpf_launch_diff2d(1, c, static_cast<const svpf_d2desc*>(d_desc), B, stream);
    return e == hipSuccess ? SVPF_RC_OK : SVPF_RC_HIP;
}

int svpf_tile_capacity_query(int tile, int block_size) { return svpf_tile_capacity(tile, block_size); }
int svpf_tile_occupancy_query(int tile, int block_size, int nmax) { return svpf_batched_wg_per_cu(block_size, tile, nmax); }
size_t svpf_lds_bytes_query(int nmax) { return svpf_lds_bytes(nmax); }
const char* svpf_version(void) { return "svpf-mi355x 0.2.0 (gfx950)"; }
float svpf_digamma_f32(float x) { return svpf_digamma(x); }

}  // extern "C"

float* svpf_internal_diag_dev(const SVPFState* s) {
    return s && s->impl ? reinterpret_cast<float*>(s->impl->dblock + kBlockDiag) : nullptr;
}
float* svpf_internal_diag_host(const SVPFState* s) {
    return s && s->impl ? reinterpret_cast<float*>(s->impl->hblock + kBlockDiag) : nullptr;
}
void svpf_internal_set_error(const SVPFState* s, int rc, hipError_t e) {
    if (!s || !s->impl) return;
    s->impl->last_error = rc;
    if (rc == SVPF_RC_HIP) s->impl->last_hip = e;
}
