// svpf_kernels.h -- host-side launch interface of the SVPF kernels.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "svpf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

// Tile shapes of the single-workgroup engine: R rows per lane, CS lanes per row.
enum {
    SVPF_TILE_R1_C4 = 0,
    SVPF_TILE_R1_C2 = 1,
    SVPF_TILE_R1_C1 = 2,
    SVPF_TILE_R2_C1 = 3,
    SVPF_TILE_R4_C1 = 4,
    SVPF_TILE_R8_C1 = 5,
    SVPF_TILE_R2_C2 = 6,
    SVPF_TILE_R4_C2 = 7,
    SVPF_NUM_TILES = 8,
    // ORed into a tile id: the instantiation with the Heun step and the Student-t
    // state, which filters using either need (without it they are rejected)
    SVPF_TILE_EXT = 0x100
};

size_t svpf_lds_bytes(int nmax);
hipError_t svpf_kernels_init(void);
hipError_t svpf_launch_init(float* h, float* hp, float* v, int n, float mu, float sd, uint64_t seed,
                            hipStream_t stream);
int svpf_tile_capacity(int tile, int bs);
int svpf_batched_wg_per_cu(int bs, int tile, int nmax);
hipError_t svpf_launch_batched(int bs, int tile, int grid, const svpf_fparams* params, svpf_fstate* fstate,
                               const svpf_fdesc* desc, float* pool_h, float* pool_hp, float* pool_v,
                               const int* order, int nf, int* queue, int* progress, int nmax, int t0, int t1,
                               hipStream_t stream);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Cooperative latency kernel for one filter (256 threads per workgroup, CS lanes per row).
int svpf_coop_rows(int cs);
size_t svpf_coop_lds_bytes(int n);
// 64-bit words of a filter's cooperative exchange buffer (records, positions, partials, commit stage).
size_t svpf_coop_xbuf_words_host(int n);
hipError_t svpf_launch_coop(int cs, int G, int stride, const svpf_fparams* params, svpf_fstate* fstate,
                            const svpf_fdesc* desc, float* pool_h, float* pool_hp, float* pool_v, uint64_t* xbuf,
                            int n, int t0, int t1, int cooperative, int bwg, hipStream_t stream);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Differentiable form: tape floats per tick slot, and the forward (backward = 0)
// or adjoint (backward = 1) launch over B filters, one workgroup each.
size_t svpf_diff_slot_floats_dev(int n, int S);
hipError_t svpf_launch_diff(int backward, const svpf_dconfig* C, const svpf_ddesc* desc, int B, hipStream_t stream);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Two-factor form: tape floats per tick slot, forward (backward = 0) or adjoint launch.
size_t svpf_diff2d_slot_floats_dev(int n, int S);
hipError_t svpf_launch_diff2d(int backward, const svpf_d2config* C, const svpf_d2desc* desc, int B,
                              hipStream_t stream);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Grid path for one large filter (n up to 16384): per tick a fixed sequence of
// grid-wide kernels on `stream`, with the Stein budget decided on the device.
size_t svpf_grid_workspace_floats(int n);
size_t svpf_grid_lds_bytes(int n);
hipError_t svpf_grid_init(void);
enum { SVPF_GRID_EXT_HEUN = 1, SVPF_GRID_EXT_STUDENT = 2 };
hipError_t svpf_launch_grid(const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D, float* h, float* hp,
                            float* v, float* lw, float* ws, int n, int stein_max, unsigned ext, int t0, int t1,
                            hipStream_t stream);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Diagnostics beside the step (one workgroup, O(n)).  HEALTH: out0[0..6] =
// NaN count, Inf count, sum w, sum w^2, min h, max h, sum h (finite h only).
// NU (a = y_t, b = nu) and SIGMA (h = h_pred, h2 = h_prev, a = mu, b = rho,
// c = sigma): *out0 = weighted mean score, *out1 = weighted mean z^2 or eps^2/sigma^2.
enum { SVPF_DIAG_HEALTH = 0, SVPF_DIAG_NU = 1, SVPF_DIAG_SIGMA = 2 };
hipError_t svpf_launch_diag(int mode, const float* h, const float* h2, const float* lw, int n, float a, float b,
                            float c, float* out0, float* out1, hipStream_t stream);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// The coordinate-wise two-factor step for one filter of n particles (one
// workgroup; the single-workgroup tile table), over ticks [t0, t1).  x2f:
// [6][n] = f | s | f_prev | s_prev | v_f | v_s; h (may be null): mu + f + s
// written at the end of the launch.
hipError_t svpf_launch_2f(int bs, int tile, const svpf_fparams* P, svpf_fstate* S, const svpf_fdesc* D, float* x2f,
                          float* h, int n, int t0, int t1, hipStream_t stream);
hipError_t svpf_launch_2f_init(float* x2f, int n, float rho_f, float sigma_f, float rho_s, float sigma_s,
                               uint64_t seed, hipStream_t stream);
// The joint state-parameter filter, one workgroup: mode 0 runs T steps at ticks
// t0 .. t0 + T - 1 on y[0 .. T-1] (on y1 when y is null and T = 1); mode 1 draws
// the initial particles.
hipError_t svpf_launch_joint(const svpf_jparams* P, float* xs, float* gl, float* summ, const float* y, float y1,
                             float* rec, int t0, int T, int mode, hipStream_t stream);
size_t svpf_joint_lds_bytes_host(int n);
#ifdef __cplusplus
}
#endif
