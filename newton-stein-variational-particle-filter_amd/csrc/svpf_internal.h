// Host helpers shared between the C-ABI translation units (svpf_capi.cpp owns
// the state; svpf_diagnostics.cpp reads it).  Not part of the public headers.
#pragma once

#include "../../include/svpf.h"

constexpr int kDiagScratchFloats = 16;
// Device and pinned host scratch of kDiagScratchFloats floats inside the
// state's parameter block; null for a state without an implementation.
float* svpf_internal_diag_dev(const SVPFState* s);
float* svpf_internal_diag_host(const SVPFState* s);
// Record a failed call as the state's last error (SVPF_RC_*, HIP status).
void svpf_internal_set_error(const SVPFState* s, int rc, hipError_t e);
// svpf_run_sequence_device with the per-tick records [T][8] in `records` (device
// memory of at least 8 T floats; null: the state's own buffer, grown with hipMalloc).
void svpf_internal_run_sequence_device(SVPFState* s, const float* d_obs, int T, const SVPFParams* p, float* d_ll,
                                       float* d_vol, float* records);
