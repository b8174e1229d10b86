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
