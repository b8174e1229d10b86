// svpf_kernels_ext.hip -- the batched kernel instantiated with the Heun
// predictor-corrector step and the Student-t state (tile ids with
// SVPF_TILE_EXT).  A translation unit of its own: the extra code paths raise the
// register count, which the common instantiations do not pay.
#include "svpf_batched.hpp"

namespace svpf {
template const void* batched_kernel<true>(int, int);
}
