// Kernels specialised for the fetch model: the generated collision program
// (robots/fetch_check.inc) instantiated into the kernel templates, split into
// parts compiled separately (bcc_robot_parts.h).
#include "bcc_dense.h"

namespace bcc {
#include "robots/fetch_check.inc"
}  // namespace bcc

// Kernel variants (waves per SIMD, obstacle-loop unroll of the high- and
// low-occupancy variants), measured on MI355X.
#ifndef BCC_FETCH_VARIANTS
#define BCC_FETCH_VARIANTS 8, 1, 4, 1
#endif
#define BCC_ROBOT fetch
#define BCC_ROBOT_T Robot_fetch
#define BCC_KNOBS BCC_FETCH_VARIANTS
#include "bcc_robot_parts.h"
