// Kernels specialised for the panda model: the generated collision program
// (robots/panda_check.inc) instantiated into the kernel templates, split into
// parts compiled separately (bcc_robot_parts.h).
#include "bcc_dense.h"

namespace bcc {
#include "robots/panda_check.inc"
}  // namespace bcc

// Kernel variants (waves per SIMD, obstacle-loop unroll of the high- and
// low-occupancy variants), measured on MI355X.
#ifndef BCC_PANDA_VARIANTS
#define BCC_PANDA_VARIANTS 6, 1, 4, 1
#endif
#define BCC_ROBOT panda
#define BCC_ROBOT_T Robot_panda
#define BCC_KNOBS BCC_PANDA_VARIANTS
#include "bcc_robot_parts.h"
