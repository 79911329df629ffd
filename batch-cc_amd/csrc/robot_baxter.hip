// Kernels specialised for the baxter model: the generated collision program
// (robots/baxter_check.inc) instantiated into the kernel templates, split into
// parts compiled separately (bcc_robot_parts.h).
#include "bcc_dense.h"

namespace bcc {
#include "robots/baxter_check.inc"
}  // namespace bcc

// Kernel variants (waves per SIMD, obstacle-loop unroll of the high- and
// low-occupancy variants), measured on MI355X.
#ifndef BCC_BAXTER_VARIANTS
#define BCC_BAXTER_VARIANTS 5, 1, 4, 1
#endif
#define BCC_ROBOT baxter
#define BCC_ROBOT_T Robot_baxter
#define BCC_KNOBS BCC_BAXTER_VARIANTS
#include "bcc_robot_parts.h"
