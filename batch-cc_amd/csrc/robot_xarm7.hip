// Kernels specialised for the xarm7 model: the generated collision program
// (robots/xarm7_check.inc) instantiated into the kernel templates, split into
// parts compiled separately (bcc_robot_parts.h).
#include "bcc_dense.h"

namespace bcc {
#include "robots/xarm7_check.inc"
}  // namespace bcc

// Kernel variants (waves per SIMD, obstacle-loop unroll of the high- and
// low-occupancy variants): baxter's, the other robot with a long ungated
// self program.
#ifndef BCC_XARM7_VARIANTS
#define BCC_XARM7_VARIANTS 5, 1, 4, 1
#endif
#define BCC_ROBOT xarm7
#define BCC_ROBOT_T Robot_xarm7
#define BCC_KNOBS BCC_XARM7_VARIANTS
#include "bcc_robot_parts.h"
