// Kernels specialised for the baxter model: the generated collision program
// (robots/baxter_check.inc) instantiated into the dense kernel template.
// Occupancy target of the dense kernel for this robot (waves per SIMD),
// chosen by measurement on MI355X; see build.py / tuning.
#ifndef BCC_DENSE_WAVES
#define BCC_DENSE_WAVES 4
#endif
#include "bcc_dense.h"

namespace bcc {
#include "robots/baxter_check.inc"
BCC_DEFINE_ROBOT(baxter, Robot_baxter)
}  // namespace bcc
