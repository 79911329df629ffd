// Kernels specialised for the panda model: the generated collision program
// (robots/panda_check.inc) instantiated into the dense kernel template.
// Occupancy target of the dense kernel for this robot (waves per SIMD),
// chosen by measurement on MI355X; see build.py / tuning.
#ifndef BCC_DENSE_WAVES
#define BCC_DENSE_WAVES 4
#endif
#include "bcc_dense.h"

namespace bcc {
#include "robots/panda_check.inc"
BCC_DEFINE_ROBOT(panda, Robot_panda)
}  // namespace bcc
