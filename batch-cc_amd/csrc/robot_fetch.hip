// Kernels specialised for the fetch model: the generated collision program
// (robots/fetch_check.inc) instantiated into the dense kernel template.
// Occupancy target of the dense kernel for this robot (waves per SIMD),
// chosen by measurement on MI355X; see build.py / tuning.
#ifndef BCC_DENSE_WAVES
#define BCC_DENSE_WAVES 6
#endif
#include "bcc_dense.h"

namespace bcc {
#include "robots/fetch_check.inc"
BCC_DEFINE_ROBOT(fetch, Robot_fetch)
}  // namespace bcc
