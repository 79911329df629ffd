// Kernels specialised for the baxter model: the generated collision program
// (robots/baxter_check.inc) instantiated into the dense kernel template.
#include "bcc_dense.h"

namespace bcc {
#include "robots/baxter_check.inc"
BCC_DEFINE_ROBOT(baxter, Robot_baxter)
}  // namespace bcc
