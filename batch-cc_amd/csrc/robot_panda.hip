// Kernels specialised for the panda model: the generated collision program
// (robots/panda_check.inc) instantiated into the dense kernel template.
#include "bcc_dense.h"

namespace bcc {
#include "robots/panda_check.inc"
BCC_DEFINE_ROBOT(panda, Robot_panda)
}  // namespace bcc
