// Kernels specialised for the fetch model: the generated collision program
// (robots/fetch_check.inc) instantiated into the dense kernel template.
#include "bcc_dense.h"

namespace bcc {
#include "robots/fetch_check.inc"
BCC_DEFINE_ROBOT(fetch, Robot_fetch)
}  // namespace bcc
