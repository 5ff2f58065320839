// k_det.hip -- deterministic mode: Q32.32 int64 atomics, compiled with
// -ffp-contract=off so every float expression keeps the reference order.
#define AEMH_NS det
#define AEMH_DET 1
#include "kernels.inc.h"
