// k_fast.hip -- FP32 mode: float atomics; compiled with -ffp-contract=off so
// results differ from the reference only by the order of the P2G sums.
#define AEMH_NS fast
#define AEMH_DET 0
#include "kernels.inc.h"
