// k_fast.hip -- FP32 mode: float atomics, compiled with -ffp-contract=fast.
#define AEMH_NS fast
#define AEMH_DET 0
#include "kernels.inc.h"
