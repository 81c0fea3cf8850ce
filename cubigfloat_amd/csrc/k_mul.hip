/* Kernel instantiations for fpa_mul (all its compiled shapes). */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(mul, 3, 0, 1)
