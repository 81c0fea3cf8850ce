/* Kernel instantiations for fpa_sub (all its compiled shapes). */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(sub, 2, 0, 1)
