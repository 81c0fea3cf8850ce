/* Kernel instantiations for fpa_neg (all its compiled shapes). */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(neg, 0, 0, 1)
