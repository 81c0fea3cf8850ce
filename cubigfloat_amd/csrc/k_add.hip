/* Kernel instantiations for fpa_add (all its compiled shapes). */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(add, 1, 0, 1)
