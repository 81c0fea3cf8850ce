/* Kernel instantiations for fpa_div, compilation part 1 of 2. */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(div, 4, 1, 2)
