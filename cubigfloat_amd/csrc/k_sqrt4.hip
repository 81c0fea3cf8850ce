/* Kernel instantiations for fpa_sqrt, compilation part 4 of 6. */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(sqrt, 5, 4, 6)
