/* Kernel instantiations for fpa_sqrt, compilation part 5 of 6. */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(sqrt, 5, 5, 6)
