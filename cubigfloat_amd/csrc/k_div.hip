/* Kernel instantiations for fpa_div (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_div(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 4>(*a, st);
FPA_DEFINE_LAUNCHER(div, 4)
