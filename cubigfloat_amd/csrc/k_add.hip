/* Kernel instantiations for fpa_add (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_add(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 1>(*a, st);
FPA_DEFINE_LAUNCHER(add, 1)
