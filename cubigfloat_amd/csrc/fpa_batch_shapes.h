/*
 * Kernel shape per (op, precision class) compiled into the batched persistent kernel and
 * the dot-chain kernel: X(op, n, K, L) with op 1 add, 2 sub, 3 mul, 4 div, 5 sqrt and
 * n = precision/32 limbs.  Written by cubigfloat_amd/autotune.py from measured
 * per-shape times (edit there, not here).
 */
#ifndef FPA_BATCH_SHAPES_H
#define FPA_BATCH_SHAPES_H

#define FPA_BATCH_SHAPES(X) \
  X(1, 8, 8, 1) X(1, 16, 8, 2) X(1, 32, 8, 4) X(1, 64, 8, 8) X(1, 128, 16, 8) X(1, 256, 16, 16) X(1, 512, 16, 32) \
  X(2, 8, 8, 1) X(2, 16, 8, 2) X(2, 32, 8, 4) X(2, 64, 8, 8) X(2, 128, 16, 8) X(2, 256, 16, 16) X(2, 512, 16, 32) \
  X(3, 8, 8, 1) X(3, 16, 8, 2) X(3, 32, 8, 4) X(3, 64, 8, 8) X(3, 128, 16, 8) X(3, 256, 16, 16) X(3, 512, 16, 32) \
  X(4, 8, 8, 1) X(4, 16, 8, 2) X(4, 32, 8, 4) X(4, 64, 8, 8) X(4, 128, 16, 8) X(4, 256, 16, 16) X(4, 512, 16, 32) \
  X(5, 8, 8, 1) X(5, 16, 8, 2) X(5, 32, 8, 4) X(5, 64, 8, 8) X(5, 128, 16, 8) X(5, 256, 16, 16) X(5, 512, 16, 32)

#endif
