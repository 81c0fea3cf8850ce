/*
 * Compiled kernel shapes: X(n, K, L) = precision n*32 bits, K limbs per lane,
 * L lanes per value (K*L = n, L divides 64).  The first shape listed for an n is the
 * dispatcher's default; the autotuner may pick any listed shape per (op, precision).
 */
#ifndef FPA_SHAPES_H
#define FPA_SHAPES_H

#define FPA_SHAPES(X) \
  X(8, 8, 1)    X(8, 4, 2)                  \
  X(16, 8, 2)   X(16, 16, 1)                \
  X(32, 8, 4)   X(32, 16, 2)                \
  X(64, 8, 8)   X(64, 16, 4)                \
  X(96, 12, 8)                              \
  X(128, 16, 8) X(128, 8, 16) X(128, 32, 4) \
  X(160, 10, 16)                            \
  X(192, 12, 16)                            \
  X(224, 14, 16)                            \
  X(256, 16, 16) X(256, 32, 8)              \
  X(512, 16, 32)

#endif
