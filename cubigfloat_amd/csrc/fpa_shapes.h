/*
 * Compiled kernel shapes: X(n, K, L) = precision n*32 bits, K limbs per lane,
 * L lanes per value (K*L = n, L divides 64, L <= 32 so a mul's low and high carry
 * ballots share one 64-bit mask).  The first shape listed for an n is the dispatcher's
 * default; the autotuner may pick any listed shape per (op, precision).
 */
#ifndef FPA_SHAPES_H
#define FPA_SHAPES_H

#define FPA_SHAPES(X)                                        \
  X(8, 8, 1)     X(8, 4, 2)     X(8, 2, 4)                   \
  X(16, 8, 2)    X(16, 16, 1)   X(16, 4, 4)                  \
  X(32, 8, 4)    X(32, 16, 2)   X(32, 4, 8)                  \
  X(64, 8, 8)    X(64, 16, 4)   X(64, 4, 16)  X(64, 32, 2)   \
  X(96, 12, 8)   X(96, 6, 16)   X(96, 24, 4)                 \
  X(128, 16, 8)  X(128, 8, 16)  X(128, 32, 4) X(128, 4, 32)  \
  X(160, 10, 16) X(160, 20, 8)  X(160, 5, 32)                \
  X(192, 12, 16) X(192, 24, 8)  X(192, 6, 32)                \
  X(224, 14, 16) X(224, 28, 8)  X(224, 7, 32)                \
  X(256, 16, 16) X(256, 32, 8)  X(256, 8, 32)                \
  X(512, 16, 32) X(512, 32, 16)

/* mul is compiled only for shapes with K <= FPA_MUL_MAX_K: with 28 or 32 limbs per lane its
 * digit window no longer fits the registers, it spills, and it ran 20-40x slower than the
 * other shapes of the same precision. */
#define FPA_MUL_MAX_K 24

/* The dot-chain kernel (one mul core and one add core per element) is compiled only for
 * shapes with K <= FPA_DOT_MAX_K: above that its add core no longer inlines and becomes a
 * real call, which no kernel of this library may contain. */
#define FPA_DOT_MAX_K 20

#endif
