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
  X(64, 8, 8)    X(64, 16, 4)   X(64, 4, 16)  X(64, 32, 2)  X(64, 2, 32)  \
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

/* Kinds of the persistent batch kernel: X(op, n, K, L), op 1 add, 2 sub, 3 mul, 4 div,
 * 5 sqrt, one shape per (op, precision) for 256 ... 8192 bits.  Every kind is inlined into
 * one kernel whose registers are the largest kind's, so the shapes are the tuned ones held
 * to at most 128 VGPRs (4 waves per SIMD): div takes 8 limbs per lane, where its 16-limb
 * shapes need up to 164. */
#define FPA_PERSIST_SHAPES(X)                                                                \
  X(1, 8, 8, 1)   X(2, 8, 8, 1)   X(3, 8, 8, 1)   X(4, 8, 8, 1)   X(5, 8, 8, 1)              \
  X(1, 16, 8, 2)  X(2, 16, 8, 2)  X(3, 16, 8, 2)  X(4, 16, 8, 2)  X(5, 16, 8, 2)             \
  X(1, 32, 8, 4)  X(2, 32, 8, 4)  X(3, 32, 8, 4)  X(4, 32, 8, 4)  X(5, 32, 8, 4)             \
  X(1, 64, 8, 8)  X(2, 64, 8, 8)  X(3, 64, 8, 8)  X(4, 64, 8, 8)  X(5, 64, 16, 4)            \
  X(1, 128, 8, 16) X(2, 128, 8, 16) X(3, 128, 16, 8) X(4, 128, 8, 16) X(5, 128, 16, 8)       \
  X(1, 256, 8, 32) X(2, 256, 8, 32) X(3, 256, 16, 16) X(4, 256, 8, 32) X(5, 256, 16, 16)

#endif
