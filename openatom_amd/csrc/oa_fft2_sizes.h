// Transform lengths compiled for the two-factor FFT engine: X(N, R1, R2),
// N = R1 * R2 with R1, R2 <= 16.  Any other length of an axis makes the
// context use the generic Stockham kernels of oa_fft.hip for all passes.
// The Makefile reads this list (one X(...) per line).
#pragma once
#define OA_FFT2_SIZES(X) \
  X(24, 4, 6)            \
  X(27, 3, 9)            \
  X(30, 5, 6)            \
  X(32, 4, 8)            \
  X(36, 6, 6)            \
  X(40, 5, 8)            \
  X(45, 5, 9)            \
  X(48, 6, 8)            \
  X(50, 5, 10)           \
  X(54, 6, 9)            \
  X(60, 6, 10)           \
  X(64, 8, 8)            \
  X(72, 8, 9)            \
  X(75, 5, 15)           \
  X(80, 8, 10)           \
  X(81, 9, 9)            \
  X(90, 9, 10)           \
  X(96, 8, 12)           \
  X(100, 10, 10)         \
  X(108, 9, 12)          \
  X(120, 10, 12)         \
  X(128, 8, 16)          \
  X(144, 12, 12)         \
  X(150, 10, 15)         \
  X(160, 10, 16)
