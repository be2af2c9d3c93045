// Size dispatch shared by the instantiation units: FN<T, n, NL>(args, stream) for the grid sizes
// this build compiles (powers of two, 16..4096). A column length that is not a power of two runs
// in the y-pass kernels of the next power of two (the kernels read the true length from Geo).
#pragma once
#define GF_SIZES(FN, T, NL, n, a, s)                     \
  switch (pow2_ceil(n)) {                                \
    case 16: FN<T, 16, NL>(a, s); break;                 \
    case 32: FN<T, 32, NL>(a, s); break;                 \
    case 64: FN<T, 64, NL>(a, s); break;                 \
    case 128: FN<T, 128, NL>(a, s); break;               \
    case 256: FN<T, 256, NL>(a, s); break;               \
    case 512: FN<T, 512, NL>(a, s); break;               \
    case 1024: FN<T, 1024, NL>(a, s); break;             \
    case 2048: FN<T, 2048, NL>(a, s); break;             \
    case 4096: FN<T, 4096, NL>(a, s); break;             \
    default: break;                                      \
  }
