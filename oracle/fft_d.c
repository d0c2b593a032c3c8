/* Oracle (test infrastructure only): FP64 instantiation of fft_impl.h */
#define REAL double
#define SFX(name) name##_d
#include "fft_impl.h"
