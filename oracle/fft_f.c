/* Oracle (test infrastructure only): FP32 instantiation of fft_impl.h */
#define REAL float
#define SFX(name) name##_f
#include "fft_impl.h"
