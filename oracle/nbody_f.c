/* Oracle (test infrastructure only): FP32 instantiation of nbody_impl.h */
#define REAL float
#define SFX(name) name##_f
#include "nbody_impl.h"
