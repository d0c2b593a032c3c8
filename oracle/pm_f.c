/* Oracle (test infrastructure only): FP32 instantiation of pm_impl.h */
#define REAL float
#define SFX(name) name##_f
#include "pm_impl.h"
