/* Oracle (test infrastructure only): FP64 instantiation of pm_impl.h */
#define REAL double
#define SFX(name) name##_d
#include "pm_impl.h"
