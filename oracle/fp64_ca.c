/* TEST INFRASTRUCTURE ONLY: the FP64 evaluation of ca.c (see fp64.h). */
#include "fp64.h"
#include "ca.c"
