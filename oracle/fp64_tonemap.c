/* TEST INFRASTRUCTURE ONLY: the FP64 evaluation of tonemap.c (see fp64.h). */
#include "fp64.h"
#include "tonemap.c"
