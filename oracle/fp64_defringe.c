/* TEST INFRASTRUCTURE ONLY: the FP64 evaluation of defringe.c (see fp64.h). */
#include "fp64.h"
#include "defringe.c"
