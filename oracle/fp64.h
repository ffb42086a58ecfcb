/*
 * TEST INFRASTRUCTURE ONLY.  An FP64 evaluation of the float stages: each
 * fp64_<stage>.c includes this header and then the oracle's restatement of the
 * stage, so the same source is compiled a second time with every float, float
 * cast and single-precision libm call turned into its double counterpart and
 * the entry points renamed oracle64_*.  The formulas and their order are the
 * oracle's; only the precision changes.  Constants written as float literals
 * (0.15f and the like) keep their float values, which are the reference's
 * constants.  The tests and bench.py report the error of the HIP output and of
 * the reference build against these evaluations (arrays of double).
 */
#ifndef LRE_ORACLE_FP64_H
#define LRE_ORACLE_FP64_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define float double
#define expf exp
#define logf log
#define sqrtf sqrt
#define ceilf ceil
#define floorf floor
#define fabsf fabs

#define oracle_defringe oracle64_defringe
#define oracle_axial_ca oracle64_axial_ca
#define oracle_lateral_ca oracle64_lateral_ca
#define oracle_detail_tonemap oracle64_detail_tonemap
#define oracle_tone_mapping oracle64_tone_mapping
#define oracle_white_balance oracle64_white_balance
#endif
