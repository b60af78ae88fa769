/*
 * kf_oracle_f32.c -- the oracle compiled with a float scalar type (TEST INFRASTRUCTURE ONLY).
 *
 * kf_oracle.c with every `double` read as `float`, floating constants single precision
 * (-fsingle-precision-constant, oracle/Makefile) and sqrt/fabs in float: the same FP64 code,
 * evaluated in FP32 throughout, structs and arrays included.  It exists to measure what FP32
 * rounding alone does to the reference optimiser over several iterations
 * (tests/test_oracle_fp32_spread.py), the spread the GPU's FP32 reference-formula path is
 * held to against the FP64 oracle; it checks nothing by itself.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define double float
#define sqrt sqrtf
#define fabs fabsf
#include "kf_oracle.c"
