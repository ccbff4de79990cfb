/*
 * ORACLE — test infrastructure only (the parity reference of the tests and the CPU baseline).
 * CPU restatement of the reference's FDTD time loop; see
 * fdtd_oracle_impl.h for the per-function reference citations.
 *
 * Build: see oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#define T float
#define SUF f32
#include "fdtd_oracle_impl.h"
#undef T
#undef SUF

#define T double
#define SUF f64
#include "fdtd_oracle_impl.h"
#undef T
#undef SUF

int oracle_abi_version(void) { return 1; }
