/*
 * oracle.c -- CPU oracle for the bendy2d solver step.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; the product library never
 * links, loads or calls it.
 *
 * Two instantiations of oracle_impl.h:
 *   or32_*  REAL = float   bit-exact restatement of Murrent/bendy2d 0.0.1
 *   or64_*  REAL = double  FP64 twin, used to state each result's error
 *
 * Build: make -C oracle   (gcc -O2 -ffp-contract=off, no -march, no fast-math)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SQRT sqrtf
#define PFX or32_
#include "oracle_impl.h"
#undef REAL
#undef SQRT
#undef PFX

#define REAL double
#define SQRT sqrt
#define PFX or64_
#include "oracle_impl.h"
#undef REAL
#undef SQRT
#undef PFX
