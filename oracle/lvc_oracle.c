/*
 * CPU ORACLE for the LocVolCalib hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of finpar/LocVolCalib.fut (the reference is Futhark
 * source; its compiler is not available, so the reference itself cannot be
 * built or run).  It exists to check the HIP implementation: only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg load it.  The
 * product library never links or calls it.
 *
 * Instantiated for float (suffix _f32, the reference's own precision) and
 * double (suffix _f64, the FP64 oracle: a type substitution f32 -> f64 of the
 * same program).  Two tridiagonal modes: the reference's scan form
 * (LocVolCalib.fut:68-124) and Thomas (misc/ocean-sim/tridiag.fut:81-95).
 *
 * Pinning (which checks fix each function) is listed in the header of
 * oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* Planted-mistake switch: non-zero values make one function of the oracle
 * wrong on purpose, so the tests can show that their checks catch it. */
static int lvc_oracle_mutation = 0;
void lvc_oracle_set_mutation(int m) { lvc_oracle_mutation = m; }
int lvc_oracle_get_mutation(void) { return lvc_oracle_mutation; }
#define MUT(k) (lvc_oracle_mutation == (k))

/* GRID32 switch: initGrid's scalars in f32 (see init_grid_ in the body). */
static int lvc_oracle_grid32 = 0;
void lvc_oracle_set_grid32(int on) { lvc_oracle_grid32 = on != 0; }
#define GRID32 (lvc_oracle_grid32 != 0)

#define R float
#define SUF f32
#define LOG logf
#define EXP expf
#define SQRT sqrtf
#include "lvc_oracle_impl.h"
#undef R
#undef SUF
#undef LOG
#undef EXP
#undef SQRT

#define R double
#define SUF f64
#define LOG log
#define EXP exp
#define SQRT sqrt
#include "lvc_oracle_impl.h"
#undef R
#undef SUF
#undef LOG
#undef EXP
#undef SQRT
