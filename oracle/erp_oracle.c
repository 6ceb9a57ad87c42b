/*
 * erp_oracle.c -- CPU oracle for ERPlacer's global-placement iteration.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
 * CPU-baseline leg; never by the product library.
 *
 * The type-generic body (erp_oracle_impl.inc) is instantiated for double
 * (the FP64 parity oracle) and float (the reference binary's precision, used
 * as the OpenMP CPU baseline).  Build: oracle/Makefile.
 */
#define _GNU_SOURCE
#include "erp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_SQRT2
#define M_SQRT2 1.41421356237309504880
#endif

/* REAL: the element type of every array; SCAL: the scalars of the loop (step
 * size, momentum, density weight, gamma, scheduler state) and the sums that
 * produce them (HPWL, overflow, energy, the step-size and L1 dot products). */
#define REAL double
#define SCAL double
#define SFX _f64
#include "erp_oracle_impl.inc"
#undef REAL
#undef SCAL
#undef SFX

#define REAL float
#define SCAL float
#define SFX _f32
#include "erp_oracle_impl.inc"
#undef REAL
#undef SCAL
#undef SFX

/* f32 arrays with FP64 scalars and sums: the precision split of the GPU path
 * (f32 coordinates and per-element arithmetic as the reference binary, FP64
 * step-size dot products, scheduler state, HPWL, overflow and energy sums). */
#define REAL float
#define SCAL double
#define SFX _f32d
#include "erp_oracle_impl.inc"
#undef REAL
#undef SCAL
#undef SFX

/* Fixed-point scale for the density accumulators: the largest power of two S
 * with mass_bound * S <= 2^62, so no bin (nor the overflow sum) can exceed the
 * int64 range.  The same rule is used by the GPU path. */
double ero_fx_scale(double mass_bound) {
    if (!(mass_bound > 0)) return 1.0;
    int e;
    frexp(4611686018427387904.0 / mass_bound, &e);   /* 2^62 / bound = f * 2^e, f in [0.5,1) */
    return ldexp(1.0, e - 1);
}

/* GPU finalize semantics restated: rho = (q_mov + q_fix) / S rounded once to
 * float; overflow numerator = sum max(q_mov + q_fix - round(coef * S), 0). */
void ero_fx_finalize_f32(int64_t nbins, const int64_t *q_mov, const int64_t *q_fix,
                         double scale, double coef, float *rho, int64_t *ovf_q) {
    const int64_t coef_q = (int64_t)llrint(coef * scale);
    const double inv = 1.0 / scale;
    int64_t acc = 0;
    for (int64_t b = 0; b < nbins; b++) {
        int64_t t = q_mov[b] + q_fix[b];
        rho[b] = (float)((double)t * inv);
        if (t > coef_q) acc += t - coef_q;
    }
    *ovf_q = acc;
}
