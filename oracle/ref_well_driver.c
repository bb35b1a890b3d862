/*
 * Driver for the reference's own RNG path, compiled from the reference
 * sources in place (oracle/Makefile target "ref"): src/WELL19937a_new.c with
 * gcc and src/pic2d_RandomNumberInterface.f90 with amdflang.  TEST
 * INFRASTRUCTURE ONLY: it exists to pin oracle.c's WELL restatement.
 */
#include <stdint.h>

extern void _QMrng_wrapperPwell_random_seed(int32_t *seed);
extern double _QMrng_wrapperPwell_random_number(void);

void ref_well_seed(int32_t seed) { _QMrng_wrapperPwell_random_seed(&seed); }

void ref_well_fill(int64_t n, double *out) {
    for (int64_t k = 0; k < n; k++) out[k] = _QMrng_wrapperPwell_random_number();
}
