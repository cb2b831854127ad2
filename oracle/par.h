/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline; never
 * linked by the product path).
 *
 * Minimal fork-join helper standing in for the reference's rayon
 * `par_chunks_mut` / `into_par_iter` loops (e.g. crypto/math/src/fft/
 * two_half_fft.rs:213-232, crypto/stark/src/prover.rs:942-949): a range
 * [0, count) is split into contiguous chunks, one pthread each.
 */
#ifndef LVM_ORACLE_PAR_H
#define LVM_ORACLE_PAR_H
#include <stddef.h>

typedef void (*par_body_fn)(void *ctx, size_t begin, size_t end, int thread);

int par_threads(void);
void par_set_threads(int t);
/* Runs body over [0, count) split into at most par_threads() chunks whose
 * boundaries are multiples of `grain` (except the last). */
void par_for(size_t count, size_t grain, par_body_fn body, void *ctx);

#endif
