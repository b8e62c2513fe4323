/*
 * TEST INFRASTRUCTURE ONLY -- a thin harness around the reference's own
 * header-only quicksort, include/sif/utils/sort.h of sif v0.5.0, compiled in
 * place from /root/reference by oracle/Makefile into oracle/_ref/. It is the
 * one piece of the reference path that builds here without FFTW (every other
 * hot-path unit includes math/fft.h -> fftw3.h), and it pins the tie order of
 * orc_sort_candidates_ref().
 *
 * The candidate type and the predicate are those of src/finder/utils.h:122-131
 * and src/finder/utils.c:24.
 */
#include <stdint.h>

#include "sif/utils/sort.h"

typedef struct {
  sif_real delta;
  uint32_t refined;
  uint64_t flat_idx;
} ref_candidate_t;

SIF_DEFINE_QUICKSORT_PARALLEL(ref_sort_candidates_impl, ref_candidate_t,
  a.delta < b.delta)

void ref_sort_candidates(ref_candidate_t* items, uint64_t count) {
  ref_sort_candidates_impl(items, count);
}
