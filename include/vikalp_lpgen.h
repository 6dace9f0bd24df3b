/*
 * Seeded synthetic LP generators (C ABI).
 *
 * No counterpart in the reference: it ships no seeded LP (its tests use
 * hand-written 1- and 2-variable models, tests/solver/continuous/test_pdhg.cpp).
 * These generators make the benchmark workloads and parity inputs, and are
 * shared by the HIP runs, the benchmark and the CPU checker so every machine
 * rebuilds the same instance from (family, sizes, seed).
 *
 * The LP is   min c'x  s.t. row_lower <= A x <= row_upper,
 *             var_lower <= x <= var_upper,
 * with A in CSR using int64 offsets and indices (the reference's
 * vikalp::Index, include/vikalp/contracts/Model.hpp:11).
 */
#ifndef VIKALP_LPGEN_H
#define VIKALP_LPGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vk_lp vk_lp;

/* family 1: m x n, exactly k nonzeros per column, all equality rows.
 * family 2: m x n, Zipf(1.5) column lengths on [2,200], 10% columns rescaled,
 *           70% equality rows, 30% <= rows (k ignored).
 * family 4: block-angular LP with k periods (m, n ignored).
 * Returns 0 on success, -1 on bad arguments, -2 on allocation failure. */
int vk_lp_generate(int family, int64_t m, int64_t n, int64_t k, uint64_t seed,
                   vk_lp **out);
int vk_lp_dims(const vk_lp *lp, int64_t *m, int64_t *n, int64_t *nnz);
/* Copies into caller buffers; any pointer may be NULL to skip that array.
 * Sizes: row_offsets m+1; column_indices, values nnz; objective, var_lower,
 * var_upper, x0 n; row_lower, row_upper, y0 m. */
int vk_lp_export(const vk_lp *lp, int64_t *row_offsets, int64_t *column_indices,
                 double *values, double *objective, double *row_lower,
                 double *row_upper, double *var_lower, double *var_upper,
                 double *x0, double *y0);
void vk_lp_free(vk_lp *lp);
/* Number of periods T ~ U{4..48} of instance `instance` in a batch. */
int64_t vk_lp_block_periods(uint64_t seed, int64_t instance);

#ifdef __cplusplus
}
#endif
#endif
