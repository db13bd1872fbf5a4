/*
 * band.h — band Cholesky of the pinned pressure operator, shared by the two
 * oracle restatements (fcfd_oracle.c, m3_oracle.c).  TEST INFRASTRUCTURE
 * ONLY, never part of the product.
 *
 * Storage: row r of the symmetric positive definite matrix keeps its columns
 * r-b..r at L[r (b+1) + (c - r + b)].  The caller fills the lower band with
 * the first unknown pinned (row and column 0 the identity), factors, and
 * solves; orc_band_solve keeps x[0] = 0.
 */
#ifndef FCFD_ORACLE_BAND_H
#define FCFD_ORACLE_BAND_H

typedef struct {
    double *L;
    long n, b;
} orc_band;

int orc_band_alloc(orc_band *B, long n, long b);   /* (re)allocate and zero */
int orc_band_cholesky(orc_band *B);                /* in place */
void orc_band_solve(const orc_band *B, double *x);

#endif
