#pragma once
/*
 * compute_e.h — Euler's number to D significant digits on the GPU.
 *
 * Same value and output format as the reference's compute_e
 * (ntt/app/compute_e/main.c:167-345): 20 guard digits, N =
 * split_terms_needed(D + 20) terms of sum 1/k!, x = (B + P) / B, the first D
 * significant digits of x truncated, written as "2.ddd..." (D digits).
 * No counterpart as one call in the reference (its main() strings together
 * binary_split, newton_reciprocal, bigint_div_newton and bigint_to_decimal).
 */
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double t_tree;        /* product tree (binary splitting), seconds */
    double t_recip;       /* Newton reciprocal of B */
    double t_div;         /* fixed-point quotient (P - B) / B */
    double t_radix;       /* radix conversion to decimal */
    double t_total;
    long n_terms;
    unsigned long products;   /* device multiplies issued */
} ntt_e_stats;

/* out receives "2." followed by D-1 digits and a NUL (out_cap >= D + 2).
 * batched_tree: 1 evaluates the product tree level-batched on the device
 * (one hipGraph), 0 uses the reference's depth-first binary_split with one
 * bigint_mul call per product (the plain path). */
int ntt_compute_e(long n_digits, int batched_tree, char *out, size_t out_cap, ntt_e_stats *stats);

#ifdef __cplusplus
}
#endif
