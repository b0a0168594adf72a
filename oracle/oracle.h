/*
 * oracle.h — CPU restatement of the reference's CRT-NTT big-integer multiply.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under cryptography_amd/ links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg do, and only as the checker (or as the timed CPU baseline).
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The arithmetic parameters follow the build spec (three primes c*2^k+1 just
 * below 2^60, 16-bit coefficients) rather than the reference's four 2^64-2^b+1
 * primes and whole-limb coefficients; the algorithms are the same.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_prime / orc_root      primality and exact root order (tests/test_oracle.py)
 *   orc_mulmod / orc_shoup_*  random and corner comparison with Python ints
 *   orc_ntt_forward/inverse   O(n^2) DFT (Python ints), reference ntt_cpu.c spectra
 *                             (oracle/_ref, committed as tests/golden/ntt_spectra.json)
 *   orc_garner3               X mod p_i == r_i and a closed-form CRT (Python)
 *   orc_pack16 / orc_pack16_cla  Python big-int sum, long carry chains
 *   orc_mul_ntt3 / orc_mul_schoolbook  GMP mpz_mul and Python int multiply
 *   orc_binary_split          reference compute_e_host output, GMP, OEIS digits
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NPRIMES 3

/* Prime i (i < 3), its 2-adicity k (p = c*2^k + 1) and a generator of Z_p^*. */
uint64_t orc_prime(int i);
int      orc_prime_k(int i);
uint64_t orc_prime_gen(int i);
/* A primitive 2^log_n-th root of unity mod prime i: g^((p-1)/2^log_n). */
uint64_t orc_root(int i, int log_n);

uint64_t orc_addmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_submod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t orc_invmod(uint64_t a, uint64_t p);
/* Shoup: w' = floor(w*2^64/p); mul returns a*w mod p in [0,p) for any a < 2^64. */
uint64_t orc_shoup_pre(uint64_t w, uint64_t p);
uint64_t orc_shoup_mul_lazy(uint64_t a, uint64_t w, uint64_t wp, uint64_t p); /* [0,2p) */
uint64_t orc_shoup_mul(uint64_t a, uint64_t w, uint64_t wp, uint64_t p);

/* In-place natural-order NTT of length 2^log_n over prime index pi (CT-DIT,
 * bit-reverse then radix-2 stages).  Values must be in [0,p). */
void orc_ntt_forward(uint64_t *a, int log_n, int pi);
/* Inverse including the 1/n scale. */
void orc_ntt_inverse(uint64_t *a, int log_n, int pi);
/* Plain O(n^2) DFT with root w of order n (tiny sizes only; test helper). */
void orc_dft_naive(const uint64_t *in, uint64_t *out, int n, uint64_t w, uint64_t p);

/* Garner CRT for three residues: x = a0 + a1*P0 + a2*P0*P1, x < P0*P1*P2,
 * returned little-endian in out[0..2]. */
void orc_garner3(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t out[3]);

/* Split n_limbs u64 limbs into 16-bit coefficients; zero-pads to n_coeffs. */
void orc_split16(const uint64_t *limbs, size_t n_limbs, uint64_t *coeffs, size_t n_coeffs);

/* out[0..n_out) = sum_k X[k] * 2^(16k) where each X[k] has 3 little-endian
 * words (X + 3k).  Sequential carry.  Returns the carry word out of the top
 * limb (0 when n_out is large enough). */
uint64_t orc_pack16(const uint64_t *X, size_t n_coeffs, uint64_t *out, size_t n_out);
/* Same value via the carry-lookahead generate/propagate decomposition
 * (assumes X[k] < 2^64, i.e. word 1 and 2 zero); returns -1 if not. */
int      orc_pack16_cla(const uint64_t *X, size_t n_coeffs, uint64_t *out, size_t n_out);

/* Full 3-prime CRT-NTT multiply: out[0..na+nb) = a*b (u64 limbs, untrimmed).
 * Transform length is the least power of two >= 4(na+nb)-1.
 * n_threads > 1 runs the per-prime transforms with OpenMP.  Returns 0 or -1. */
int orc_mul_ntt3(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                 uint64_t *out, int n_threads);
/* The same product with every transform spread over n_threads threads (the
 * all-core CPU baseline; identical results). */
int orc_mul_ntt3_par(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                     uint64_t *out, int n_threads);
/* O(na*nb) schoolbook multiply with a u128 carry; out has na+nb limbs. */
void orc_mul_schoolbook(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                        uint64_t *out);

/* Transform length used for a product of na and nb u64 limbs (log2). */
int orc_log_len(size_t na, size_t nb);

/* Binary splitting for sum_{k=a}^{b-1} 1/k!: P and B as u64 limb arrays.
 * *P / *B are malloc'd; lengths returned in *np / *nb.  Uses the same split
 * point m = a + (b-a)/2 and merge formula as the reference. */
void orc_binary_split(long a, long b, uint64_t **P, size_t *np, uint64_t **B, size_t *nb);
long orc_split_terms_needed(long n_digits);

/* splitmix64 stream: fills n words from seed (the fixture PRNG). */
void orc_splitmix_fill(uint64_t seed, uint64_t *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
