/*
 * CPU oracle for the finite-field NTT -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's host FFT
 * (test/fft_host.h:7-118) over correct prime-field arithmetic.  It exists to
 * check the GPU library and to time the CPU baseline.  Only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it; the
 * library itself never links or calls it.
 *
 * Elements are arrays of little-endian u32 limbs (1 limb for the 32-bit
 * field, 24 for the 753-bit fields) holding residues in [0, p) in Montgomery
 * form with radix 2^(32*limbs), the layout of fields::Scalar
 * (cuda/device_field.h:75).
 *
 * Deviations from the reference, each deliberate (see the header of oracle.c):
 *   - the Montgomery constant, final reduction, one() and pow() are correct
 *     (the reference's are not: device_field_operators.h:47,139-149,216-258;
 *     device_field.h:86-91);
 *   - the radix-2 butterfly multiplies the odd input by the twiddle
 *     (test/fft_host.h:44-48 adds and subtracts the twiddle itself).  The
 *     original butterfly is kept behind `literal_butterfly` so the loop
 *     structure can be compared bit for bit with the reference source.
 */
#ifndef GSN_ORACLE_H
#define GSN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field identifiers; the same numbering as the library's C ABI. */
enum { OR_FIELD_FP32 = 0, OR_FIELD_MNT4753_FR = 1, OR_FIELD_MNT6753_FR = 2 };

#define OR_MAX_LIMBS 24

typedef struct {
    int id;
    int nlimbs;                     /* 1 or 24 */
    uint32_t mod[OR_MAX_LIMBS];     /* p */
    uint32_t mprime;                /* -p^-1 mod 2^32 */
    uint32_t one[OR_MAX_LIMBS];     /* R mod p, R = 2^(32*nlimbs) */
    uint32_t r2[OR_MAX_LIMBS];      /* R^2 mod p */
    unsigned two_adicity;
} or_field;

/* Fills *f for the given field id; returns 0 on success. */
int or_field_init(or_field *f, int field_id);
const or_field *or_field_get(int field_id);

/* Field operations on canonical residues.  r may alias a or b. */
void or_add(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void or_sub(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void or_mul(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b);
void or_pow(const or_field *f, uint32_t *r, const uint32_t *a, uint64_t e);
void or_to_mont(const or_field *f, uint32_t *r, const uint32_t *a);
void or_from_mont(const or_field *f, uint32_t *r, const uint32_t *a);
int or_is_canonical(const or_field *f, const uint32_t *a);

/* test/fft_host.h:7-16 */
size_t or_bitreverse(size_t n, unsigned l);

/* test/fft_host.h:18-54.  In place on n = 2^k elements of f->nlimbs limbs. */
void or_serial_radix2_fft(const or_field *f, uint32_t *a, size_t n,
                          const uint32_t *omega, const uint32_t *one,
                          int literal_butterfly);

/* test/fft_host.h:56-118.  `threads` <= 0 means the OpenMP default. */
void or_parallel_radix2_fft_inner(const or_field *f, uint32_t *a, size_t n,
                                  const uint32_t *omega, unsigned log_cpus,
                                  const uint32_t *one, int literal_butterfly,
                                  int threads);

/* Phase 1 of the parallel variant (test/fft_host.h:76-102) restricted to
 * the partial sums j in [j_begin, j_end): used to time a bounded sample. */
void or_parallel_phase1_sample(const or_field *f, const uint32_t *a, size_t n,
                               const uint32_t *omega, unsigned log_cpus,
                               size_t j_begin, size_t j_end, uint32_t *tmp_out,
                               int threads);

/* Inverse transform: forward transform with omega^-1, then scale by n^-1. */
void or_inverse_fft(const or_field *f, uint32_t *a, size_t n, const uint32_t *omega);
/* Extended radix-2 domain H u shift*H (|H| = n/2, omega_half generates H):
 * forward evaluates on H then shift*H; inverse interpolates back. */
/* G1: projective points (X, Y, Z) of 3 * nlimbs limbs, Montgomery form, Z = 0
 * the identity; coeff_a in Montgomery form; scalars plain integers. */
void or_g1_dbl(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p);
void or_g1_add(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p, const uint32_t *q);
void or_g1_mul(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p, const uint32_t *k,
               int klimbs);
void or_g1_msm(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *points,
               const uint32_t *k, int klimbs, size_t n);
/* Mixed-radix transforms: n = n1 * n2 (n1 odd, n2 = 2^a), omega = w2 * w1. */
void or_dft_direct(const or_field *f, uint32_t *out, const uint32_t *x, size_t n, const uint32_t *omega);
/* sum_j x[j] omega^(j k) at the nk indices ks (O(n) each, OpenMP over chunks
 * of j); `threads` <= 0 means the OpenMP default. */
void or_dft_points(const or_field *f, uint32_t *out, const uint32_t *x, size_t n, const uint32_t *omega,
                   const uint64_t *ks, size_t nk, int threads);
void or_mixed_fft(const or_field *f, uint32_t *a, size_t n1, size_t n2, const uint32_t *w2, const uint32_t *w1,
                  int inverse);
void or_extended_radix2_fft(const or_field *f, uint32_t *a, size_t n, const uint32_t *omega_half,
                            const uint32_t *shift, int inverse);

/* omega^-1 (Fermat inversion). */
void or_inverse(const or_field *f, uint32_t *r, const uint32_t *a);

#ifdef __cplusplus
}
#endif

#endif
