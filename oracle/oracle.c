/*
 * CPU oracle for the finite-field NTT -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's host FFT (test/fft_host.h) and of
 * the field arithmetic it runs on (cuda/device_field_operators.h), written
 * to be checked against golden vectors, the DFT definition and the
 * reference source compiled from /root/reference.  The GPU library never
 * links or calls this file.
 *
 * Deliberate deviations from the reference (all are reference defects):
 *   1. Montgomery constant: -p^-1 mod 2^32 is computed (0xE45E7FFF for
 *      MNT4-753 Fr); the reference hard-codes 0xFFFFFFFF
 *      (device_field_operators.h:47).
 *   2. Final reduction: CIOS ends with a full conditional subtraction of p;
 *      the reference's montyNormalize copies 24 bytes instead of 24 limbs
 *      (device_field_operators.h:142,147).
 *   3. one() is R mod p and pow() is square-and-multiply inside the
 *      Montgomery domain; the reference returns raw 1 and leaves the domain
 *      (device_field.h:86-91, device_field_operators.h:222-258).
 *   4. The butterfly computes t = w * a[k+j+m]; the reference sets t = w and
 *      multiplies w by the odd input (test/fft_host.h:44-48).  Passing
 *      literal_butterfly = 1 restores the reference's statements exactly.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
static int nthreads_for(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }
#else
static int nthreads_for(int threads) { (void)threads; return 1; }
#endif

/* MNT4-753 Fr limbs as stored in cuda/device_field.h:62-65. */
static const uint32_t MNT4_LIMBS[24] = {
    610172929u, 1586521054u, 752685471u, 3818738770u, 2596546032u, 1669861489u,
    1987204260u, 1750781161u, 3411246648u, 3087994277u, 4061660573u, 2971133814u,
    2707093405u, 2580620505u, 3902860685u, 134068517u, 1821890675u, 1589111033u,
    1536143341u, 3086587728u, 4007841197u, 270700578u, 764593169u, 115910u};

/* MNT6-753 Fr, the decimal printed at cuda/device_field.h:446. */
static const char *MNT6_DECIMAL =
    "41898490967918953402344214791240637128170709919953949071783502921025352812571106773"
    "05889376379033892141807097188845847732317305749159385506969624185479639616572141632"
    "5350064441470418137846398469611935719059908164220784476160001";

/* ---------------------------------------------------------------- limbs */

static int geq_n(const uint32_t *a, const uint32_t *b, int s)
{
    for (int i = s - 1; i >= 0; --i) {
        if (a[i] != b[i])
            return a[i] > b[i];
    }
    return 1;
}

static uint32_t add_n(uint32_t *r, const uint32_t *a, const uint32_t *b, int s)
{
    uint64_t c = 0;
    for (int i = 0; i < s; ++i) {
        c += (uint64_t)a[i] + b[i];
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    return (uint32_t)c;
}

static uint32_t sub_n(uint32_t *r, const uint32_t *a, const uint32_t *b, int s)
{
    uint64_t borrow = 0;
    for (int i = 0; i < s; ++i) {
        uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        r[i] = (uint32_t)d;
        borrow = (d >> 63) & 1;
    }
    return (uint32_t)borrow;
}

/* x = (2x) mod p for x < p. */
static void dbl_mod(const or_field *f, uint32_t *x)
{
    uint32_t c = add_n(x, x, x, f->nlimbs);
    if (c || geq_n(x, f->mod, f->nlimbs))
        sub_n(x, x, f->mod, f->nlimbs);
}

static void parse_decimal(const char *s, uint32_t *out, int nlimbs)
{
    memset(out, 0, sizeof(uint32_t) * nlimbs);
    for (; *s; ++s) {
        uint64_t carry = (uint64_t)(*s - '0');
        for (int i = 0; i < nlimbs; ++i) {
            uint64_t v = (uint64_t)out[i] * 10u + carry;
            out[i] = (uint32_t)v;
            carry = v >> 32;
        }
    }
}

/* ---------------------------------------------------------------- fields */

int or_field_init(or_field *f, int field_id)
{
    memset(f, 0, sizeof(*f));
    f->id = field_id;
    switch (field_id) {
    case OR_FIELD_FP32:
        f->nlimbs = 1;
        f->mod[0] = 3u * (1u << 30) + 1u;
        break;
    case OR_FIELD_MNT4753_FR:
        f->nlimbs = 24;
        memcpy(f->mod, MNT4_LIMBS, sizeof(MNT4_LIMBS));
        break;
    case OR_FIELD_MNT6753_FR:
        f->nlimbs = 24;
        parse_decimal(MNT6_DECIMAL, f->mod, 24);
        break;
    default:
        return -1;
    }
    /* -p^-1 mod 2^32 by Newton iteration on the low limb. */
    uint32_t inv = 1;
    for (int i = 0; i < 6; ++i)
        inv *= 2u - f->mod[0] * inv;
    f->mprime = 0u - inv;
    /* R mod p and R^2 mod p by repeated doubling from 1. */
    uint32_t x[OR_MAX_LIMBS] = {0};
    x[0] = 1;
    for (int i = 0; i < 32 * f->nlimbs; ++i)
        dbl_mod(f, x);
    memcpy(f->one, x, sizeof(uint32_t) * f->nlimbs);
    for (int i = 0; i < 32 * f->nlimbs; ++i)
        dbl_mod(f, x);
    memcpy(f->r2, x, sizeof(uint32_t) * f->nlimbs);
    /* 2-adicity of p - 1 */
    uint32_t pm1[OR_MAX_LIMBS];
    memcpy(pm1, f->mod, sizeof(pm1));
    pm1[0] -= 1; /* p is odd, no borrow */
    unsigned s = 0;
    for (int i = 0; i < f->nlimbs; ++i) {
        if (pm1[i] == 0) {
            s += 32;
            continue;
        }
        s += (unsigned)__builtin_ctz(pm1[i]);
        break;
    }
    f->two_adicity = s;
    return 0;
}

const or_field *or_field_get(int field_id)
{
    static or_field fields[3];
    static int ready[3];
    if (field_id < 0 || field_id > 2)
        return NULL;
    if (!ready[field_id]) {
        or_field tmp;
        or_field_init(&tmp, field_id);
        fields[field_id] = tmp;
        ready[field_id] = 1;
    }
    return &fields[field_id];
}

int or_is_canonical(const or_field *f, const uint32_t *a)
{
    return !geq_n(a, f->mod, f->nlimbs);
}

/* Scalar::add, device_field_operators.h:190-195, reducing to [0, p). */
void or_add(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    uint32_t c = add_n(r, a, b, f->nlimbs);
    if (c || geq_n(r, f->mod, f->nlimbs))
        sub_n(r, r, f->mod, f->nlimbs);
}

/* Scalar::subtract, device_field_operators.h:198-204. */
void or_sub(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    uint32_t borrow = sub_n(r, a, b, f->nlimbs);
    if (borrow)
        add_n(r, r, f->mod, f->nlimbs);
}

/* ciosMontgomeryMultiply, device_field_operators.h:151-187 (CIOS, Koc et
 * al.), with the correct m' and a full final subtraction.  The limb count s
 * is a parameter of an always-inlined body so that the 24-limb and 1-limb
 * calls compile to fully unrolled loops (the same statements). */
static inline __attribute__((always_inline)) void cios(const or_field *f, uint32_t *r, const uint32_t *a,
                                                       const uint32_t *b, const int s)
{
    uint32_t t[OR_MAX_LIMBS + 2];
    memset(t, 0, sizeof(t));
    for (int i = 0; i < s; ++i) {
        uint64_t C = 0, x;
        for (int j = 0; j < s; ++j) {
            x = (uint64_t)t[j] + (uint64_t)a[j] * b[i] + C;
            t[j] = (uint32_t)x;
            C = x >> 32;
        }
        x = (uint64_t)t[s] + C;
        t[s] = (uint32_t)x;
        t[s + 1] = (uint32_t)(x >> 32);
        const uint32_t m = t[0] * f->mprime;
        x = (uint64_t)t[0] + (uint64_t)m * f->mod[0];
        C = x >> 32;
        for (int j = 1; j < s; ++j) {
            x = (uint64_t)t[j] + (uint64_t)m * f->mod[j] + C;
            t[j - 1] = (uint32_t)x;
            C = x >> 32;
        }
        x = (uint64_t)t[s] + C;
        t[s - 1] = (uint32_t)x;
        t[s] = t[s + 1] + (uint32_t)(x >> 32);
    }
    if (t[s] || geq_n(t, f->mod, s))
        sub_n(t, t, f->mod, s);
    memcpy(r, t, sizeof(uint32_t) * s);
}

void or_mul(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    if (f->nlimbs == 24)
        cios(f, r, a, b, 24);
    else if (f->nlimbs == 1)
        cios(f, r, a, b, 1);
    else
        cios(f, r, a, b, f->nlimbs);
}

/* Scalar::pow (device_field_operators.h:229-258) restated as
 * square-and-multiply in the Montgomery domain. */
void or_pow(const or_field *f, uint32_t *r, const uint32_t *a, uint64_t e)
{
    uint32_t acc[OR_MAX_LIMBS], base[OR_MAX_LIMBS];
    memcpy(acc, f->one, sizeof(acc));
    memcpy(base, a, sizeof(uint32_t) * f->nlimbs);
    while (e) {
        if (e & 1)
            or_mul(f, acc, acc, base);
        e >>= 1;
        if (e)
            or_mul(f, base, base, base);
    }
    memcpy(r, acc, sizeof(uint32_t) * f->nlimbs);
}

static void pow_big(const or_field *f, uint32_t *r, const uint32_t *a,
                    const uint32_t *e, int elimbs)
{
    uint32_t acc[OR_MAX_LIMBS];
    memcpy(acc, f->one, sizeof(acc));
    for (int i = elimbs - 1; i >= 0; --i) {
        for (int b = 31; b >= 0; --b) {
            or_mul(f, acc, acc, acc);
            if ((e[i] >> b) & 1)
                or_mul(f, acc, acc, a);
        }
    }
    memcpy(r, acc, sizeof(uint32_t) * f->nlimbs);
}

void or_inverse(const or_field *f, uint32_t *r, const uint32_t *a)
{
    uint32_t e[OR_MAX_LIMBS], two[OR_MAX_LIMBS] = {2};
    sub_n(e, f->mod, two, f->nlimbs);
    pow_big(f, r, a, e, f->nlimbs);
}

void or_to_mont(const or_field *f, uint32_t *r, const uint32_t *a)
{
    or_mul(f, r, a, f->r2);
}

void or_from_mont(const or_field *f, uint32_t *r, const uint32_t *a)
{
    uint32_t one_raw[OR_MAX_LIMBS] = {1};
    or_mul(f, r, a, one_raw);
}

/* ------------------------------------------------------------------- FFT */

/* bitreverse_host, test/fft_host.h:7-16 */
size_t or_bitreverse(size_t n, unsigned l)
{
    size_t r = 0;
    for (unsigned k = 0; k < l; ++k) {
        r = (r << 1) | (n & 1);
        n >>= 1;
    }
    return r;
}

static unsigned ilog2(size_t n)
{
    unsigned l = 0;
    while (((size_t)1 << l) < n)
        ++l;
    return l;
}

#define EL(base, i, s) ((base) + (size_t)(i) * (size_t)(s))

/* _basic_serial_radix2_FFT, test/fft_host.h:18-54 */
void or_serial_radix2_fft(const or_field *f, uint32_t *a, size_t n,
                          const uint32_t *omega, const uint32_t *one,
                          int literal_butterfly)
{
    const int s = f->nlimbs;
    const unsigned logn = ilog2(n);
    uint32_t tmp[OR_MAX_LIMBS], t[OR_MAX_LIMBS], w[OR_MAX_LIMBS], w_m[OR_MAX_LIMBS];

    /* in-place bit-reversal swap (fft_host.h:24-29) */
    for (size_t k = 0; k < n; ++k) {
        const size_t rk = or_bitreverse(k, logn);
        if (k < rk) {
            memcpy(tmp, EL(a, k, s), sizeof(uint32_t) * s);
            memcpy(EL(a, k, s), EL(a, rk, s), sizeof(uint32_t) * s);
            memcpy(EL(a, rk, s), tmp, sizeof(uint32_t) * s);
        }
    }

    size_t m = 1; /* invariant: m = 2^{stage-1} (fft_host.h:31) */
    for (unsigned st = 1; st <= logn; ++st) {
        or_pow(f, w_m, omega, n / (2 * m)); /* fft_host.h:35-36 */
        for (size_t k = 0; k < n; k += 2 * m) {
            memcpy(w, one, sizeof(uint32_t) * s);
            for (size_t j = 0; j < m; ++j) {
                uint32_t *lo = EL(a, k + j, s);
                uint32_t *hi = EL(a, k + j + m, s);
                if (literal_butterfly) {
                    /* fft_host.h:44-48 as written */
                    memcpy(t, w, sizeof(uint32_t) * s);
                    or_mul(f, w, w, hi);
                } else {
                    or_mul(f, t, w, hi);
                }
                or_sub(f, hi, lo, t);
                or_add(f, lo, lo, t);
                or_mul(f, w, w, w_m);
            }
        }
        m *= 2;
    }
}

/* Phase 1 of _basic_parallel_radix2_FFT_inner, test/fft_host.h:76-102, for
 * the partial sums j in [j_begin, j_end).  tmp rows are (n >> log_cpus)
 * elements long and are written (not accumulated) here. */
static void phase1_rows(const or_field *f, const uint32_t *a, size_t n,
                        const uint32_t *omega, unsigned log_cpus,
                        size_t j_begin, size_t j_end, uint32_t *tmp, int threads)
{
    const int s = f->nlimbs;
    const size_t num_cpus = (size_t)1 << log_cpus;
    const unsigned log_m = ilog2(n);
    const size_t rows = (size_t)1 << (log_m - log_cpus);
    long jb = (long)j_begin, je = (long)j_end;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_for(threads))
    for (long jj = jb; jj < je; ++jj) {
        const size_t j = (size_t)jj;
        uint32_t omega_j[OR_MAX_LIMBS], omega_step[OR_MAX_LIMBS], elt[OR_MAX_LIMBS];
        uint32_t temp[OR_MAX_LIMBS];
        uint32_t *row = tmp + (j - j_begin) * rows * s;
        memset(row, 0, sizeof(uint32_t) * rows * s); /* FieldT::zero(), fft_host.h:73 */
        or_pow(f, omega_j, omega, j);
        or_pow(f, omega_step, omega, j << (log_m - log_cpus));
        memcpy(elt, f->one, sizeof(uint32_t) * s); /* FieldT::one(), fft_host.h:85 */
        for (size_t i = 0; i < rows; ++i) {
            for (size_t q = 0; q < num_cpus; ++q) {
                /* invariant: elt = omega^(j * idx) */
                const size_t idx = (i + (q << (log_m - log_cpus))) % ((size_t)1 << log_m);
                or_mul(f, temp, EL(a, idx, s), elt);
                or_add(f, EL(row, i, s), EL(row, i, s), temp);
                or_mul(f, elt, elt, omega_step);
            }
            or_mul(f, elt, elt, omega_j);
        }
    }
}

void or_parallel_phase1_sample(const or_field *f, const uint32_t *a, size_t n,
                               const uint32_t *omega, unsigned log_cpus,
                               size_t j_begin, size_t j_end, uint32_t *tmp_out,
                               int threads)
{
    phase1_rows(f, a, n, omega, log_cpus, j_begin, j_end, tmp_out, threads);
}

/* _basic_parallel_radix2_FFT_inner, test/fft_host.h:56-118 */
void or_parallel_radix2_fft_inner(const or_field *f, uint32_t *a, size_t n,
                                  const uint32_t *omega, unsigned log_cpus,
                                  const uint32_t *one, int literal_butterfly,
                                  int threads)
{
    const int s = f->nlimbs;
    const size_t num_cpus = (size_t)1 << log_cpus;
    const unsigned log_m = ilog2(n);

    if (log_m < log_cpus) { /* fft_host.h:64-68 */
        or_serial_radix2_fft(f, a, n, omega, one, literal_butterfly);
        return;
    }
    const size_t rows = (size_t)1 << (log_m - log_cpus);
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * s * rows * num_cpus);
    if (!tmp)
        abort();

    phase1_rows(f, a, n, omega, log_cpus, 0, num_cpus, tmp, threads);

    uint32_t omega_num_cpus[OR_MAX_LIMBS];
    or_pow(f, omega_num_cpus, omega, num_cpus); /* fft_host.h:103 */
    const long nc = (long)num_cpus;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads_for(threads))
    for (long j = 0; j < nc; ++j) /* fft_host.h:104-108 */
        or_serial_radix2_fft(f, tmp + (size_t)j * rows * s, rows, omega_num_cpus, one,
                             literal_butterfly);

#pragma omp parallel for num_threads(nthreads_for(threads))
    for (long i = 0; i < nc; ++i) /* fft_host.h:110-117 */
        for (size_t j = 0; j < rows; ++j)
            memcpy(EL(a, (j << log_cpus) + (size_t)i, s), tmp + ((size_t)i * rows + j) * s,
                   sizeof(uint32_t) * s);
    free(tmp);
}

void or_inverse_fft(const or_field *f, uint32_t *a, size_t n, const uint32_t *omega)
{
    const int s = f->nlimbs;
    uint32_t winv[OR_MAX_LIMBS], nm[OR_MAX_LIMBS] = {0}, nraw[OR_MAX_LIMBS] = {0}, ninv[OR_MAX_LIMBS];
    or_inverse(f, winv, omega);
    or_serial_radix2_fft(f, a, n, winv, f->one, 0);
    nraw[0] = (uint32_t)n; /* n <= 2^30 */
    or_to_mont(f, nm, nraw);
    or_inverse(f, ninv, nm);
    for (size_t i = 0; i < n; ++i)
        or_mul(f, EL(a, i, s), EL(a, i, s), ninv);
}

/* Extended radix-2 domain H u shift*H, |H| = m = n/2, the construction of
 * libfqfft's extended_radix2_domain (not vendored with the reference; its
 * published algorithm): it lets a field of 2-adicity s evaluate on 2^(s+1)
 * points.  omega_half is a primitive m-th root, shift^m != 1.
 * Forward: with G = shift^m, A(w^i) is the size-m DFT of a_j + a_(j+m) and
 * A(shift w^i) that of shift^j (a_j + G a_(j+m)); outputs in that order.
 * Inverse: the two size-m inverse DFTs give u_j = c_j + c_(j+m) and
 * shift^j (c_j + G c_(j+m)); then c_(j+m) = (u - v) / (1 - G) with
 * v = shift^-j * (second), and c_j = u - c_(j+m). */
void or_extended_radix2_fft(const or_field *f, uint32_t *a, size_t n, const uint32_t *omega_half,
                            const uint32_t *shift, int inverse)
{
    const int s = f->nlimbs;
    const size_t m = n / 2;
    uint32_t G[OR_MAX_LIMBS], t[OR_MAX_LIMBS], u[OR_MAX_LIMBS], v[OR_MAX_LIMBS], sh_i[OR_MAX_LIMBS];
    or_pow(f, G, shift, (uint64_t)m);
    if (!inverse) {
        memcpy(sh_i, f->one, sizeof(uint32_t) * s);
        for (size_t i = 0; i < m; ++i) {
            uint32_t *lo = EL(a, i, s), *hi = EL(a, i + m, s);
            or_mul(f, t, G, hi);
            or_add(f, t, t, lo);
            or_mul(f, t, t, sh_i);
            or_add(f, lo, lo, hi);
            memcpy(hi, t, sizeof(uint32_t) * s);
            or_mul(f, sh_i, sh_i, shift);
        }
        or_serial_radix2_fft(f, a, m, omega_half, f->one, 0);
        or_serial_radix2_fft(f, EL(a, m, s), m, omega_half, f->one, 0);
        return;
    }
    or_inverse_fft(f, a, m, omega_half);
    or_inverse_fft(f, EL(a, m, s), m, omega_half);
    uint32_t shinv[OR_MAX_LIMBS], k[OR_MAX_LIMBS];
    or_inverse(f, shinv, shift);
    or_sub(f, k, f->one, G);
    or_inverse(f, k, k); /* 1 / (1 - G) */
    memcpy(sh_i, f->one, sizeof(uint32_t) * s);
    for (size_t i = 0; i < m; ++i) {
        uint32_t *lo = EL(a, i, s), *hi = EL(a, i + m, s);
        memcpy(u, lo, sizeof(uint32_t) * s);
        or_mul(f, v, hi, sh_i);
        or_sub(f, t, u, v);
        or_mul(f, hi, t, k);
        or_sub(f, lo, u, hi);
        or_mul(f, sh_i, sh_i, shinv);
    }
}

/* ------------------------------------------------------------ G1 curves
 * Short Weierstrass y^2 = x^3 + a x + b over the base field f, points in
 * homogeneous projective coordinates (X : Y : Z), x = X/Z, y = Y/Z, all in
 * Montgomery form; Z = 0 is the identity.  The reference's mnt4753_G1
 * (cuda/device_field.h:296-437) uses these formulas -- addition
 * "add-1998-cmo-2" (:335-352) and doubling "dbl-2007-bl" (:354-378) of the
 * Explicit-Formulas Database -- but over the wrong field (its coordinates
 * are Scalar, the curve's scalar field) and without the exceptional cases;
 * here the coordinates are in the base field and P + P, P + (-P) and the
 * identity are handled.  b never enters the formulas. */
#define PT(p, k) ((p) + (size_t)(k) * (size_t)f->nlimbs)

static int is_zero_el(const or_field *f, const uint32_t *a)
{
    for (int i = 0; i < f->nlimbs; ++i)
        if (a[i])
            return 0;
    return 1;
}

void or_g1_dbl(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p)
{
    const int s = f->nlimbs;
    const size_t eb = sizeof(uint32_t) * (size_t)s;
    if (is_zero_el(f, PT(p, 2))) {
        memmove(r, p, 3 * eb);
        return;
    }
    uint32_t XX[OR_MAX_LIMBS], ZZ[OR_MAX_LIMBS], w[OR_MAX_LIMBS], t[OR_MAX_LIMBS], sv[OR_MAX_LIMBS],
        ss[OR_MAX_LIMBS], sss[OR_MAX_LIMBS], R[OR_MAX_LIMBS], RR[OR_MAX_LIMBS], B[OR_MAX_LIMBS],
        h[OR_MAX_LIMBS], X3[OR_MAX_LIMBS], Y3[OR_MAX_LIMBS];
    const uint32_t *X = PT(p, 0), *Y = PT(p, 1), *Z = PT(p, 2);
    or_mul(f, XX, X, X);
    or_mul(f, ZZ, Z, Z);
    or_mul(f, w, coeff_a, ZZ);          /* w = a ZZ + 3 XX */
    or_add(f, w, w, XX);
    or_add(f, w, w, XX);
    or_add(f, w, w, XX);
    or_mul(f, t, Y, Z);                 /* s = 2 Y Z */
    or_add(f, sv, t, t);
    or_mul(f, ss, sv, sv);
    or_mul(f, sss, sv, ss);
    or_mul(f, R, Y, sv);                /* R = Y s */
    or_mul(f, RR, R, R);
    or_add(f, t, X, R);                 /* B = (X + R)^2 - XX - RR */
    or_mul(f, B, t, t);
    or_sub(f, B, B, XX);
    or_sub(f, B, B, RR);
    or_mul(f, h, w, w);                 /* h = w^2 - 2 B */
    or_sub(f, h, h, B);
    or_sub(f, h, h, B);
    or_mul(f, X3, h, sv);               /* X3 = h s */
    or_sub(f, t, B, h);                 /* Y3 = w (B - h) - 2 RR */
    or_mul(f, Y3, w, t);
    or_sub(f, Y3, Y3, RR);
    or_sub(f, Y3, Y3, RR);
    memcpy(PT(r, 0), X3, eb);
    memcpy(PT(r, 1), Y3, eb);
    memcpy(PT(r, 2), sss, eb);          /* Z3 = s^3 */
}

void or_g1_add(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p, const uint32_t *q)
{
    const int s = f->nlimbs;
    const size_t eb = sizeof(uint32_t) * (size_t)s;
    if (is_zero_el(f, PT(p, 2))) {
        memmove(r, q, 3 * eb);
        return;
    }
    if (is_zero_el(f, PT(q, 2))) {
        memmove(r, p, 3 * eb);
        return;
    }
    const uint32_t *X1 = PT(p, 0), *Y1 = PT(p, 1), *Z1 = PT(p, 2);
    const uint32_t *X2 = PT(q, 0), *Y2 = PT(q, 1), *Z2 = PT(q, 2);
    uint32_t X1Z2[OR_MAX_LIMBS], Y1Z2[OR_MAX_LIMBS], Z1Z2[OR_MAX_LIMBS], u[OR_MAX_LIMBS], uu[OR_MAX_LIMBS],
        v[OR_MAX_LIMBS], vv[OR_MAX_LIMBS], vvv[OR_MAX_LIMBS], R[OR_MAX_LIMBS], A[OR_MAX_LIMBS], t[OR_MAX_LIMBS],
        X3[OR_MAX_LIMBS], Y3[OR_MAX_LIMBS], Z3[OR_MAX_LIMBS];
    or_mul(f, X1Z2, X1, Z2);
    or_mul(f, Y1Z2, Y1, Z2);
    or_mul(f, Z1Z2, Z1, Z2);
    or_mul(f, t, Y2, Z1);               /* u = Y2 Z1 - Y1 Z2 */
    or_sub(f, u, t, Y1Z2);
    or_mul(f, t, X2, Z1);               /* v = X2 Z1 - X1 Z2 */
    or_sub(f, v, t, X1Z2);
    if (is_zero_el(f, v)) {
        if (is_zero_el(f, u)) {         /* P == Q */
            or_g1_dbl(f, coeff_a, r, p);
        } else {                        /* P == -Q */
            memset(r, 0, 3 * eb);
        }
        return;
    }
    or_mul(f, uu, u, u);
    or_mul(f, vv, v, v);
    or_mul(f, vvv, vv, v);
    or_mul(f, R, vv, X1Z2);
    or_mul(f, A, uu, Z1Z2);             /* A = uu Z1Z2 - vvv - 2 R */
    or_sub(f, A, A, vvv);
    or_sub(f, A, A, R);
    or_sub(f, A, A, R);
    or_mul(f, X3, v, A);
    or_sub(f, t, R, A);                 /* Y3 = u (R - A) - vvv Y1Z2 */
    or_mul(f, Y3, u, t);
    or_mul(f, t, vvv, Y1Z2);
    or_sub(f, Y3, Y3, t);
    or_mul(f, Z3, vvv, Z1Z2);
    memcpy(PT(r, 0), X3, eb);
    memcpy(PT(r, 1), Y3, eb);
    memcpy(PT(r, 2), Z3, eb);
}

/* [k] P by double-and-add from the top bit; k has klimbs 32-bit limbs (a
 * plain integer, not Montgomery form). */
void or_g1_mul(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *p, const uint32_t *k,
               int klimbs)
{
    uint32_t acc[3 * OR_MAX_LIMBS];
    memset(acc, 0, sizeof(acc));
    for (int i = klimbs - 1; i >= 0; --i)
        for (int b = 31; b >= 0; --b) {
            or_g1_dbl(f, coeff_a, acc, acc);
            if ((k[i] >> b) & 1u)
                or_g1_add(f, coeff_a, acc, acc, p);
        }
    memcpy(r, acc, sizeof(uint32_t) * 3 * (size_t)f->nlimbs);
}

/* sum_i [k_i] P_i, the naive way (the reference's host multi_exp,
 * test/multiexp.h:3-14, with a correct scalar multiplication); k_i plain
 * integers of klimbs limbs each. */
void or_g1_msm(const or_field *f, const uint32_t *coeff_a, uint32_t *r, const uint32_t *points,
               const uint32_t *k, int klimbs, size_t n)
{
    const size_t ps = 3 * (size_t)f->nlimbs;
    uint32_t acc[3 * OR_MAX_LIMBS], t[3 * OR_MAX_LIMBS];
    memset(acc, 0, sizeof(acc));
    for (size_t i = 0; i < n; ++i) {
        or_g1_mul(f, coeff_a, t, points + i * ps, k + i * (size_t)klimbs, klimbs);
        or_g1_add(f, coeff_a, acc, acc, t);
    }
    memcpy(r, acc, sizeof(uint32_t) * ps);
}

/* ------------------------------------------------------------ mixed radix
 * Transforms of n = n2 * n1 points, n2 = 2^a and n1 odd (a product of 3, 5
 * and 7 dividing p - 1), over omega = w2 * w1 (w2 a primitive n2-th root,
 * w1 a primitive n1-th root; their product is a primitive n-th root).  The
 * reference has no such transform (its FFT is radix-2 only); these restate
 * the DFT's definition and its four-step factorisation. */

/* out[k] = sum_j x[j] omega^(j k), the definition, O(n^2) */
void or_dft_direct(const or_field *f, uint32_t *out, const uint32_t *x, size_t n, const uint32_t *omega)
{
    const int s = f->nlimbs;
    uint32_t wk[OR_MAX_LIMBS], wjk[OR_MAX_LIMBS], t[OR_MAX_LIMBS], acc[OR_MAX_LIMBS];
    memcpy(wk, f->one, sizeof(uint32_t) * s); /* omega^k */
    for (size_t k = 0; k < n; ++k) {
        memset(acc, 0, sizeof(acc));
        memcpy(wjk, f->one, sizeof(uint32_t) * s); /* omega^(j k) */
        for (size_t j = 0; j < n; ++j) {
            or_mul(f, t, EL(x, j, s), wjk);
            or_add(f, acc, acc, t);
            or_mul(f, wjk, wjk, wk);
        }
        memcpy(EL(out, k, s), acc, sizeof(uint32_t) * s);
        or_mul(f, wk, wk, omega);
    }
}

/* out[i] = sum_j x[j] omega^(j ks[i]) for each of the nk indices ks[i]: the
 * DFT's definition at chosen outputs, O(n) per index, so outputs of
 * transforms far too large for or_dft_direct can be checked.  The j range
 * is cut into chunks [b, e); each chunk evaluates its part by Horner's rule,
 * sum_{j in [b,e)} x[j] w^(j - b) with w = omega^k, and is shifted by w^b;
 * the chunks' parts are summed.  The nk Horner chains of a chunk run side by
 * side over the same x[j]. */
static void mul32(const or_field *f, uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    /* one-limb Montgomery product, the s = 1 case of or_mul */
    const uint64_t t = (uint64_t)a[0] * b[0];
    const uint32_t m = (uint32_t)t * f->mprime;
    const uint64_t mp = (uint64_t)m * f->mod[0];
    /* t + m p overflows 64 bits for p > 2^31; its low words sum to 0 mod
     * 2^32, with a carry exactly when they are non-zero */
    const uint64_t u = (t >> 32) + (mp >> 32) + ((uint32_t)t != 0u); /* < 2p */
    r[0] = (uint32_t)(u >= f->mod[0] ? u - f->mod[0] : u);
}

void or_dft_points(const or_field *f, uint32_t *out, const uint32_t *x, size_t n, const uint32_t *omega,
                   const uint64_t *ks, size_t nk, int threads)
{
    const int s = f->nlimbs;
    const size_t es = sizeof(uint32_t) * (size_t)s;
    const int T = nthreads_for(threads);
    const size_t chunks = (size_t)T * 4 < n ? (size_t)T * 4 : (n ? n : 1);
    uint32_t *w = (uint32_t *)malloc(es * (nk ? nk : 1));
    uint32_t *part = (uint32_t *)calloc(chunks * (nk ? nk : 1), es);
    if (!w || !part)
        abort();
    for (size_t i = 0; i < nk; ++i)
        or_pow(f, EL(w, i, s), omega, (uint64_t)(ks[i] % (n ? n : 1)));
    const long nc = (long)chunks;
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
    for (long c = 0; c < nc; ++c) {
        const size_t b = n * (size_t)c / chunks, e = n * (size_t)(c + 1) / chunks;
        uint32_t *acc = EL(part, (size_t)c * nk, s);
        uint32_t t[OR_MAX_LIMBS];
        if (s == 1) {
            const uint32_t p = f->mod[0];
            for (size_t j = e; j-- > b;) {
                const uint32_t xj = x[j];
                for (size_t i = 0; i < nk; ++i) {
                    mul32(f, t, acc + i, w + i);
                    const uint64_t v = (uint64_t)t[0] + xj;
                    acc[i] = (uint32_t)(v >= p ? v - p : v);
                }
            }
        } else {
            for (size_t j = e; j-- > b;) {
                const uint32_t *xj = EL(x, j, s);
                for (size_t i = 0; i < nk; ++i) {
                    uint32_t *ai = EL(acc, i, s);
                    or_mul(f, t, ai, EL(w, i, s));
                    or_add(f, ai, t, xj);
                }
            }
        }
        for (size_t i = 0; i < nk; ++i) {
            uint32_t sh[OR_MAX_LIMBS];
            or_pow(f, sh, EL(w, i, s), (uint64_t)b);
            or_mul(f, EL(acc, i, s), EL(acc, i, s), sh);
        }
    }
    for (size_t i = 0; i < nk; ++i) {
        uint32_t sum[OR_MAX_LIMBS] = {0};
        for (size_t c = 0; c < chunks; ++c)
            or_add(f, sum, sum, EL(part, c * nk + i, s));
        memcpy(EL(out, i, s), sum, es);
    }
    free(w);
    free(part);
}

/* The same transform by the four-step factorisation n = n1 * n2, with
 * j = n2 j1 + j2 and k = k1 + n1 k2:
 *   Y[k1][j2] = omega^(j2 k1) sum_j1 x[n2 j1 + j2] w1'^(j1 k1),  w1' = omega^n2
 *   X[k1 + n1 k2] = sum_j2 Y[k1][j2] w2'^(j2 k2),                w2' = omega^n1
 * (odd DFTs directly, the rows by the radix-2 FFT).  inverse != 0: the
 * inverse transform (omega^-1, then n^-1). */
void or_mixed_fft(const or_field *f, uint32_t *a, size_t n1, size_t n2, const uint32_t *w2, const uint32_t *w1,
                  int inverse)
{
    const int s = f->nlimbs;
    const size_t n = n1 * n2;
    uint32_t om[OR_MAX_LIMBS], w1p[OR_MAX_LIMBS], w2p[OR_MAX_LIMBS], t[OR_MAX_LIMBS], acc[OR_MAX_LIMBS];
    or_mul(f, om, w2, w1);
    if (inverse)
        or_inverse(f, om, om);
    or_pow(f, w1p, om, (uint64_t)n2);
    or_pow(f, w2p, om, (uint64_t)n1);
    uint32_t *Y = (uint32_t *)malloc(sizeof(uint32_t) * s * n);
    uint32_t *col = (uint32_t *)malloc(sizeof(uint32_t) * s * n1);
    uint32_t *dft = (uint32_t *)malloc(sizeof(uint32_t) * s * n1);
    for (size_t j2 = 0; j2 < n2; ++j2) {
        for (size_t j1 = 0; j1 < n1; ++j1)
            memcpy(EL(col, j1, s), EL(a, n2 * j1 + j2, s), sizeof(uint32_t) * s);
        or_dft_direct(f, dft, col, n1, w1p);
        uint32_t tw[OR_MAX_LIMBS], step[OR_MAX_LIMBS];
        or_pow(f, step, om, (uint64_t)j2); /* omega^j2 */
        memcpy(tw, f->one, sizeof(uint32_t) * s);
        for (size_t k1 = 0; k1 < n1; ++k1) {
            or_mul(f, EL(Y, k1 * n2 + j2, s), EL(dft, k1, s), tw);
            or_mul(f, tw, tw, step);
        }
    }
    for (size_t k1 = 0; k1 < n1; ++k1)
        or_serial_radix2_fft(f, EL(Y, k1 * n2, s), n2, w2p, f->one, 0);
    uint32_t ninv[OR_MAX_LIMBS], nraw[OR_MAX_LIMBS] = {0}, nm[OR_MAX_LIMBS];
    nraw[0] = (uint32_t)n;
    or_to_mont(f, nm, nraw);
    or_inverse(f, ninv, nm);
    for (size_t k1 = 0; k1 < n1; ++k1)
        for (size_t k2 = 0; k2 < n2; ++k2) {
            memcpy(t, EL(Y, k1 * n2 + k2, s), sizeof(uint32_t) * s);
            if (inverse)
                or_mul(f, t, t, ninv);
            memcpy(EL(a, k1 + n1 * k2, s), t, sizeof(uint32_t) * s);
        }
    (void)acc;
    free(Y);
    free(col);
    free(dft);
}
