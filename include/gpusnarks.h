/*
 * gpusnarks-amd: C ABI of the MI355X (gfx950) finite-field NTT.
 *
 * This is the boundary beneath the reference's only entry point for this
 * path,
 *     template <typename FieldT> void best_fft(std::vector<FieldT> &a, const FieldT &omg);
 * declared at cuda/fft_kernel.h:24-25 and defined at cuda/fft_kernel.cu:117-147
 * (host vector in, in place, natural order in and out).  fft_kernel.h in this
 * directory re-creates that template on top of these functions.
 *
 * Conventions
 *   - Every function returns an int status: GSN_OK (0) or a gsn_status code;
 *     nothing here exits or aborts (the reference prints and calls exit(-1),
 *     cuda/fft_kernel.cu:137-142).  gsn_strerror() names a code and
 *     gsn_last_error() gives the calling thread's last detailed message.
 *   - Elements are little-endian u32 limbs in Montgomery form: 1 limb for
 *     GSN_FIELD_FP32, 24 limbs (96 bytes, the layout of the reference's
 *     fields::Scalar, cuda/device_field.h:75) for the 753-bit fields.
 *   - omega is an element of the same field and representation; a NULL
 *     omega selects the field's standard primitive n-th root of unity.
 *     A given omega must satisfy omega^n = 1 and omega^(n/2) = -1.
 *   - The caller owns its buffers; the library owns plans, twiddle tables,
 *     device workspaces and streams.  `stream` arguments are hipStream_t
 *     values passed as void* (NULL = the default stream).
 *   - Outputs are canonical residues in [0, p), bit-identical to the CPU
 *     restatement of the reference host FFT (test/fft_host.h) with a correct
 *     butterfly.
 */
#ifndef GPUSNARKS_AMD_H
#define GPUSNARKS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field identifiers. */
enum {
    GSN_FIELD_FP32 = 0,        /* p = 3*2^30 + 1, 2-adicity 30 */
    GSN_FIELD_MNT4753_FR = 1,  /* the reference's modulus, cuda/device_field.h:62-65; 2-adicity 15 */
    GSN_FIELD_MNT6753_FR = 2   /* cuda/device_field.h:445-446; 2-adicity 30 */
};

/* Status codes. */
typedef enum {
    GSN_OK = 0,
    GSN_ERR_INVALID_ARGUMENT = 1,
    GSN_ERR_NOT_POWER_OF_TWO = 2,
    GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY = 3,
    GSN_ERR_BAD_OMEGA = 4,
    GSN_ERR_NO_DEVICE = 5,
    GSN_ERR_DEVICE = 6,
    GSN_ERR_OUT_OF_MEMORY = 7,
    GSN_ERR_ADMISSION = 8,
    GSN_ERR_NOT_READY = 9,
    GSN_ERR_INJECTED_FAULT = 10,
    GSN_ERR_SHUTDOWN = 11
} gsn_status;

/* --- library ---------------------------------------------------------- */

/* Name of a status code (static string). */
const char *gsn_strerror(int status);
/* Detailed message of the calling thread's last failing call. */
const char *gsn_last_error(void);
/* 10000*major + 100*minor + patch */
int gsn_version(void);
/* Selects and initialises a HIP device (default: device 0 on first use). */
int gsn_device_init(int device);
/* Bytes per element of a field (4 or 96); 0 for an unknown field. */
size_t gsn_element_bytes(int field);
/* log2 of the largest power-of-two transform the field supports; -1 if unknown. */
int gsn_two_adicity(int field);
/* Writes the standard primitive 2^log_n-th root of unity (Montgomery form). */
int gsn_root_of_unity(int field, unsigned log_n, void *omega_out);
/* Fault injection for tests: the next `count` transform launches fail with
 * GSN_ERR_INJECTED_FAULT before touching the device. */
int gsn_debug_inject_faults(int count);

/* --- device plumbing: memory, streams, events ----------------------------
 * Thin wrappers over the HIP runtime this library links, for callers that
 * have no HIP runtime of their own (or a different one). */
int gsn_device_malloc(void **ptr, size_t bytes);
int gsn_device_free(void *ptr);
int gsn_device_memory(size_t *free_bytes, size_t *total_bytes);
int gsn_device_synchronize(void);
/* Copies; stream may be NULL.  Synchronous with respect to the host. */
int gsn_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream);
int gsn_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream);
int gsn_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream);
int gsn_stream_create(void **stream);
int gsn_stream_destroy(void *stream);
int gsn_stream_synchronize(void *stream);
int gsn_event_create(void **event);
int gsn_event_destroy(void *event);
int gsn_event_record(void *event, void *stream);
int gsn_event_synchronize(void *event);
int gsn_event_elapsed_ms(float *ms, void *start, void *stop);
/* Copies count u32 words with the library's elementwise kernel (one read and
 * one write per word): a known-traffic reference for hardware counters. */
int gsn_debug_copy_u32(void *d_out, const void *d_in, size_t count, void *stream);
/* Runs the 753-bit product on every CU (four waves per SIMD, about a
 * millisecond) and reports the engine clock it ran at, from the core cycle
 * counter against the constant-rate wall clock: the clock behind a measured
 * multiply rate, which under this load sits below the 2.4 GHz peak. */
int gsn_debug_clock_ghz(double *ghz);
/* Measures the v_mad_u64_u32 issue rate (32x32->64 multiply-adds per
 * second, the unit of the 753-bit roofline) with independent multiply-add
 * chains on every CU (eight waves per SIMD, HIP events around one launch of
 * about a millisecond), and the engine clock during that launch. */
int gsn_debug_mad_rate(double *mads_per_s, double *ghz);

/* --- plain transform: the best_fft replacement ---------------------------
 * cuda/fft_kernel.cu:117-147: host vector in, transform in place,
 * synchronous.  inverse != 0 computes the inverse transform (omega^-1 and
 * n^-1 scaling), which the reference does not provide. */
int gsn_fft_host(int field, void *data, size_t n, const void *omega, int inverse);
/* Elements must be canonical (below p) on every transform entry point: the
 * kernels keep intermediate values partly reduced and their bounds start
 * from inputs below p.  gsn_fft_host checks this (GSN_ERR_INVALID_ARGUMENT);
 * the plan, batch and server entry points do not, and give unspecified
 * (memory-safe) results for inputs at or above p. */

/* --- domain helpers: the elementwise steps around a transform ----------
 * None has a counterpart in the reference (its FFT, cuda/fft_kernel.cu:117-147,
 * is the only operation on the path).  Device buffers of n elements (d_out may
 * equal the input), asynchronous on `stream`, canonical inputs and outputs.
 * Elements and constants are in Montgomery form except where a name says
 * otherwise. */
/* d_out[i] = d_a[i] * d_b[i] (evaluation-form product; the reference's
 * Scalar operator*, cuda/device_field.h:115-122, over a vector). */
int gsn_pointwise_mul(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream);
/* Vector forms of the field operators, which the reference has only per
 * element: Scalar operator+ / operator- (cuda/device_field.h:124-140) and
 * operator^ (cuda/device_field.h:151-158, a uint32 exponent there; 64 bits
 * here).  Inputs canonical; x^0 = 1 for every x, 0 included. */
int gsn_vec_add(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream);
int gsn_vec_sub(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream);
int gsn_vec_pow(int field, void *d_out, const void *d_in, size_t n, uint64_t exponent, void *stream);
/* *d_out = sum_i d_a[i] * d_b[i], one element written to device memory:
 * the reference's multiexp over field elements (cuda/multi_exp.cu:104-135,
 * instantiated for Scalar x Scalar at :139; its host version is
 * test/multiexp.h:3-14), on device buffers and asynchronous on `stream`
 * (block sums live in stream-ordered scratch, hipMallocAsync). */
int gsn_inner_product(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream);
/* d_out[i] = d_in[i] * s, s one host element (e.g. n^-1). */
int gsn_scale(int field, void *d_out, const void *d_in, size_t n, const void *s, void *stream);
/* d_out[i] = d_in[i] * g^i (inverse != 0: g^-i), n a power of two, g a nonzero
 * host element: before a forward transform it evaluates on the coset g*H;
 * after an inverse transform with inverse != 0 it undoes the shift. */
int gsn_coset_shift(int field, void *d_out, const void *d_in, size_t n, const void *g, int inverse, void *stream);
/* Plain residues in [0, p) -> Montgomery form (x * R mod p), and back. */
int gsn_to_montgomery(int field, void *d_out, const void *d_in, size_t n, void *stream);
int gsn_from_montgomery(int field, void *d_out, const void *d_in, size_t n, void *stream);

/* --- plans: device-resident twiddles, asynchronous execution ------------ */
typedef struct gsn_plan gsn_plan;
int gsn_plan_create(gsn_plan **plan, int field, unsigned log_n, const void *omega, int inverse);
int gsn_plan_destroy(gsn_plan *plan);
/* A plan over the extended radix-2 domain H u shift*H of n = 2^log_n points
 * (|H| = n/2, H generated by the standard primitive (n/2)-th root): forward
 * evaluates a polynomial of n coefficients on H (outputs 0..n/2-1, natural
 * order) and then on shift*H (outputs n/2..n-1); inverse interpolates back.
 * It reaches one size past the field's 2-adicity, e.g. MNT4-753 Fr at 2^16.
 * shift NULL: the field's smallest quadratic non-residue; shift^(n/2) must
 * not be 1.  No counterpart in the reference, whose FFT is power-of-two
 * radix-2 only (cuda/fft_kernel.cu:117-147); the construction is libfqfft's
 * extended_radix2_domain.  Executes with gsn_plan_execute. */
int gsn_plan_create_ext(gsn_plan **plan, int field, unsigned log_n, const void *shift, int inverse);
/* A plan for n = n1 * 2^a points with an odd n1 > 1 dividing the 3/5/7-smooth
 * part of p - 1 (3 for the 32-bit field; 3 * 5^2 * 7 for MNT4-753 Fr, up to
 * 525; 3^2 * 5^4 * 7^2 for MNT6-753 Fr, up to n1 = 735), a <= the 2-adicity:
 * X_k = sum_j x_j w^(j k) over the primitive n-th root w = w_{2^a} w_{n1}
 * (the standard 2^a-th root times ROOT_ODD^(ODD_MAX / n1) of
 * gsn_field_params.h); inverse != 0: the inverse, with n^-1.  A four-step
 * factorisation: odd DFTs of the n1-element columns (radix 7, 5, 3 passes
 * over all columns), radix-2 transforms of the rows, a transpose; the plan
 * holds two workspaces of n elements.  No counterpart in the
 * reference (radix-2 only); it takes MNT4-753 Fr past 2^16 (up to
 * 2^15 * 525).  Executes with gsn_plan_execute (d_out may equal d_in). */
int gsn_plan_create_mixed(gsn_plan **plan, int field, size_t n, int inverse);
/* Transforms d_in into d_out (device pointers; d_out may equal d_in) on
 * `stream`.  Asynchronous; the plan's workspace is reused, so calls on one
 * plan must not overlap in time. */
int gsn_plan_execute(gsn_plan *plan, const void *d_in, void *d_out, void *stream);
/* gsn_plan_execute without any device allocation: no counterpart in the
 * reference (whose best_fft allocates on every call, cuda/fft_kernel.cu:129-132).
 * It is for callers that cannot allocate -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The intermediate
 * passes run in one library-owned workspace per device, a static device array
 * of GSN_NOALLOC_WORKSPACE_BYTES (64 MiB, held until the process ends), instead
 * of the plan's own workspace, which gsn_plan_execute allocates on first use.
 * Plans of gsn_plan_create only (not _ext / _mixed), whose
 * gsn_plan_workspace_bytes() is at most that limit (e.g. 753-bit fields up to
 * 2^18, the 32-bit field up to 2^23 with the default schedules); larger ones
 * fail with GSN_ERR_INVALID_ARGUMENT.  The workspace needs no clearing (every
 * pass writes what the next reads).  Every call on one GPU shares its
 * workspace: calls that overlap in time on one GPU race, from any plan,
 * stream or thread, so serialise them (one stream, or events). */
#define GSN_NOALLOC_WORKSPACE_BYTES ((size_t)64 << 20)
int gsn_plan_execute_noalloc(gsn_plan *plan, const void *d_in, void *d_out, void *stream);
/* Workspace bytes an execution of the plan needs (0 for one pass). */
size_t gsn_plan_workspace_bytes(const gsn_plan *plan);
/* Device bytes the plan holds (twiddles + workspace). */
size_t gsn_plan_device_bytes(const gsn_plan *plan);
/* Number of kernel launches (passes) per execution. */
int gsn_plan_num_passes(const gsn_plan *plan);

/* --- batches: many transforms, mixed fields, sizes and directions -------- */
typedef struct {
    int field;
    unsigned log_n;
    int inverse;
    const void *omega; /* NULL: standard root */
    const void *d_in;
    void *d_out;       /* may equal d_in */
} gsn_batch_item;

typedef struct gsn_batch gsn_batch;
/* Builds a batch: plans for every (field, size, direction, omega), a
 * schedule of work tiles and, if use_graph != 0, a captured hipGraph. */
int gsn_batch_create(gsn_batch **batch, const gsn_batch_item *items, size_t count, int use_graph);
int gsn_batch_launch(gsn_batch *batch, void *stream);
int gsn_batch_destroy(gsn_batch *batch);
/* One-shot convenience: create, launch, synchronise, destroy. */
int gsn_batch_execute(const gsn_batch_item *items, size_t count, void *stream);
/* Kernel launches one replay of the batch issues. */
int gsn_batch_num_launches(const gsn_batch *batch);

/* --- multi-scalar multiplication on G1 ------------------------------------
 * SWITCHED OFF in this build: gsn_msm_bases_create and gsn_msm return
 * GSN_ERR_INVALID_ARGUMENT ("switched off") -- the first GPU run did not
 * finish and its cause is not yet known.
 *
 * sum_i [s_i] P_i: the curve instantiation of the reference's multiexp
 * (cuda/multi_exp.cu:140, mnt4753_G1 x Scalar), rebuilt as a bucket
 * (Pippenger) method over a prepared table of fixed bases.  Points are
 * homogeneous projective (X, Y, Z), 3 x 96 bytes, coordinates in Montgomery
 * form over the curve's base field, Z = 0 the identity (the result is
 * (0, 0, 0) then); scalars are canonical elements of the scalar field in
 * Montgomery form.  The formulas are the reference's (add-1998-cmo-2 and
 * dbl-2007-bl, cuda/device_field.h:335-378) over the base field, with the
 * exceptional cases (P + P, P + (-P), identity) handled. */
enum {
    GSN_CURVE_MNT4753_G1 = 0, /* y^2 = x^3 + 2x + b over MNT6-753 Fr; scalars in MNT4-753 Fr */
    GSN_CURVE_MNT6753_G1 = 1  /* y^2 = x^3 + 11x + b over MNT4-753 Fr; scalars in MNT6-753 Fr */
};
typedef struct gsn_msm_bases gsn_msm_bases;
/* Prepares n bases (device pointer) for repeated MSMs: a table of
 * [2^(c w)] P_i for every window of c scalar bits (window_bits 0: chosen
 * from n), ceil(753 / c) * n * 288 bytes of device memory.  Asynchronous on
 * `stream`; the bases may be freed once the stream has run. */
int gsn_msm_bases_create(gsn_msm_bases **bases, int curve, const void *d_points, size_t n, unsigned window_bits,
                         void *stream);
int gsn_msm_bases_destroy(gsn_msm_bases *bases);
int gsn_msm_bases_info(const gsn_msm_bases *bases, unsigned *window_bits, size_t *table_bytes);
/* *d_out (one point, device memory) = sum_i [d_scalars[i]] P_i over the
 * prepared bases; asynchronous on `stream` (scratch is stream-ordered). */
int gsn_msm(gsn_msm_bases *bases, void *d_out, const void *d_scalars, void *stream);

/* --- autotuning: per-size pass schedules -----------------------------------
 * A schedule is num_passes triples (log_r, log_w, log_e): pass radix 2^log_r,
 * 2^log_w columns per workgroup, register radix 2^log_e; the log_r sum to
 * log_n.  Setting one drops cached plans from the cache (plans already in
 * use, by batches or by calls running on other threads, stay valid).  At
 * start-up the library reads tuning.txt beside itself (or $GSN_TUNING_FILE):
 * lines "field log_n r,w,e r,w,e ...", and "batch field log_n ..." for the
 * schedule of transforms fused into batches. */
int gsn_tuning_set(int field, unsigned log_n, const int *shapes, int num_passes);
int gsn_tuning_get(int field, unsigned log_n, int *shapes, int max_passes, int *num_passes);
int gsn_tuning_clear(void);
int gsn_tuning_load(const char *path);

/* --- serving: continuous batching of NTT requests ------------------------ */
typedef struct {
    size_t hbm_budget_bytes;   /* admission limit on device bytes in flight (0: 1/2 of free HBM) */
    int max_inflight;          /* max requests on the device at once (0: 4096) */
    int num_streams;           /* device streams (0: 4) */
    int max_batch;             /* max requests fused into one launch group (0: 256) */
} gsn_server_config;

typedef struct gsn_server gsn_server;
int gsn_server_create(gsn_server **server, const gsn_server_config *config);
/* Queues a request; h_in/h_out are host buffers of n elements (may alias)
 * that must stay valid until the request completes.  Returns a ticket. */
int gsn_server_submit(gsn_server *server, int field, unsigned log_n, int inverse, const void *omega,
                      const void *h_in, void *h_out, uint64_t *ticket);
/* Same with operands already resident in device memory (d_in/d_out may
 * alias); nothing is copied. */
int gsn_server_submit_device(gsn_server *server, int field, unsigned log_n, int inverse, const void *omega,
                             const void *d_in, void *d_out, uint64_t *ticket);
/* *done = 1 when the ticket has completed (its status is returned). */
int gsn_server_poll(gsn_server *server, uint64_t ticket, int *done);
/* Blocks until the ticket completes; returns its status. */
int gsn_server_wait(gsn_server *server, uint64_t ticket);
/* Arrival, start and completion times of a finished ticket (seconds on the
 * server's monotonic clock). */
int gsn_server_ticket_times(gsn_server *server, uint64_t ticket, double *t_submit, double *t_start,
                            double *t_done);
int gsn_server_destroy(gsn_server *server);

#ifdef __cplusplus
}
#endif

#endif /* GPUSNARKS_AMD_H */
