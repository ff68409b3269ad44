/*
 * hs_gpu_msm.h -- C ABI of the MI355X (gfx950) variable-time multi-scalar multiplication for the
 * Helios / Selene curve cycle.  Plain pointers and sizes only; every entry point is callable from
 * C, C++ (the reference's HeliosPoint / SelenePoint layer), ctypes or any FFI.
 *
 * Curves: HS_GPU_HELIOS (coordinates in F_p, p = 2^255-19, group order q),
 *         HS_GPU_SELENE (coordinates in F_q, q = 2^255-gamma, group order p).
 *
 * Formats (identical to the reference):
 *   scalars  n x 32 bytes little-endian, contiguous (reference api_point.cpp:103-109); any
 *            256-bit value is accepted.
 *   points   hs_gpu_jacobian[n]: the reference's helios_jacobian / selene_jacobian struct,
 *            X, Y, Z as uint64_t[5] radix-2^51 limbs (lazy limbs allowed), affine
 *            (X/Z^2, Y/Z^3), Z = 0 is the identity (helios/include/helios.h:41-46,
 *            helios_ops.h:44-64).
 *   out32    32-byte compressed point: canonical x little-endian, bit 255 = parity of canonical
 *            y, identity = 32 zero bytes (helios/src/x64/helios_tobytes.cpp:42-66).
 *
 * Result semantics: by default the result equals the reference CPU MSM byte for byte, including
 * the reference's behaviour for 864 <= n < 2592 (its window-8 recoding drops a final carry,
 * helios_msm_vartime.cpp:127-162 with :256-257).  HS_GPU_EXACT computes sum s_i * P_i exactly.
 *
 * Errors: functions returning int give 0 on success or a negative HS_GPU_E* code;
 * hs_gpu_last_error() describes the most recent failure of the calling thread.
 * Threading: every entry point may be called from several host threads.  Host-pointer MSMs from
 * concurrent callers each take a staging slot of their own (stream, upload buffers, workspace),
 * so they overlap on the device; only their launch sequences are serialised.
 * Sizes: a plain MSM of 2^24 points or more runs as chunks of 2^23 points whose sums are added on
 * the device (same results); batch elements, server requests and fixed-base sets are limited to
 * fewer than 2^24 points each.
 */
#ifndef HS_GPU_MSM_H
#define HS_GPU_MSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HS_GPU_HELIOS 0
#define HS_GPU_SELENE 1

/* flags */
#define HS_GPU_EXACT 1u /* exact sum; do not reproduce the reference's window-8 carry drop */
#define HS_GPU_ERASE 2u /* zero the device copies of the scalars and of their digits after use (the
                           Pedersen and msm_fixed drop-ins always do; helios_pedersen.h:85,
                           helios_msm_fixed.h:151 erase theirs) */

/* status codes */
#define HS_GPU_OK 0
#define HS_GPU_EINVAL (-1)   /* bad argument (curve id, null pointer with n > 0, size) */
#define HS_GPU_EDEVICE (-2)  /* HIP runtime / kernel failure, or no gfx950 device */
#define HS_GPU_ENOMEM (-3)   /* device allocation failed */
#define HS_GPU_EDECODE (-4)  /* a compressed point failed to decode */

/* Layout-identical to the reference's helios_jacobian and selene_jacobian (120 bytes). */
typedef struct hs_gpu_jacobian
{
    uint64_t X[5];
    uint64_t Y[5];
    uint64_t Z[5];
} hs_gpu_jacobian;

/* ---------------------------------------------------------------- drop-in slot replacements
 * Same signature and semantics as the reference's dispatch slots
 *   void helios_msm_vartime(helios_jacobian *result, const unsigned char *scalars,
 *                           const helios_jacobian *points, size_t n)
 * (helios/include/helios_msm_vartime.h:43-72; slot types include/helioselene_dispatch.h:53,56;
 * backends helios_msm_vartime_{x64,avx2,ifma}, src/dispatch.cpp:133-179) and the selene_ twin.
 * n == 0 gives the identity (helios_msm_vartime.cpp:402-406).  The result is written last, so it
 * may alias the inputs.  A device failure (no gfx950 device, a refused allocation, a HIP error)
 * never yields a wrong point: the slot calls the CPU function registered with
 * hs_gpu_set_cpu_fallback (the reference's own backend, e.g. helios_msm_vartime_x64, which
 * returns the reference's bytes), and with none registered it prints the cause and aborts, as the
 * reference's std::bad_alloc would end a C caller.  hs_gpu_msm_ex returns a status instead. */
void hs_gpu_helios_msm_vartime(hs_gpu_jacobian *result, const unsigned char *scalars,
                               const hs_gpu_jacobian *points, size_t n);
void hs_gpu_selene_msm_vartime(hs_gpu_jacobian *result, const unsigned char *scalars,
                               const hs_gpu_jacobian *points, size_t n);

/* CPU fallbacks of the void drop-ins (no reference counterpart: the reference has no device to
 * fail).  msm has the slot signature; msm_fixed the helios_msm_fixed signature over hs_gpu_affine
 * tables.  Either may be NULL (no fallback for that entry point).  hs_gpu_fallback_count()
 * reports how many drop-in calls were answered by a fallback. */
typedef void (*hs_gpu_cpu_msm_fn)(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_jacobian *points,
                                  size_t n);
struct hs_gpu_affine;
typedef void (*hs_gpu_cpu_msm_fixed_fn)(hs_gpu_jacobian *result, const unsigned char *scalars,
                                        const struct hs_gpu_affine *const *tables, size_t n);
int hs_gpu_set_cpu_fallback(int curve, hs_gpu_cpu_msm_fn msm, hs_gpu_cpu_msm_fixed_fn msm_fixed);
uint64_t hs_gpu_fallback_count(void);

/* Status-returning form of the slot, with flags.  Either output may be NULL. */
int hs_gpu_msm_ex(int curve, hs_gpu_jacobian *result, unsigned char out32[32], const unsigned char *scalars,
                  const hs_gpu_jacobian *points, size_t n, unsigned flags);

/* Device-resident form: all pointers are device pointers (HBM), the call is asynchronous on
 * `stream` (a hipStream_t, NULL = the library's stream for this device).  d_out32 receives the
 * 32-byte encoding, d_out_jac (may be NULL) the reference struct. */
int hs_gpu_msm_device(int curve, const void *d_scalars, const void *d_points, size_t n, unsigned flags,
                      void *d_out32, void *d_out_jac, void *stream);

/* No counterpart in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  hs_gpu_msm_device with every
 * scratch buffer carved from one static 64 MiB device array per device: it calls no hipMalloc,
 * hipFree or allocating library routine once the library is initialised (hs_gpu_init; otherwise the
 * first call initialises it).  It takes n <= 65536 points (HS_GPU_EINVAL beyond), uses the built-in
 * plan parameters (not the tuning table) and gives the same results as hs_gpu_msm_device.  The
 * array keeps the last call's contents; every call on a device shares it, so calls that overlap on
 * one device (on different streams, or from several threads without ordering) race. */
int hs_gpu_msm_device_noalloc(int curve, const void *d_scalars, const void *d_points, size_t n, unsigned flags,
                              void *d_out32, void *d_out_jac, void *stream);

/* Waits for all work the library queued on its own stream(s). */
int hs_gpu_synchronize(void);

/* Bytes form: points as 32-byte compressed encodings, decoded on the device like the
 * reference's helios_frombytes / selene_frombytes (helios/src/x64/helios_frombytes.cpp:59-109,
 * selene_frombytes.cpp:52-111), except that 0^32 is read as the identity (the reference test
 * loader's convention, src/tests.cpp:7479-7496).  Returns HS_GPU_EDECODE if any point is invalid. */
int hs_gpu_msm_bytes(int curve, unsigned char out32[32], const unsigned char *scalars, const unsigned char *points32,
                     size_t n, unsigned flags);

/* ---------------------------------------------------------------- batched MSMs
 * Many independent MSMs of uneven sizes, both curves mixed, computed by one launch sequence:
 * MSMs are grouped by curve and window width, every group's work runs as one set of kernels over
 * all of its MSMs, and the groups run on parallel streams.  The reference has no batch call: each
 * element is one helios_msm_vartime / selene_msm_vartime (same formats and result semantics).
 * A descriptor's scalars/points are n x 32 bytes / n reference Jacobian structs; result and out32
 * may be NULL.  hs_gpu_msm_batch takes host pointers and is synchronous.  The plan form takes
 * DEVICE pointers: create once (allocates and uploads the layout), launch many times
 * asynchronously; hs_gpu_batch_capture records the whole launch sequence into one hipGraph,
 * after which hs_gpu_batch_launch replays that graph. */
typedef struct hs_gpu_msm_desc
{
    int curve;
    size_t n;
    const unsigned char *scalars;
    const hs_gpu_jacobian *points;
    hs_gpu_jacobian *result;
    unsigned char *out32;
} hs_gpu_msm_desc;

typedef struct hs_gpu_batch hs_gpu_batch;

int hs_gpu_msm_batch(const hs_gpu_msm_desc *descs, size_t count, unsigned flags);
int hs_gpu_batch_create(const hs_gpu_msm_desc *device_descs, size_t count, unsigned flags, hs_gpu_batch **out);
int hs_gpu_batch_launch(hs_gpu_batch *batch, void *stream);
int hs_gpu_batch_capture(hs_gpu_batch *batch);
size_t hs_gpu_batch_points(const hs_gpu_batch *batch);
void hs_gpu_batch_destroy(hs_gpu_batch *batch);

/* ---------------------------------------------------------------- Pedersen and fixed-base MSM
 * Layout-identical to the reference's helios_affine / selene_affine (x, y as radix-2^51 limbs,
 * 80 bytes; helios/include/helios.h:48-52). */
typedef struct hs_gpu_affine
{
    uint64_t x[5];
    uint64_t y[5];
} hs_gpu_affine;

/* Drop-in for helios_pedersen_commit / selene_pedersen_commit
 * (helios/include/helios_pedersen.h:55-86): blinding * H + sum values[i] * generators[i], one MSM
 * of n + 1 terms with the plain MSM's result semantics; n > SIZE_MAX/32 - 1 gives the identity. */
void hs_gpu_helios_pedersen_commit(hs_gpu_jacobian *result, const unsigned char *blinding, const hs_gpu_jacobian *H,
                                   const unsigned char *values, const hs_gpu_jacobian *generators, size_t n);
void hs_gpu_selene_pedersen_commit(hs_gpu_jacobian *result, const unsigned char *blinding, const hs_gpu_jacobian *H,
                                   const unsigned char *values, const hs_gpu_jacobian *generators, size_t n);

/* Width-5 fixed-base tables [1P .. 16P] for n bases, 16 entries per base written to
 * tables[16 * i + k] = (k + 1) * P_i (canonical limbs); the n-base form of
 * helios_scalarmult_fixed_precompute (helios/include/helios_scalarmult_fixed.h:54-65).  An identity
 * base gets sixteen (0, 0) entries. */
int hs_gpu_fixed_precompute(int curve, const hs_gpu_jacobian *points, size_t n, hs_gpu_affine *tables);

/* Drop-in for helios_msm_fixed / selene_msm_fixed (helios/include/helios_msm_fixed.h:60-61):
 * sum scalars[i] * P_i from the callers' 16-entry tables.  Always the exact sum.  This departs
 * from the reference where its madd-2007-bl (helios_msm_fixed.h:125-136) meets an accumulator
 * equal to the table entry it adds: the reference then drops the pair and continues from the
 * identity (for example bases [P, P] with scalars [s, s] give the identity there, 2 s P here;
 * tests/golden/workload_runs.json holds the reference's bytes for such inputs).  Random inputs
 * never meet the case. */
void hs_gpu_helios_msm_fixed(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_affine *const *tables,
                             size_t n);
void hs_gpu_selene_msm_fixed(hs_gpu_jacobian *result, const unsigned char *scalars, const hs_gpu_affine *const *tables,
                             size_t n);

/* Resident fixed-base sets (no reference counterpart: the reference recomputes per call).  A set
 * keeps n bases and their tables in HBM, at most 9 GiB of tables per set: a signed-binary comb
 * with 16 teeth (2^15 combinations of 2^(16 j) P per base, 2 MiB) up to 4608 bases, else the
 * widest window width w of 14, 13, 12, 10 and 8 bits whose 2^(w-1) multiples per base (64 B
 * each) fit, else 5 bits.  2^12 + 1 bases get the comb (8 GiB), 2^20 bases width 8 (8 GiB).  Each
 * MSM over the set takes one row of n scalars (n x 32 bytes) and has the plain MSM's result
 * semantics (HS_GPU_EXACT as there).
 * method:
 * HS_GPU_FIXED_COMB    the tables with shared doublings (16 columns of 1 doubling for the comb,
 *                      ceil(256/w) windows of w doublings otherwise: helios_msm_fixed's algorithm
 *                      at w = 5), one wavefront per (MSM, column or window);
 * HS_GPU_FIXED_PIPPENGER  the batched Pippenger over the resident bases;
 * HS_GPU_FIXED_AUTO    the faster of the two for this shape.
 * A Pedersen set holds [H, G_0 .. G_{m-1}] (n = m + 1); hs_gpu_pedersen_commit_batch takes
 * count blindings (count x 32 bytes) and count value vectors (count x m x 32 bytes).  Calls on one
 * set are serialised; the device form is asynchronous on `stream`. */
#define HS_GPU_FIXED_AUTO 0
#define HS_GPU_FIXED_COMB 1
#define HS_GPU_FIXED_PIPPENGER 2
typedef struct hs_gpu_fixed hs_gpu_fixed;
int hs_gpu_fixed_create(int curve, const hs_gpu_jacobian *points, size_t n, hs_gpu_fixed **out);
int hs_gpu_fixed_create_device(int curve, const void *d_points, size_t n, void *stream, hs_gpu_fixed **out);
size_t hs_gpu_fixed_size(const hs_gpu_fixed *set);
int hs_gpu_fixed_msm(hs_gpu_fixed *set, const unsigned char *scalars, size_t count, unsigned flags, int method,
                     unsigned char *out32, hs_gpu_jacobian *results);
int hs_gpu_fixed_msm_device(hs_gpu_fixed *set, const void *d_scalars, size_t count, unsigned flags, int method,
                            void *d_out32, void *d_out_jac, void *stream);
int hs_gpu_pedersen_commit_batch(hs_gpu_fixed *set, const unsigned char *blindings, const unsigned char *values,
                                 size_t count, unsigned flags, int method, unsigned char *out32,
                                 hs_gpu_jacobian *results);
void hs_gpu_fixed_destroy(hs_gpu_fixed *set);

/* ---------------------------------------------------------------- serving loop
 * No reference counterpart (the reference has no request loop).  A server owns `lanes` device
 * lanes (a stream plus device_bytes / lanes of HBM each) and a scheduler thread.  Requests are
 * plain MSM descriptors.  With device_inputs = 0 the scalars/points are host buffers that must
 * stay valid until the ticket completes; with 1 they are device pointers.  out32 / result are
 * host buffers written before the ticket completes.  The scheduler admits queued requests, in
 * order, into any idle lane while their summed device footprint fits the lane (at most max_batch
 * per lane).  Each lane's requests run as one batch, and lanes overlap on the device.  A request
 * whose footprint exceeds a lane fails with HS_GPU_ENOMEM.  Results equal the plain MSM's (same
 * flags semantics).
 * hs_gpu_server_wait returns the request's status (0 or HS_GPU_E*), or 1 if timeout_ms (>= 0)
 * expired first; a completed ticket is released by the wait that returns its status.
 * Latencies are submit-to-completion, measured on the host.  hs_gpu_server_get_stats counts,
 * averages and takes the maximum over every completion since the server started or its stats were
 * reset; p50 and p99 are taken over the most recent 2^20 completions (the server keeps no more). */
typedef struct hs_gpu_server_config
{
    size_t device_bytes; /* HBM for in-flight requests, all lanes together; 0 = 8 GiB */
    int lanes;           /* 0 = 4 */
    int max_batch;       /* 0 = 256 */
    unsigned flags;      /* HS_GPU_EXACT or 0 */
} hs_gpu_server_config;

typedef struct hs_gpu_server_stats
{
    uint64_t completed, batches, points;
    double p50_us, p99_us, max_us, mean_us, elapsed_s;
} hs_gpu_server_stats;

typedef struct hs_gpu_server hs_gpu_server;
int hs_gpu_server_create(const hs_gpu_server_config *config, hs_gpu_server **out);
int hs_gpu_server_submit(hs_gpu_server *server, const hs_gpu_msm_desc *request, int device_inputs, uint64_t *ticket);
int hs_gpu_server_wait(hs_gpu_server *server, uint64_t ticket, int timeout_ms);
int hs_gpu_server_get_stats(hs_gpu_server *server, hs_gpu_server_stats *out);
int hs_gpu_server_reset_stats(hs_gpu_server *server);
void hs_gpu_server_destroy(hs_gpu_server *server); /* drains queued and running requests first */

/* ---------------------------------------------------------------- autotuner
 * No reference counterpart (the reference's autotune picks CPU backends, src/dispatch.cpp:200-360).
 * hs_gpu_autotune measures, on this device, candidate plan parameters for MSMs of 2^lg points for
 * each lg in [lg_lo, lg_hi] (6 <= lg_lo <= lg_hi <= 22): the window width c, the run length L of
 * the flat bucket kernel (sorted entries per thread) and the reduction segment G (buckets per
 * running-sum segment).  mode 0 tunes the plain MSM (latency of one MSM); mode 1 tunes a group of
 * a batch (time of many MSMs of that size in one batch).  It searches coordinate-wise (c, then L,
 * then G) from the built-in defaults (plain MSMs: c = floor(log2 n) - 4 clamped to [4, 16];
 * L = 16 and G = 4 below 2^15 points, L = 32 and G = 8 up to 2^19, L = 64 and G = 8 from 2^19),
 * over L in {8, 16, 32, 64} and G in
 * {2, 4, 8, 16}, keeping a candidate only when it is more than 3% faster
 * (best of three timed runs), installs the result into the library's table (used by all
 * later plans, batches and server admissions of that size class +-1) and appends lines
 * "mode curve lg c L G ms" to `path` when it is not NULL.  hs_gpu_tune_load reads such a file (the
 * environment variable HS_GPU_TUNE_FILE names one to load at first use); hs_gpu_tune_get reports
 * the parameters a plan of n points would use; hs_gpu_tune_clear restores the defaults.
 * Tuning does not change results, only speed. */
int hs_gpu_autotune(int curve, int mode, int lg_lo, int lg_hi, const char *path);
int hs_gpu_tune_load(const char *path);
int hs_gpu_tune_get(int curve, int mode, size_t n, int *c, int *L, int *G);
int hs_gpu_tune_clear(void);

/* ---------------------------------------------------------------- hash-to-curve
 * Simplified SWU map of n 32-byte inputs (RFC 9380 6.6.2; reference helios_map_to_curve /
 * selene_map_to_curve, helios/src/x64/helios_map_to_curve.cpp:184-189).  Outputs the reference
 * Jacobian struct with Z = 1 and/or the compressed encoding; either output may be NULL. */
int hs_gpu_map_to_curve(int curve, const unsigned char *u, size_t n, hs_gpu_jacobian *out_points,
                        unsigned char *out32);
int hs_gpu_map_to_curve_device(int curve, const void *d_u, size_t n, void *d_points, void *d_out32, void *stream);

/* ---------------------------------------------------------------- measurement
 * No reference counterpart.  hs_gpu_profile(1) (re)starts profiling: every plain MSM then records
 * HIP events between its phases on the stream it runs on.  hs_gpu_phase_times() returns the last
 * MSM's phase durations (ms), hs_gpu_phase_summary() the per-phase totals over all profiled MSMs
 * of one curve (curve = -1: both) and their count; phases are named by hs_gpu_phase_name(i).
 * hs_gpu_probe_int_mul_peak() measures the device's 32x32->64 multiply throughput
 * (v_mad_u64_u32 issue rate, partial products per second). */
int hs_gpu_profile(int on);
int hs_gpu_phase_times(float *ms, int cap);
int hs_gpu_phase_summary(int curve, float *ms, int cap, int *calls);
const char *hs_gpu_phase_name(int i);
int hs_gpu_probe_int_mul_peak(double *products_per_second);
/* Compute units, peak engine clock (kHz) and name of the library's device: the nominal issue
 * rate of the multiply is compute_units x 4 SIMDs x 16 lanes x clock. */
int hs_gpu_device_info(int *compute_units, int *max_clock_khz, char *name, size_t name_cap);
/* Copies intermediate state of the last plain MSM on this device (inspection in tests): which = 0
 * window sums, 1 segment sums (per window, partials over groups of up to 256 segments), 2 bucket sums, 5 partial sums of the bucket runs cut by chunk
 * edges (as 32-byte encodings), 3 sorted bucket entries, 4 bucket offsets (uint32).  *count receives the number of items written. */
int hs_gpu_debug_dump(int curve, int which, void *out, size_t cap_bytes, size_t *count);

/* ---------------------------------------------------------------- device memory
 * No reference counterpart.  The library keeps its workspaces between calls (no allocation on the
 * steady-state path).  hs_gpu_set_device_limit caps the device bytes it may hold (0 = no cap; the
 * environment variable HS_GPU_DEVICE_LIMIT sets the initial cap); an allocation past the cap fails
 * with HS_GPU_ENOMEM.  hs_gpu_device_bytes reports the bytes held; hs_gpu_trim waits for the
 * library's streams and frees every cached workspace (resident sets, batch plans and servers keep
 * theirs). */
int hs_gpu_set_device_limit(size_t bytes);
size_t hs_gpu_device_bytes(void);
int hs_gpu_trim(void);

/* ---------------------------------------------------------------- library */
int hs_gpu_init(void);              /* selects device 0 (or the current device); idempotent */
const char *hs_gpu_last_error(void);
const char *hs_gpu_version(void);
/* Window width the library uses for an n-point MSM on `curve`. */
int hs_gpu_window_bits(int curve, size_t n);
/* Window width the reference CPU MSM uses (helios_msm_vartime.cpp:248-263), 0 for n <= 16. */
int hs_gpu_reference_window_bits(size_t n);

#ifdef __cplusplus
}
#endif
#endif
