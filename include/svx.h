/* libsvx -- MI355X (gfx950) statevector engine for the QSV2 MUX2 record stream.
 *
 * C ABI: plain pointers and sizes, no torch or HIP types in the signatures.  `state` and the other
 * buffer arguments are device pointers (e.g. a torch tensor's data_ptr()); `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Every function returns 0 on success or a
 * negative SVX_E_* code; svx_last_error() returns the calling thread's last message.  Launches are
 * asynchronous on `stream` unless the comment says the call synchronises.
 *
 * dtype: SVX_C64 = complex64 (interleaved float32 re, im -- the numpy / torch layout of the
 * reference, kernels/sv/sv_engine.cpp:4-5), SVX_C128 = complex128 (the error judge).
 *
 * Ownership (as in the reference, host/sv/runtime.py:173-203): the caller owns every state /
 * scratch buffer; the library never frees caller memory.  The library copies each QSV2 stream it
 * is given into its own device-resident descriptor cache before returning, so the host array may
 * be reused at once; a stream already seen is not uploaded again (replays launch kernels only).
 */
#ifndef SVX_H
#define SVX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVX_C64 0
#define SVX_C128 1

#define SVX_OK 0
#define SVX_E_ARG (-1)      /* bad argument (maps to ValueError) */
#define SVX_E_MAGIC (-2)    /* stream does not start with the QSV2 magic (AssertionError, host/sv/model.py:9) */
#define SVX_E_FORMAT (-3)   /* malformed stream: truncated, bad opcode, window out of range (ValueError) */
#define SVX_E_CAPACITY (-4) /* tile / slice does not fit this device (RuntimeError, host/sv/runtime.py:220-221) */
#define SVX_E_HIP (-5)      /* HIP runtime error (RuntimeError) */
#define SVX_E_NODEV (-6)    /* no usable gfx950 device (RuntimeError) */

/* How a segment is prepared.  Every configuration is per call (no process-wide state): two
 * threads preparing the same stream with different options get separate handles.
 * svx_default_options() fills the defaults; `size` must be sizeof(svx_options). */
#define SVX_FUSE_AUTO 0      /* cluster a pass's MUX2 ops into dense units where the units are cheaper */
#define SVX_FUSE_OFF 1       /* every op on the MUX2 path */
#define SVX_FUSE_DENSE4 2    /* every block of <= 4 tile bits a dense 4-qubit unit */
#define SVX_DENSE_F32 0      /* dense units on FP32 matrix cores, default: 4-qubit units as three real
                                16x16 products on v_mfma_f32_16x16x4f32, 5-qubit units and units whose
                                entries are all 0, +-1, +-i as the real form on v_mfma_f32_32x32x2f32 */
#define SVX_DENSE_SPLIT16 1  /* opt-in: split-FP16 matrix cores (hi 11 bits + lo FP16, lo x lo dropped,
                                block-scaled below 2^15): ~2^-21 error per product relative to the
                                tile/column maximum, so an amplitude 1e-9 below its tile maximum keeps
                                only ~3 significant digits; faster than FP32 matrix cores */
#define SVX_KERNEL_AUTO 0    /* per-pass choice (single-buffer / dense-only kernels) */
#define SVX_KERNEL_SINGLE 1  /* the same choice (kept for callers that set it) */
#define SVX_ORDER_TUNABLE 0  /* XCD-contiguous tile order; svx_tune_prepared may switch passes to identity */
#define SVX_ORDER_IDENTITY 1 /* fixed identity workgroup -> tile order */
#define SVX_ORDER_XCD 2      /* fixed XCD-contiguous order (not tunable) */
#define SVX_MUX_AUTO 0       /* general uncontrolled 2x2 ops on the FP32 matrix cores (v_mfma_f32_4x4x1_16b_f32,
                                FP32 products, one FMA rounding per step), the rest on the FP32 vector units */
#define SVX_MUX_VALU 1       /* every MUX2 op on the FP32 vector units */

typedef struct svx_options {
    int32_t size;            /* sizeof(svx_options) */
    int32_t fuse;            /* SVX_FUSE_* */
    int32_t dense_math;      /* SVX_DENSE_* */
    int32_t kernel;          /* SVX_KERNEL_* */
    int32_t diag_tables;     /* 1 (default): runs of diagonal ops as one table of phases; 0: never */
    int32_t dense_only;      /* 1 (default): split-FP16 passes of 4-qubit units only on the 512-thread
                                dense-only kernel (FP32 passes run on the MUX2 kernel); 0: never */
    int32_t tile_scale;      /* split-FP16 only: 1 (default) per-tile pre-scaling, 0 per-column scaling */
    int32_t tile_order;      /* SVX_ORDER_* */
    int32_t mux_engine;      /* SVX_MUX_* */
    int32_t fuse_width;      /* largest fused block in tile bits (1..5); 0: default 4 */
    float fuse_cost4;        /* cost of one dense 4-qubit unit in FP32 FMAs per amplitude (a general MUX2
                                op costs 8); 0: default; negative: zero cost (every block of >= 2 ops fuses) */
    float fuse_cost5;        /* same for 5-qubit units */
    int32_t reserved[4];
} svx_options;

typedef struct svx_caps {
    int32_t device;          /* HIP device ordinal */
    int32_t cu_count;        /* compute units */
    int64_t lds_per_block;   /* bytes of LDS one workgroup may use */
    int64_t hbm_bytes;       /* total device memory */
    int32_t wavefront;       /* lanes per wavefront */
    int32_t max_k_c64;       /* largest tile (log2 amplitudes) the pass kernel holds, complex64 */
    int32_t max_k_c128;      /* same, complex128 */
    int32_t max_high;        /* largest number of gathered high qubits per pass */
    char arch[64];           /* gcnArchName */
} svx_caps;

/* Library / device set-up.  No reference counterpart (XRT device + xclbin load,
 * host/sv/runtime.py:93-97); fills `out` when non-NULL. */
int svx_init(int32_t device, svx_caps *out);

/* Run one QSV2 segment on `n_slices` contiguous slices of `n_slice_amps` = 2^L amplitudes each
 * (every slice runs the same segment independently).  Replaces the reference's run call
 * `sv_engine(st, ins_off)` started on every slice's compute unit and waited on
 * (kernels/sv/sv_engine.cpp:302, host/sv/runtime.py:227-233; semantics host/sv/model.py:59-71).
 * Accepts reference streams (k <= 12, m <= 4) and retargeted ones (k <= max_k, m <= max_high;
 * a pass with m > 4 carries an 8-word [0x48, h4..h10] record after its header). */
int svx_run_segment(void *state, int64_t n_slice_amps, int32_t n_slices, int32_t dtype,
                    const int32_t *qsv2, int64_t n_words, void *stream);

/* The reference kernel ABI verbatim: `st` is one slice buffer holding 2^L amplitudes followed by
 * the instruction region starting at 256-bit word `ins_off` (kernels/sv/sv_engine.cpp:290-302,
 * host/sv/runtime.py:218-233).  Reads the stream from device memory (synchronises `stream` once to
 * fetch it), then runs it like svx_run_segment with one slice, complex64. */
int svx_engine(void *st, int64_t ins_off, void *stream);

/* svx_run_segment with explicit options (NULL = defaults). */
int svx_run_segment_ex(void *state, int64_t n_slice_amps, int32_t n_slices, int32_t dtype,
                       const int32_t *qsv2, int64_t n_words, const svx_options *opts, void *stream);

/* Fill `out` with the default options. */
int svx_default_options(svx_options *out);

/* Validate and upload a segment once; returns a handle for svx_run_prepared (no reference
 * counterpart: the reference re-uploads instructions per launch, host/sv/runtime.py:218-225).
 * Equal streams with equal options share one handle (reference counted).  The caller owns one
 * reference per successful call and gives it back with svx_release_prepared. */
int svx_prepare(const int32_t *qsv2, int64_t n_words, int32_t dtype, int32_t state_bits, int64_t *handle);
int svx_prepare_ex(const int32_t *qsv2, int64_t n_words, int32_t dtype, int32_t state_bits,
                   const svx_options *opts, int64_t *handle);
/* Give back one reference.  A handle without references stays cached (so replays of the same
 * stream need no upload) until the cache of unreferenced descriptors exceeds its byte limit; its
 * device memory is freed once the last launch that used it has completed.  Using a handle after
 * its last release is an error (SVX_E_ARG), never a fault. */
int svx_release_prepared(int64_t handle);
/* Byte limit of the cache of unreferenced descriptors (default 256 MiB); returns the bytes held by
 * all prepared descriptors (referenced or not) in *held when non-NULL. */
int svx_cache_limit(int64_t limit_bytes, int64_t *held);
/* dtype and state bits a handle was prepared for (callers check their state against them). */
int svx_prepared_shape(int64_t handle, int32_t *dtype, int32_t *state_bits);
/* Launch a prepared segment: kernels only, no allocation or copy (safe under stream capture).
 * `state` must hold 2^state_bits amplitudes of the handle's dtype (svx_prepared_shape). */
int svx_run_prepared(void *state, int64_t handle, void *stream);

/* Pick each pass's workgroup-to-tile order by measurement: every pass of the prepared segment that
 * runs on the single-buffer or dense-only kernel over >= 64 workgroup tiles is launched 1 + reps
 * times with each candidate order (each XCD a contiguous range of tiles, or the identity) on
 * `state`, timed with HIP events (synchronously), and keeps the identity only when it is >= 3%
 * faster.  The order changes which workgroup takes which tile, never the arithmetic, so results
 * stay bit-identical.  `state` must be a state the segment may run on (its contents change: each
 * tunable pass is applied 2 (1 + reps) times).  The chosen orders belong to the handle, which is
 * shared by every caller that prepared the same stream with the same options (svx_run_segment
 * included); launches running concurrently with the tuning use either order, with the same
 * results.  No reference counterpart (the FPGA engine's tile order is fixed by its reader). */
int svx_tune_prepared(void *state, int64_t handle, int32_t reps, void *stream);
/* Continuous batching: run pass passes[i] of prepared segment handles[i] on states[i] (complex64,
 * i < n_items <= 4096) with one launch per kernel shape, the items' tiles side by side in one grid.
 * Each state gets the same arithmetic as in svx_run_prepared, so results are bit-identical alone and
 * in a batch.  Items must touch distinct states.  Asynchronous on `stream`.  No reference
 * counterpart (the reference runs one circuit at a time, host/sv/runtime.py:286-334). */
int svx_run_batch(void *const *states, const int64_t *handles, const int32_t *passes, int32_t n_items,
                  void *stream);

/* Number of passes (kernel launches) of a prepared segment, and the kernel index of pass p. */
int svx_prepared_info(int64_t handle, int32_t *n_passes, int32_t *n_launches);

/* Name of the kernel that runs pass `pass` of a prepared segment ("k_mux<13,4>", "k_pass<double,13,false>",
 * ..., "none" for a pass without effective ops), for profiling reports.  No reference counterpart. */
int svx_prepared_kernel(int64_t handle, int32_t pass, char *name, int32_t cap);

/* Work of pass `pass` of a prepared segment after classification and pass fusion:
 * out[0] MUX2 register groups, out[1] MUX2 ops, out[2] dense units on 4 qubits, out[3] dense units on
 * 5 qubits, out[4] dense units whose entries are all 0, +-1 or +-i (FP32 products exact, so
 * permutations stay bit-exact), out[5] MUX2 ops applied on the FP32 matrix cores (the MUX2 kernel
 * runs general 2x2 ops on v_mfma_f32_4x4x1_16b_f32 unless svx_options.mux_engine is SVX_MUX_VALU),
 * out[6] diagonal tables
 * (runs of diagonal ops applied as one table of phases), out[7] the workgroup-to-tile order of the
 * pass (-2: each XCD takes a contiguous range of tiles, 0: identity; see svx_tune_prepared)
 * (out has 8 entries).  No reference counterpart. */
int svx_prepared_pass_stats(int64_t handle, int32_t pass, int32_t *out);

/* Persistent queue of whole-state complex64 circuits of 2^state_bits amplitudes (state_bits 10..13):
 * a resident kernel (`workgroups` workgroups on a stream of the library's own) takes the jobs from a
 * host-visible ring of `capacity` slots, each workgroup one circuit at a time with the state in LDS,
 * so many circuits run in one launch with no launch per circuit.  A job is a prepared segment of one
 * whole-state pass on the MUX2 kernel (compiled with k = state_bits; FP32 dense units included) and a
 * state it runs on (zero_init: the state starts as |0...0>, nothing is read); the handle must stay
 * referenced until the job is done.  Results are bit-identical to svx_run_prepared on the same state
 * (same device code).  The kernel leaves after idle_us microseconds without a job and is relaunched
 * by submit / done / wait when needed; close stops it.  svx_queue_submit returns SVX_E_CAPACITY when
 * every slot holds an unfinished job.  No reference counterpart (the reference's latency study runs
 * one circuit per call, bench/latency_jitter.py:24-58). */
int svx_queue_open(int32_t state_bits, int32_t capacity, int32_t workgroups, int32_t idle_us, int64_t *queue);
int svx_queue_submit(int64_t queue, void *state, int64_t handle, int32_t zero_init, int64_t *ticket);
int svx_queue_done(int64_t queue, int64_t ticket, int32_t *done);
int svx_queue_wait(int64_t queue, int64_t ticket);
int svx_queue_close(int64_t queue);

/* Qubit permutation perm[new_bit] = old_bit of a 2^n state, bit-exact.
 * Replaces the reference's remap paths: model.remap (host/sv/model.py:74-78), the host transpose
 * of local qubits (host/sv/runtime.py:261-268) and the sv_swap bank exchange
 * (host/sv/runtime.py:235-259, kernels/sv/sv_swap.cpp:10).
 * Out of place (dst and src disjoint): any permutation in one HBM sweep.
 * In place (dst == src, or dst NULL -- the reference's sv_swap also exchanges in place): a sequence
 * of sweeps, each moving only bits inside one 12-bit tile (the low 6 bits plus at most 6 others), so
 * every workgroup reads and rewrites its own tile; needs no second state.  A permutation of bits
 * that all lie in one such tile is one sweep; a bit reversal of n >= 12 qubits takes
 * 1 + ceil((n/2 - 6) / 3) sweeps (svx_permute_plan). */
int svx_permute(void *dst, const void *src, int32_t n, const int32_t *perm_new_to_old,
                int32_t dtype, void *stream);

/* The sweeps svx_permute makes for perm: *n_steps = their number (1 out of place); when steps is not
 * NULL, step i's permutation (position b after it holds what position steps[i n + b] held before)
 * goes to steps[i n .. i n + n - 1] (at most max_steps steps).  Host only.  No reference counterpart. */
int svx_permute_plan(int32_t n, const int32_t *perm_new_to_old, int32_t in_place, int32_t *steps,
                     int32_t max_steps, int32_t *n_steps);

/* The reference swap kernel's contract (kernels/sv/sv_swap.cpp:10): exchange nwords 64-byte words
 * a[offa..offa+nwords) <-> b[offb..).  Ranges must not overlap. */
int svx_swap_words(void *a, void *b, int64_t offa, int64_t offb, int64_t nwords, void *stream);

/* |0...0>: every amplitude 0, amplitude 0 = 1 (host/sv/runtime.py:205-217, device-side zeroing
 * pass + 32-byte write).  A fill kernel, no hipMemset (safe under stream capture). */
int svx_zero_state(void *state, int32_t n, int32_t dtype, void *stream);

/* Product state prod_q (a_q |0> + b_q |1>); amps = n x 4 doubles (a.re, a.im, b.re, b.im) per
 * qubit q.  Computed in FP64 on the device and rounded once.  No reference counterpart (the
 * reference always starts from |0...0>). */
int svx_product_state(void *state, int32_t n, int32_t dtype, const double *amps, void *stream);

/* FP64 reductions (the reference samples from FP32 chunk sums accumulated in FP64,
 * host/sv/runtime.py:47-80): norm = sum |a|^2 and, when n_low > 0, the 2^n_low marginal
 * probabilities of logical qubits 0..n_low-1 located at physical bits fmap[0..n_low-1]
 * (out_marginals[x], bit i of x = logical qubit i).  Deterministic order; synchronises `stream`;
 * outputs are host pointers (out_marginals may be NULL when n_low = 0). */
int svx_reduce_f64(const void *state, int32_t n, int32_t dtype, const int32_t *fmap, int32_t n_low,
                   double *out_norm, double *out_marginals, void *stream);

/* svx_reduce_f64 without allocation, for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc; it synchronises `stream`, so it is not for use while the
 * stream is being captured into a graph).  No counterpart in the
 * reference.  Same arguments, kernels, reduction order and results as svx_reduce_f64; the FP64
 * partials, the bin configuration and the bins live in a library-owned static device array of
 * 4 MiB + 6 KiB per device (held until the process ends), so n_low is limited to 8 (SVX_E_CAPACITY
 * beyond).  Every call on one device shares that array: calls that overlap on one GPU race.
 * Synchronises `stream`. */
int svx_reduce_f64_noalloc(const void *state, int32_t n, int32_t dtype, const int32_t *fmap, int32_t n_low,
                           double *out_norm, double *out_marginals, void *stream);

/* Judge metrics in FP64 (tests/functional_suite.py:90-94, bench/fpga/sv_maxwidth_ref.py:15-25):
 * out[0] = F = |<a|b>|^2 / (<a|a><b|b>), out[1] = 1-F = d^2 - d^4/4 after global-phase alignment
 * of the normalised states (no cancellation), out[2] = max |b/|b| - ph a/|a||, out[3] = d.
 * `a` is the reference.  Synchronises `stream`. */
int svx_fidelity_f64(const void *a, int32_t dta, const void *b, int32_t dtb, int32_t n, double *out,
                     void *stream);

/* Copy / convert between dtypes (c128 -> c64 rounds once; c64 -> c128 exact). */
int svx_convert(void *dst, int32_t dt_dst, const void *src, int32_t dt_src, int64_t n_amps, void *stream);

int svx_sync(void *stream);
const char *svx_last_error(void);
const char *svx_version(void);

#ifdef __cplusplus
}
#endif
#endif
