/**
 *  @file include/stringzillas_amd.h
 *  @brief C ABI of the MI355X-native similarity and fingerprint engines: a drop-in for the `szs_*` symbols of
 *         StringZillas (include/stringzillas/stringzillas.h) plus batch, serving and on-disk tape extensions.
 *
 *  Every reference symbol keeps its name, signature, argument meaning and status codes. The type layouts below
 *  are ABI-identical to include/stringzilla/types.h so a caller compiled against the reference header can link
 *  against `libstringzillas_amd.so` unchanged.
 *
 *  Execution model: every engine runs on an AMD GPU (gfx950). A GPU scope or the default scope (device 0) runs
 *  it; a CPU-cores scope is rejected with `sz_device_code_mismatch_k`, as the reference does for a GPU engine on
 *  a CPU scope (c/stringzillas/levenshtein.cuh:76-79). Inputs and outputs may live in host memory (they are
 *  staged through HBM) or in device memory (used in place).
 */
#ifndef STRINGZILLAS_AMD_H_
#define STRINGZILLAS_AMD_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#pragma region Types shared with include/stringzilla/types.h

#ifndef STRINGZILLA_TYPES_H_
typedef uint8_t sz_u8_t;
typedef uint32_t sz_u32_t;
typedef uint64_t sz_u64_t;
typedef int8_t sz_i8_t;
typedef size_t sz_size_t;                 /* include/stringzilla/types.h:538 */
typedef ptrdiff_t sz_ssize_t;             /* include/stringzilla/types.h:539 */
typedef char const *sz_cptr_t;            /* include/stringzilla/types.h:633 */
typedef sz_i8_t sz_error_cost_t;          /* include/stringzilla/types.h:634 */
typedef sz_u64_t sz_sorted_idx_t;

/** Status codes, include/stringzilla/types.h:699-722. */
typedef enum sz_status_t {
    sz_success_k = 0,
    sz_bad_alloc_k = -10,
    sz_invalid_utf8_k = -12,
    sz_contains_duplicates_k = -13,
    sz_overflow_risk_k = -14,
    sz_unexpected_dimensions_k = -15,
    sz_missing_gpu_k = -16,
    sz_device_code_mismatch_k = -17,
    sz_device_memory_mismatch_k = -18,
    sz_authentication_failed_k = -19,
    sz_status_unknown_k = -1,
} sz_status_t;

/** Capability bits, include/stringzilla/types.h:727-767, plus the ROCm bit this library adds (bit 25). */
typedef enum sz_capability_t {
    sz_cap_serial_k = 1,
    sz_cap_parallel_k = 1 << 2,
    sz_cap_any_k = 0x7FFFFFFF,
    sz_cap_cuda_k = 1 << 21,
    sz_cap_kepler_k = 1 << 22,
    sz_cap_hopper_k = 1 << 23,
    sz_cap_rocm_k = 1 << 25, /**< new: AMD GPU (HIP) capability, no reference counterpart */
    sz_caps_none_k = 0,
} sz_capability_t;

/** Sequence of strings behind two callbacks, include/stringzilla/types.h:1168-1182. */
typedef sz_cptr_t (*sz_sequence_member_start_t)(void const *, sz_sorted_idx_t);
typedef sz_size_t (*sz_sequence_member_length_t)(void const *, sz_sorted_idx_t);
typedef struct sz_sequence_t {
    void const *handle;
    sz_size_t count;
    sz_sequence_member_start_t get_start;
    sz_sequence_member_length_t get_length;
} sz_sequence_t;

/** Memory allocator, include/stringzilla/types.h:859-870. */
typedef void *(*sz_memory_allocate_t)(sz_size_t, void *);
typedef void (*sz_memory_free_t)(void *, sz_size_t, void *);
typedef struct sz_memory_allocator_t {
    sz_memory_allocate_t allocate;
    sz_memory_free_t free;
    void *handle;
} sz_memory_allocator_t;
#endif

/** Arrow-like tapes of `count` strings with `count + 1` offsets; include/stringzillas/stringzillas.h:68-82. */
typedef struct sz_sequence_u32tape_t {
    sz_cptr_t data;
    sz_u32_t const *offsets;
    sz_size_t count;
} sz_sequence_u32tape_t;
typedef struct sz_sequence_u64tape_t {
    sz_cptr_t data;
    sz_u64_t const *offsets;
    sz_size_t count;
} sz_sequence_u64tape_t;

#pragma endregion

#pragma region Version and capabilities (stringzillas.h:33-62)

int szs_version_major(void);                          /* replaces stringzillas.h:33 */
int szs_version_minor(void);                          /* replaces stringzillas.h:38 */
int szs_version_patch(void);                          /* replaces stringzillas.h:43 */
sz_capability_t szs_capabilities_comptime(void);      /* replaces stringzillas.h:49 */
sz_capability_t szs_capabilities_runtime(void);       /* replaces stringzillas.h:56 */
sz_capability_t szs_capabilities(void);               /* replaces stringzillas.h:62 */

/** Replaces stringzillas.h:88. Fills @p alloc with a `hipMallocManaged`-backed allocator (malloc without a GPU). */
sz_status_t sz_memory_allocator_init_unified(sz_memory_allocator_t *alloc, char const **error_message);

#pragma endregion

#pragma region Device scopes (stringzillas.h:101-152)

typedef void *szs_device_scope_t;
sz_status_t szs_device_scope_init_default(szs_device_scope_t *scope, char const **error_message);      /* :107 */
sz_status_t szs_device_scope_init_cpu_cores(sz_size_t cpu_cores, szs_device_scope_t *scope,            /* :114 */
                                            char const **error_message);
sz_status_t szs_device_scope_init_gpu_device(sz_size_t gpu_device, szs_device_scope_t *scope,          /* :122 */
                                             char const **error_message);
sz_status_t szs_device_scope_get_cpu_cores(szs_device_scope_t scope, sz_size_t *cpu_cores,             /* :130 */
                                           char const **error_message);
sz_status_t szs_device_scope_get_gpu_device(szs_device_scope_t scope, sz_size_t *gpu_device,           /* :138 */
                                            char const **error_message);
sz_status_t szs_device_scope_get_capabilities(szs_device_scope_t scope, sz_capability_t *capabilities, /* :146 */
                                              char const **error_message);
void szs_device_scope_free(szs_device_scope_t scope);                                                  /* :152 */

#pragma endregion

#pragma region Levenshtein distances (stringzillas.h:153-292)

typedef void *szs_levenshtein_distances_t;
typedef void *szs_levenshtein_distances_utf8_t;

/** Replaces stringzillas.h:170. Costs are non-negative here (the reference stores distances in unsigned cells). */
sz_status_t szs_levenshtein_distances_init(sz_error_cost_t match, sz_error_cost_t mismatch, sz_error_cost_t open,
                                           sz_error_cost_t extend, sz_memory_allocator_t const *alloc,
                                           sz_capability_t capabilities, szs_levenshtein_distances_t *engine,
                                           char const **error_message);
/** Replaces stringzillas.h:184: `results[q * results_row_stride + c]`; `candidates == NULL` = symmetric. */
sz_status_t szs_levenshtein_distances(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                      sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                      sz_size_t *results, sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:199. */
sz_status_t szs_levenshtein_distances_u32tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                              sz_sequence_u32tape_t const *queries,
                                              sz_sequence_u32tape_t const *candidates, sz_size_t *results,
                                              sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:214. */
sz_status_t szs_levenshtein_distances_u64tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                              sz_sequence_u64tape_t const *queries,
                                              sz_sequence_u64tape_t const *candidates, sz_size_t *results,
                                              sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:223. */
void szs_levenshtein_distances_free(szs_levenshtein_distances_t engine);

/** Replaces stringzillas.h:239. Unlike the reference's CUDA path, affine and non-unit costs run on the GPU too. */
sz_status_t szs_levenshtein_distances_utf8_init(sz_error_cost_t match, sz_error_cost_t mismatch,
                                                sz_error_cost_t open, sz_error_cost_t extend,
                                                sz_memory_allocator_t const *alloc, sz_capability_t capabilities,
                                                szs_levenshtein_distances_utf8_t *engine,
                                                char const **error_message);
/** Replaces stringzillas.h:253. */
sz_status_t szs_levenshtein_distances_utf8(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                           sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                           sz_size_t *results, sz_size_t results_row_stride,
                                           char const **error_message);
/** Replaces stringzillas.h:268. */
sz_status_t szs_levenshtein_distances_utf8_u32tape(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                                   sz_sequence_u32tape_t const *queries,
                                                   sz_sequence_u32tape_t const *candidates, sz_size_t *results,
                                                   sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:283. */
sz_status_t szs_levenshtein_distances_utf8_u64tape(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                                   sz_sequence_u64tape_t const *queries,
                                                   sz_sequence_u64tape_t const *candidates, sz_size_t *results,
                                                   sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:292. */
void szs_levenshtein_distances_utf8_free(szs_levenshtein_distances_utf8_t engine);

#pragma endregion

#pragma region Needleman-Wunsch and Smith-Waterman scores (stringzillas.h:293-434)

typedef void *szs_needleman_wunsch_scores_t;
typedef void *szs_smith_waterman_scores_t;

/** Replaces stringzillas.h:310: `byte_to_class[256]` + row-major `class_substitution_costs[32 * 32]`. */
sz_status_t szs_needleman_wunsch_scores_init(sz_u8_t const *byte_to_class,
                                             sz_error_cost_t const *class_substitution_costs, sz_error_cost_t open,
                                             sz_error_cost_t extend, sz_memory_allocator_t const *alloc,
                                             sz_capability_t capabilities, szs_needleman_wunsch_scores_t *engine,
                                             char const **error_message);
/** Replaces stringzillas.h:325. */
sz_status_t szs_needleman_wunsch_scores(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                        sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                        sz_ssize_t *results, sz_size_t results_row_stride,
                                        char const **error_message);
/** Replaces stringzillas.h:340. */
sz_status_t szs_needleman_wunsch_scores_u32tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                sz_sequence_u32tape_t const *queries,
                                                sz_sequence_u32tape_t const *candidates, sz_ssize_t *results,
                                                sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:355. */
sz_status_t szs_needleman_wunsch_scores_u64tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                sz_sequence_u64tape_t const *queries,
                                                sz_sequence_u64tape_t const *candidates, sz_ssize_t *results,
                                                sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:364. */
void szs_needleman_wunsch_scores_free(szs_needleman_wunsch_scores_t engine);

/** Replaces stringzillas.h:380. */
sz_status_t szs_smith_waterman_scores_init(sz_u8_t const *byte_to_class,
                                           sz_error_cost_t const *class_substitution_costs, sz_error_cost_t open,
                                           sz_error_cost_t extend, sz_memory_allocator_t const *alloc,
                                           sz_capability_t capabilities, szs_smith_waterman_scores_t *engine,
                                           char const **error_message);
/** Replaces stringzillas.h:395. */
sz_status_t szs_smith_waterman_scores(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                      sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                      sz_ssize_t *results, sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:410. */
sz_status_t szs_smith_waterman_scores_u32tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                              sz_sequence_u32tape_t const *queries,
                                              sz_sequence_u32tape_t const *candidates, sz_ssize_t *results,
                                              sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:425. */
sz_status_t szs_smith_waterman_scores_u64tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                              sz_sequence_u64tape_t const *queries,
                                              sz_sequence_u64tape_t const *candidates, sz_ssize_t *results,
                                              sz_size_t results_row_stride, char const **error_message);
/** Replaces stringzillas.h:434. */
void szs_smith_waterman_scores_free(szs_smith_waterman_scores_t engine);

#pragma endregion

#pragma region Rolling fingerprints: MinHash + Count-Min (stringzillas.h:436-536)

typedef void *szs_fingerprints_t;      /* stringzillas.h:457 */
typedef void *szs_fingerprints_utf8_t; /* stringzillas.h:458 */

/**
 *  Replaces stringzillas.h:476. `dimensions` hash dimensions per text; `window_widths` (NULL or count 0: the
 *  defaults {3, 4, 5, 7, 9, 11, 15, 31}); `seed` derives every dimension's multiplier and modulo. When every width
 *  gets the same multiple of 64 dimensions, dimension `d` uses `window_widths[(d / 64) % count]`, otherwise
 *  `window_widths[d % count]` (c/stringzillas/fingerprints.cuh:39-140). `alphabet_size`, `alloc` and
 *  `capabilities` are accepted and, as in the reference, do not change the results. A zero window width is
 *  rejected with `sz_unexpected_dimensions_k`.
 */
sz_status_t szs_fingerprints_init(sz_size_t dimensions, sz_size_t alphabet_size, sz_size_t const *window_widths,
                                  sz_size_t window_widths_count, sz_u64_t seed, sz_memory_allocator_t const *alloc,
                                  sz_capability_t capabilities, szs_fingerprints_t *engine,
                                  char const **error_message);
/**
 *  Replaces stringzillas.h:492. Row `i` of `min_hashes` (at `i * min_hashes_stride` bytes) receives the low 32
 *  bits of each dimension's smallest window hash, or 0xFFFFFFFF for a text shorter than that dimension's window;
 *  `min_counts` the number of windows that reach it (0 when skipped). Outputs may be host or device memory.
 */
sz_status_t szs_fingerprints_sequence(szs_fingerprints_t engine, szs_device_scope_t device, sz_sequence_t const *texts,
                                      sz_u32_t *min_hashes, sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                      sz_size_t min_counts_stride, char const **error_message);
/** Replaces stringzillas.h:509. */
sz_status_t szs_fingerprints_u64tape(szs_fingerprints_t engine, szs_device_scope_t device,
                                     sz_sequence_u64tape_t const *texts, sz_u32_t *min_hashes,
                                     sz_size_t min_hashes_stride, sz_u32_t *min_counts, sz_size_t min_counts_stride,
                                     char const **error_message);
/** Replaces stringzillas.h:526. */
sz_status_t szs_fingerprints_u32tape(szs_fingerprints_t engine, szs_device_scope_t device,
                                     sz_sequence_u32tape_t const *texts, sz_u32_t *min_hashes,
                                     sz_size_t min_hashes_stride, sz_u32_t *min_counts, sz_size_t min_counts_stride,
                                     char const **error_message);
/** Replaces stringzillas.h:536. */
void szs_fingerprints_free(szs_fingerprints_t engine);

/* The UTF-8 engine hashes the same bytes (c/stringzillas/fingerprints.cuh:144-190 forwards every call). */
sz_status_t szs_fingerprints_utf8_init(sz_size_t dimensions, sz_size_t alphabet_size, sz_size_t const *window_widths,
                                       sz_size_t window_widths_count, sz_u64_t seed,
                                       sz_memory_allocator_t const *alloc, sz_capability_t capabilities,
                                       szs_fingerprints_utf8_t *engine, char const **error_message);
sz_status_t szs_fingerprints_utf8_sequence(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                           sz_sequence_t const *texts, sz_u32_t *min_hashes,
                                           sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                           sz_size_t min_counts_stride, char const **error_message);
sz_status_t szs_fingerprints_utf8_u32tape(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                          sz_sequence_u32tape_t const *texts, sz_u32_t *min_hashes,
                                          sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                          sz_size_t min_counts_stride, char const **error_message);
sz_status_t szs_fingerprints_utf8_u64tape(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                          sz_sequence_u64tape_t const *texts, sz_u32_t *min_hashes,
                                          sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                          sz_size_t min_counts_stride, char const **error_message);
void szs_fingerprints_utf8_free(szs_fingerprints_utf8_t engine);

/** New: dimensions of an engine (0 for a null handle). */
sz_size_t szs_fingerprints_dimensions(szs_fingerprints_t engine);
/** New: HIP-event time of the last call's hashing kernels on the device, in milliseconds. */
double szs_fingerprints_last_kernel_ms(szs_fingerprints_t engine);
/**
 *  New, tuning: `kernel` 0 picks automatically (the wave kernel when every window width holds at least 64
 *  dimensions, else the lane kernel), 1 forces the wave-per-text kernel, 2 the lane-per-text kernel;
 *  `segment_bytes` (0: automatic, else within [64, 2^30]) caps the window starts one work unit covers; longer
 *  texts are split and merged. Results never depend on either setting.
 */
sz_status_t szs_fingerprints_tune(szs_fingerprints_t engine, int kernel, sz_size_t segment_bytes,
                                  char const **error_message);

#pragma endregion

#pragma region Unified memory (stringzillas.h:545-551)

void *szs_unified_alloc(sz_size_t size_bytes);           /* replaces stringzillas.h:545 */
void szs_unified_free(void *ptr, sz_size_t size_bytes);  /* replaces stringzillas.h:551 */

#pragma endregion

#pragma region Extensions: independent pairs, mixed batches, serving (no reference counterpart)

/**
 *  Pair-list mode: `results[i] = score(firsts[i], seconds[i])` for `i < firsts->count` (both tapes must hold the
 *  same count). The reference offers only the cross-product shape; its callers score independent pairs as N
 *  1x1 calls. `results` is `sz_size_t*` for the Levenshtein engines and `sz_ssize_t*` for NW/SW, as above.
 */
sz_status_t szs_levenshtein_distances_pairs_u64tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                                    sz_sequence_u64tape_t const *firsts,
                                                    sz_sequence_u64tape_t const *seconds, sz_size_t *results,
                                                    char const **error_message);
sz_status_t szs_levenshtein_distances_utf8_pairs_u64tape(szs_levenshtein_distances_utf8_t engine,
                                                         szs_device_scope_t device,
                                                         sz_sequence_u64tape_t const *firsts,
                                                         sz_sequence_u64tape_t const *seconds, sz_size_t *results,
                                                         char const **error_message);
sz_status_t szs_needleman_wunsch_scores_pairs_u64tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                      sz_sequence_u64tape_t const *firsts,
                                                      sz_sequence_u64tape_t const *seconds, sz_ssize_t *results,
                                                      char const **error_message);
sz_status_t szs_smith_waterman_scores_pairs_u64tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                                    sz_sequence_u64tape_t const *firsts,
                                                    sz_sequence_u64tape_t const *seconds, sz_ssize_t *results,
                                                    char const **error_message);

/** One homogeneous slice of a mixed batch: an engine handle of the given kind plus its pair list. */
typedef enum szs_engine_kind_t {
    szs_engine_levenshtein_k = 0,
    szs_engine_levenshtein_utf8_k = 1,
    szs_engine_needleman_wunsch_k = 2,
    szs_engine_smith_waterman_k = 3,
} szs_engine_kind_t;

typedef struct szs_pairs_segment_t {
    szs_engine_kind_t kind;
    void *engine;                          /**< handle returned by the matching `szs_*_init` */
    sz_sequence_u64tape_t const *firsts;   /**< host or device tapes */
    sz_sequence_u64tape_t const *seconds;
    void *results;                         /**< `count` x 8-byte results (u64 or i64 by kind) */
    /** Optional `offsets[count]` of each tape. For tapes in HBM it spares a device read per tape (0: read it).
     *  When nonzero it must equal `offsets[count]`. */
    sz_size_t firsts_bytes;
    sz_size_t seconds_bytes;
} szs_pairs_segment_t;

/**
 *  Mixed batch: all segments are scored in ONE persistent, length-binned launch. A prepared batch can be replayed
 *  (as a captured HIP graph) when its inputs are device-resident; the "plain" mode instead issues one launch per
 *  pair, the one-launch-per-call baseline.
 */
typedef void *szs_batch_t;
typedef enum szs_batch_mode_t {
    szs_batch_persistent_k = 0, /**< device-side binning + one persistent launch (default) */
    szs_batch_graph_k = 1,      /**< the same, captured once into a HIP graph and replayed */
    szs_batch_per_pair_k = 2,   /**< one kernel launch per pair, no binning: the per-call baseline */
} szs_batch_mode_t;

sz_status_t szs_batch_init(szs_device_scope_t device, szs_pairs_segment_t const *segments, sz_size_t segments_count,
                           szs_batch_mode_t mode, szs_batch_t *batch, char const **error_message);
/**
 *  No reference counterpart; for callers that cannot allocate device memory (tight serving loops, harnesses that
 *  forbid `hipMalloc`). Scores `segments` as `szs_batch_init` (persistent mode) + `szs_batch_run` + `szs_batch_free`
 *  would, with the same kernels and results, but calls no `hipMalloc`/`hipFree`: every device buffer comes from a
 *  static array of `SZS_NOALLOC_SCRATCH_BYTES` per GPU, held from the library's load on that GPU until the process
 *  ends. Tapes and results must be in device memory (`sz_device_memory_mismatch_k` otherwise); an input needing
 *  more than the array holds fails with `sz_bad_alloc_k` (wide multi-stripe pairs run fewer workgroups first).
 *  Calls on one GPU share the array, so they would race: the library serializes them per GPU. The first call on a
 *  device looks the array up there (and, like any call, the first on a scope creates the scope's stream).
 */
#define SZS_NOALLOC_SCRATCH_BYTES (256ull << 20)
sz_status_t szs_batch_run_noalloc(szs_device_scope_t device, szs_pairs_segment_t const *segments,
                                  sz_size_t segments_count, char const **error_message);

/** Runs the prepared batch once and waits for it; results land where the segments point. */
sz_status_t szs_batch_run(szs_batch_t batch, char const **error_message);
/** Queues the batch on its stream without waiting (`szs_batch_wait` joins it). */
sz_status_t szs_batch_launch(szs_batch_t batch, char const **error_message);
sz_status_t szs_batch_wait(szs_batch_t batch, char const **error_message);
/** Total DP cells (sum of n*m, runes for UTF-8) the batch scores per run. */
sz_size_t szs_batch_cells(szs_batch_t batch);
/** Core recurrence work per run in 32-bit VALU lane-ops: each cell weighted by the op count of the cell update in
 *  the tier that scores it (bit-parallel words, packed 16-bit or int32 DP), data movement excluded; a launch
 *  issues at least this many lane-ops, so this over its duration is a lower bound on the VALU work rate. */
sz_size_t szs_batch_core_ops(szs_batch_t batch);
/** Bytes of device memory the prepared batch holds: staged tapes, tasks, bins, stripe-boundary scratch, chain and
 *  int64 rows. Stripe-boundary scratch is capped (4 GiB or a quarter of the free HBM; a server round, its share of
 *  the pool's margin) by running fewer workgroups when a batch holds pairs of millions of columns. */
sz_size_t szs_batch_device_bytes(szs_batch_t batch);
/** Device time of the last run in milliseconds from HIP events on the batch stream: the scoring kernel alone for the
 *  persistent and per-pair modes, the whole replay for the graph mode. */
double szs_batch_last_kernel_ms(szs_batch_t batch);
void szs_batch_free(szs_batch_t batch);

/**
 *  Serving: a continuous-batching request loop. Requests (one segment each) are admitted while their footprint fits
 *  the HBM budget; a dispatcher thread coalesces everything admitted into the next device round while earlier rounds
 *  are still running, so requests that arrive during a launch overlap with it. The server reserves its pool once
 *  (the budget plus an eighth, at least 256 MiB) at init and prepares every round from it.
 */
typedef void *szs_server_t;
typedef sz_u64_t szs_ticket_t;
sz_status_t szs_server_init(szs_device_scope_t device, sz_size_t hbm_budget_bytes, sz_size_t max_inflight_rounds,
                            szs_server_t *server, char const **error_message);
/** Queues a pair-list request. Device-resident tapes are scored in place; host tapes are copied into the pool with
 *  the round. Fails with `sz_bad_alloc_k` if the request alone exceeds the budget. */
sz_status_t szs_server_submit(szs_server_t server, szs_pairs_segment_t const *request, szs_ticket_t *ticket,
                              char const **error_message);
/** Blocks until @p ticket completes; its results are then in the request's host `results` buffer. */
sz_status_t szs_server_wait(szs_server_t server, szs_ticket_t ticket, char const **error_message);
/** Like `szs_server_wait`, and reports when the server saw the request's round finish, in CLOCK_MONOTONIC seconds
 *  (the clock of Python's `time.perf_counter` on Linux), so latency needs no polling by the caller. */
sz_status_t szs_server_wait_timed(szs_server_t server, szs_ticket_t ticket, double *completed_seconds,
                                  char const **error_message);
/** Non-blocking completion check: 1 if done, 0 if pending. */
int szs_server_poll(szs_server_t server, szs_ticket_t ticket);
sz_size_t szs_server_bytes_in_use(szs_server_t server);

/** Serving counters since init. */
typedef struct szs_server_stats_t {
    sz_u64_t rounds;          /**< device rounds launched */
    sz_u64_t requests;        /**< requests admitted into rounds */
    sz_u64_t cells;           /**< DP cells scored by completed rounds */
    sz_u64_t largest_round;   /**< most requests in one round */
    sz_u64_t pending;         /**< requests waiting for admission now */
    sz_u64_t inflight;        /**< rounds on the device now */
    double prepare_seconds;   /**< host time spent staging, binning and queuing rounds */
    double device_seconds;    /**< scoring-kernel time of completed rounds (HIP events) */
} szs_server_stats_t;
sz_status_t szs_server_get_stats(szs_server_t server, szs_server_stats_t *stats);
void szs_server_free(szs_server_t server);

/**
 *  On-disk tapes (no reference counterpart): a 64-byte header ("SZSTAPE1", u64 count, u64 bytes), then u64
 *  offsets[count + 1] and the u8 data, each section 64-byte aligned; the two sections are an Arrow `large_binary`
 *  array's buffers. `szs_tape_load` maps the file and streams it into HBM through page-locked staging; the view
 *  it fills holds device pointers that every scoring call and the server take in place.
 */
typedef void *szs_device_tape_t;
sz_status_t szs_tape_save(char const *path, sz_sequence_u64tape_t const *tape, char const **error_message);
sz_status_t szs_tape_load(szs_device_scope_t device, char const *path, szs_device_tape_t *tape,
                          sz_sequence_u64tape_t *view, char const **error_message);
/** Bytes of string data of a loaded tape (`offsets[count]`), e.g. for `szs_pairs_segment_t.firsts_bytes`. */
sz_size_t szs_tape_bytes(szs_device_tape_t tape);
void szs_tape_free(szs_device_tape_t tape);

/**
 *  Tuning (no reference counterpart): the kernel shape chosen for a pair depends on its rows (the shorter side)
 *  through a per-slot ladder of (shape, most rows) steps plus a shape for anything taller. Slots are
 *  `family * 2 + symbols` for the six recurrence families x {bytes, runes} (the class-cost families score bytes
 *  only: their second slot holds engines whose costs minus the gap bias leave int8, which run the plain
 *  recurrences), then the unit-cost Levenshtein over bytes (12) and over runes (13), the only slots whose ladders
 *  may use bit-parallel shapes.
 *  An autotuner measures alternatives and installs the fastest ladder; changes apply to batches prepared later.
 */
int szs_tuning_slots(void);
int szs_tuning_shapes(void);
/**
 *  Shape `index` as (lanes per pair, rows per lane, cooperative, bit-parallel, packed, chained). Packed shapes keep
 *  two 16-bit cells per register; they score the class-cost families (byte slots of `nw_*` and `sw_*`) when the
 *  engine's range bound admits their rows per lane, and their int32 twin otherwise. Chained shapes
 *  split a tall pair's stripes over 16 workgroups that pass bottom rows through HBM (global families only).
 */
sz_status_t szs_tuning_shape(int index, int *group, int *rows_per_lane, int *cooperative, int *bit_parallel,
                             int *packed, int *chained);
/** 1 if shape `index` keeps int64 cells (the shape pairs beyond the int32 range bound take; a ladder may also
 *  list it for any slot), else 0. */
int szs_tuning_shape_wide(int index);
/** Most rows per lane of a packed 16-bit shape an engine handle admits: 8, 6, 4, 2, or 0 when scores could leave
 *  the 16-bit window (the engine then runs int32 cells everywhere). */
int szs_tuning_packed_rows(void const *engine);
sz_status_t szs_tuning_get_ladder(int slot, int *shapes, int *max_rows, int capacity, int *count, int *beyond_shape);
sz_status_t szs_tuning_set_ladder(int slot, int const *shapes, int const *max_rows, int count, int beyond_shape,
                                  char const **error_message);

#pragma endregion

#ifdef __cplusplus
}
#endif
#endif
