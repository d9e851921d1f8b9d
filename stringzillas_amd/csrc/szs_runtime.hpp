/**
 *  @file szs_runtime.hpp
 *  @brief Host-side objects of the runtime: device buffers, scopes, engines and prepared batches.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/stringzillas_amd.h"
#include "szs_device.hpp"

namespace szs_amd {

sz_status_t fail(sz_status_t status, char const **error_message, char const *text);
sz_status_t ok(char const **error_message);
sz_status_t from_hip(hipError_t error, char const **error_message, char const *context);
bool is_device_pointer(void const *pointer);

/** Runs `body` so that no C++ exception crosses the C ABI: a failed host allocation becomes `sz_bad_alloc_k`. */
template <typename callable_type_>
sz_status_t guarded(char const **error_message, callable_type_ &&body) noexcept {
    try {
        return body();
    }
    catch (std::bad_alloc const &) {
        return fail(sz_bad_alloc_k, error_message, "Host memory allocation failed");
    }
    catch (...) {
        return fail(sz_status_unknown_k, error_message, "Unexpected failure on the host");
    }
}

/**
 *  Makes `stream` wait for the work already queued on the device's default (null) stream: device-resident inputs
 *  written there (by torch's default stream, for one) are complete before this library's non-blocking stream
 *  reads them. Called at the start of every call that may read caller-owned device memory.
 */
hipError_t order_after_default_stream(hipStream_t stream);

/**
 *  A device memory pool carved from one allocation: first-fit over an offset-ordered free list, coalescing on
 *  release. The serving loop prepares its rounds from a pool so that it never calls `hipMalloc` or `hipFree`
 *  (which waits for the whole device) while other rounds are running.
 */
struct arena_t {
    char *base = nullptr;
    size_t capacity = 0;
    std::map<size_t, size_t> free_blocks; /**< offset -> bytes */
    std::mutex mutex;
    bool owned = true;     /**< `base` came from `init` (freed with the arena), not from `adopt` */
    bool exclusive = false; /**< buffers that do not fit fail instead of falling back to `hipMalloc` */
    hipError_t init(size_t bytes);
    /** Carves the pool from memory the caller keeps (e.g. a static device array); never freed here. */
    void adopt(void *memory, size_t bytes);
    void *allocate(size_t bytes);
    void deallocate(void *pointer, size_t bytes);
    ~arena_t();
};

/** The pool `device_buffer_t::reserve` draws from on the calling thread (null: plain `hipMalloc`). */
arena_t *&thread_arena();

/** Grow-only device allocation. */
struct device_buffer_t {
    void *data = nullptr;
    size_t capacity = 0;
    arena_t *arena = nullptr; /**< the pool `data` came from, if any */
    hipError_t reserve(size_t bytes);
    void release();
    template <typename type_>
    type_ *as() const noexcept {
        return reinterpret_cast<type_ *>(data);
    }
    ~device_buffer_t() { release(); }
    device_buffer_t() = default;
    device_buffer_t(device_buffer_t const &) = delete;
    device_buffer_t &operator=(device_buffer_t const &) = delete;
};

enum scope_kind_t { scope_default_k = 0, scope_cpu_k = 1, scope_gpu_k = 2 };

struct batch_t;

/** `szs_device_scope_t` target (c/stringzillas/stringzillas.cuh:241-321 in the reference). */
struct device_scope_t {
    scope_kind_t kind = scope_default_k;
    int gpu_device = 0;
    size_t cpu_cores = 0;
    hipStream_t stream = nullptr;
    std::mutex mutex;
    /** Reference-shaped and pair-list calls on this scope run one at a time (as the reference's executor admits one
     *  driver) and restage into `workspace`, whose device buffers only grow: after the first calls no call
     *  allocates device memory (the reference keeps such scratch per engine, serial.hpp:3030-3040). */
    std::mutex call_mutex;
    batch_t *workspace = nullptr;
    ~device_scope_t();
};

sz_status_t resolve_gpu(device_scope_t *scope, int *device, hipStream_t *stream, char const **error_message);

/** Engine handle: host parameters plus the reach-bound inputs. */
struct engine_t {
    int kind = 0; /**< `szs_engine_kind_t` */
    engine_params_t params;
    int substitution_magnitude = 0;
    int gap_magnitude = 0;
    int packed_rows = 0; /**< most rows per lane of a packed 16-bit sweep whose range bound holds (0: none) */
    std::mutex mutex; /**< one caller at a time, as the reference's engines are not re-entrant */
};

/** Most rows per lane (8, 6, 4, 2) a packed 16-bit sweep admits for these parameters, 0 for none. */
int packed_rows_limit(engine_params_t const &params);

sz_status_t make_engine(int kind, int match, int mismatch, int open, int extend, uint8_t const *byte_to_class,
                        int8_t const *class_costs, void **handle, char const **error_message);

/** A tape normalized for the device: u64 offsets, data and offsets both device-resident. */
struct staged_tape_t {
    uint8_t const *data = nullptr;
    uint64_t const *offsets = nullptr;
    uint64_t count = 0;
    uint64_t bytes = 0;             /**< offsets[count], the caller's stated count, or `unknown_bytes_k` */
    uint32_t segment = 0;           /**< the batch segment this tape belongs to */
    uint32_t *runes = nullptr;      /**< UTF-8 only: u32[bytes] */
    uint32_t *rune_counts = nullptr;
    device_buffer_t data_copy, offsets_copy, runes_buffer, rune_counts_buffer;
    std::vector<uint64_t> host_offsets; /**< host offsets widened for upload, kept until the tape is staged again */
    /** Forgets the staged view (keeping the buffers' capacity) so the tape can be staged again. */
    void reset_view() {
        data = nullptr, offsets = nullptr, count = 0, bytes = 0, segment = 0, runes = nullptr, rune_counts = nullptr;
    }
};

/** Stages a host or device tape of raw bytes on `stream` (device data and u64 offsets; copies only what is needed). */
sz_status_t stage_bytes_tape(sz_sequence_u64tape_t const *input, staged_tape_t &tape, hipStream_t stream,
                             char const **error_message);
sz_status_t stage_bytes_tape(sz_sequence_u32tape_t const *input, staged_tape_t &tape, hipStream_t stream,
                             char const **error_message);

/** How a segment's cells are enumerated. */
enum segment_mode_t { segment_pairs_k = 0, segment_cross_k = 1, segment_symmetric_k = 2 };

/** Host description of one segment before staging. */
struct segment_input_t {
    engine_t *engine = nullptr;
    segment_mode_t mode = segment_pairs_k;
    sz_sequence_u64tape_t const *queries = nullptr;
    sz_sequence_u32tape_t const *queries32 = nullptr;
    sz_sequence_u64tape_t const *candidates = nullptr;
    sz_sequence_u32tape_t const *candidates32 = nullptr;
    void *results = nullptr;
    uint64_t results_row_stride = 0;
    uint64_t queries_bytes = 0;    /**< `offsets[count]` of the queries when known (0: read it from the offsets) */
    uint64_t candidates_bytes = 0; /**< the same for the candidates */
};

/** How a batch is prepared beyond its inputs and mode. */
struct batch_options_t {
    hipStream_t stream = nullptr;   /**< overrides the scope's stream when set */
    bool defer_overflow = false;    /**< flag failing segments (overflow, bad tape) instead of failing the batch */
    char *pinned_results = nullptr; /**< page-locked host staging for pair-list results (null: copy directly) */
    size_t pinned_capacity = 0;     /**< bytes at `pinned_results` */
    /** Latency over throughput: a pair that a one-wave shape would sweep in more than this many stripes takes the
     *  cooperative shape of the same geometry, so no single wave sets the end of a short launch (0: off). */
    int32_t latency_stripes = 0;
    /** Most bytes of stripe-boundary scratch the launch may use (0: 4 GiB or a quarter of the free HBM, whichever
     *  is less): with wide multi-stripe pairs the persistent launch then runs fewer workgroups, each holding its
     *  rows. */
    size_t max_scratch_bytes = 0;
};

/** A batch staged on one device: everything `run` needs, ready to be launched (or replayed from a graph). */
struct batch_t {
    int device = 0;
    hipStream_t stream = nullptr;
    szs_batch_mode_t mode = szs_batch_persistent_k;
    std::vector<segment_input_t> inputs;
    std::vector<staged_tape_t *> tapes;        /**< owned; queries then candidates per segment */
    std::vector<staged_tape_t *> spare_tapes;  /**< owned; staged by an earlier call, reused by the next one */
    std::vector<device_buffer_t *> spare_results; /**< owned; result buffers of an earlier call */
    bool static_tables_ready = false;          /**< the bin priority tables are already in HBM */
    /** The staging pass's runes, tasks and bin histogram are on the device and no launch has used them yet. */
    bool materialized = false;
    std::vector<void *> host_results;          /**< per segment: the caller's host pointer, or null if device */
    std::vector<uint64_t> results_elements;    /**< per segment: elements spanned by the result matrix */
    std::vector<uint64_t> result_rows, result_columns;
    std::vector<int64_t *> device_results;     /**< per segment: where the kernels write */
    std::vector<device_buffer_t *> result_buffers;

    /** One device block for the per-call tables, uploaded from `table_image` in one copy. */
    device_buffer_t tables;
    std::vector<char> table_image;
    engine_params_t *params_table = nullptr;       /**< per segment */
    segment_view_t *segment_table = nullptr;       /**< per segment */
    materialize_args_t *materialize_table = nullptr; /**< per segment */
    utf8_tape_t *decode_table = nullptr;           /**< per UTF-8 tape */
    shape_table_t *shape_table = nullptr;          /**< the tuned shape ladders, as of staging */
    device_buffer_t tasks, bin_of_task, order, shape_of_task, identity;
    device_buffer_t histogram, cursor, task_start, item_start, priority, position_of_bin, queue_head, scratch, stats;
    device_buffer_t segment_overflow, results_block, scan_blocks;
    device_buffer_t chain_rows, chain_progress;  /**< chained pairs' bottom rows between groups, and progress */
    device_buffer_t wide_rows;                 /**< the int64 shape's per-pair bottom rows */
    uint32_t wide_pairs = 0;                   /**< pairs scored by the int64 shape */
    uint64_t wide_pitch = 0;
    uint32_t chained_pairs = 0;                /**< pairs scored by a chained shape (slots of the chain buffers) */
    uint32_t chain_pitch = 0;
    uint32_t decode_tapes_count = 0;           /**< UTF-8 tapes transcoded by each run */
    uint64_t decode_strings = 0;               /**< strings across those tapes */
    std::vector<char> materialize_args;        /**< one `materialize_args_t` per segment (opaque here) */
    std::vector<uint32_t> overflowed;          /**< per segment: `segment_flag_t` bits raised on the device */
    char *pinned_results = nullptr;            /**< staging of pair-list results, see `batch_options_t` */
    std::vector<size_t> pinned_offsets;        /**< per segment: byte offset in the staging, or SIZE_MAX */
    size_t pinned_bytes = 0;                   /**< bytes of results staged through `pinned_results` */
    uint64_t tasks_count = 0;
    uint64_t cells = 0;                        /**< sum of n*m over non-empty tasks */
    uint64_t core_ops = 0;                     /**< core recurrence lane-ops of those cells (`core_ops_per_64_cells`) */
    uint32_t scratch_capacity = 0;
    uint32_t waves = 0;
    uint32_t lds_bytes = 0;
    std::vector<uint32_t> per_pair_tasks;      /**< per-pair mode: non-empty task ids */
    std::vector<int32_t> per_pair_variants;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t kernel_begin = nullptr, kernel_end = nullptr;
    double last_kernel_ms = 0;

    /** Drops everything one staging set up (tapes and result buffers become spares, the graph is destroyed)
     *  while every device buffer keeps its capacity, so the next `batch_stage` allocates only what grows. */
    void reset();
    /** `reset`, then returns every device buffer (spares included) to where it came from. */
    void release_device_memory();
    /** Bytes of device memory the batch holds (every buffer's capacity). */
    size_t device_bytes() const;
    ~batch_t();
};

/** `sz_unexpected_dimensions_k` if a host-resident offsets array of the input decreases somewhere. */
sz_status_t check_host_offsets(segment_input_t const &input, char const **error_message);

/** Stages and bins a batch on the scope's device (or on `options.stream`). */
sz_status_t batch_prepare(device_scope_t *scope, std::vector<segment_input_t> const &inputs, szs_batch_mode_t mode,
                          batch_t **out, char const **error_message, batch_options_t const &options = {});
/** The same into an existing batch, reusing its device buffers (grown where the new inputs need more). */
sz_status_t batch_stage(batch_t &batch, device_scope_t *scope, std::vector<segment_input_t> const &inputs,
                        szs_batch_mode_t mode, char const **error_message, batch_options_t const &options = {});
/** Queues a staged batch. `ordered`: the caller staged it just now, with no device work of its own in between,
 *  so the staging's ordering after the default stream still holds. */
sz_status_t batch_launch(batch_t *batch, char const **error_message, bool ordered = false);
sz_status_t batch_wait(batch_t *batch, char const **error_message);

} // namespace szs_amd
