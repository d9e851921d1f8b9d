/**
 *  @file batch.hip
 *  @brief Batch preparation and launch: tape staging into HBM, device-side materialization and length binning,
 *         the persistent scoring launch, its HIP-graph replay, and the one-launch-per-pair baseline.
 *
 *  `batch_prepare` may synchronize (it sizes buffers from the materialized tasks); `batch_launch` only enqueues
 *  device work on the scope's stream and is what a graph captures.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <utility>
#include <vector>

#include "szs_runtime.hpp"


extern "C" {
hipError_t szs_amd_launch_fill(uint32_t *data, uint32_t count, uint32_t value, hipStream_t stream);
hipError_t szs_amd_launch_fill_many(szs_amd::fill_list_t const *list, hipStream_t stream);
hipError_t szs_amd_launch_utf8_decode(szs_amd::utf8_tape_t const *tapes, uint32_t tapes_count, uint64_t strings_total,
                                      szs_amd::prepare_stats_t *stats, uint32_t *segment_flags, hipStream_t stream);
hipError_t szs_amd_launch_widen_offsets(uint32_t const *narrow, uint64_t *wide, uint64_t count, hipStream_t stream);
hipError_t szs_amd_launch_materialize(szs_amd::materialize_args_t const *args, uint32_t segments_count,
                                      uint64_t tasks_total, szs_amd::shape_table_t const *table, szs_amd::task_t *tasks,
                                      uint32_t *bin_of_task, uint32_t *histogram,
                                      szs_amd::segment_view_t const *segments, szs_amd::prepare_stats_t *stats,
                                      uint32_t *segment_overflow, hipStream_t stream);
hipError_t szs_amd_launch_scan_bins(szs_amd::bins_view_t const *bins, hipStream_t stream);
hipError_t szs_amd_launch_scatter_bins(szs_amd::bins_view_t const *bins, uint32_t const *bin_of_task,
                                       uint32_t const *position_of_bin, uint32_t tasks_count, hipStream_t stream);
hipError_t szs_amd_launch_persistent(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                     szs_amd::bins_view_t const *bins, int32_t *scratch, uint32_t scratch_capacity,
                                     uint32_t waves, uint32_t lds_bytes, szs_amd::chain_view_t const *chain,
                                     hipStream_t stream);
hipError_t szs_amd_launch_single(szs_amd::task_t const *tasks, szs_amd::segment_view_t const *segments,
                                 uint32_t const *identity, uint32_t task_index, int variant, int32_t *scratch,
                                 uint32_t scratch_capacity, uint32_t lds_bytes, szs_amd::chain_view_t const *chain,
                                 hipStream_t stream);
hipError_t szs_amd_persistent_occupancy(int *blocks_per_cu, uint32_t lds_bytes);
}

namespace szs_amd {

#define SZS_TRY_HIP(expr, context)                                                                                   \
    do {                                                                                                             \
        hipError_t const error_ = (expr);                                                                            \
        if (error_ != hipSuccess) return from_hip(error_, error_message, context);                                   \
    } while (0)

#pragma region Static tables

/**
 *  Default shape choice. Uniform-cost (Levenshtein) families: the smallest single-stripe shape that holds all
 *  rows (65-128 rows on 16 lanes of 8 rows, four pairs per wave, rather than one pair on 64 lanes of 2: a quarter
 *  of the ramp steps, and each step's overhead spread over 8 rows), 1024-row stripes on one wave up to 4096 rows,
 *  then the cooperative 1024-row shape, chained 512-row groups beyond 24,576 rows. Class-cost families likewise,
 *  keeping LDS profiles small: at most 8 rows per lane, 512-row stripes on one wave up to 2048 rows, then the
 *  cooperative 512-row shape (chained beyond 24,576 rows for the global families); the byte slots of affine NW
 *  and of SW take the packed 16-bit shapes where they measured faster than int32 (affine NW from 193 rows, up to
 *  16,384 on one wave; affine SW from 257 rows and linear SW from 513, up to 6,144 on one wave), and an engine
 *  outside the packed range bound runs their int32 twins; linear SW takes 16 lanes of 10 rows for 129-160 rows. Unit-cost bytes and runes: DP up
 *  to 64 rows, then bit-parallel words: 4 lanes of 32-row words up to 128 rows (bytes only), 4 lanes up to 256, 16
 *  lanes of 32-row words up to 512, 16 lanes up to 1024, one wave of 32-row words up to 2048, 4096-row stripes on
 *  one wave up to 16384, then cooperative.
 */
/** Pairs taller than this go to the chained shape (global families): one workgroup each would set the batch's
 *  critical path, 16 share the pair instead. */
constexpr int chained_rows_k = 24576;

static shape_table_t default_shape_table() {
    shape_table_t table;
    std::memset(&table, 0, sizeof(table));
    for (int slot = 0; slot < shape_slots_k; ++slot) {
        for (int shape = 0; shape < shapes_count_k; ++shape) table.limits[slot][shape] = -1;
        std::vector<std::pair<int, int>> ladder; // (shape, most rows)
        int multi = 10;
        int const family = slot / 2;
        if (slot == unit_bytes_slot_k) {
            ladder = {{0, 16}, {1, 32}, {2, 64}, {25, 128}, {11, 256}, {21, 512}, {12, 1024}, {22, 2048}, {13, 16384}};
            multi = 14;
        }
        else if (slot == unit_runes_slot_k) { // 32-row words over runes were slower than 64-row ones at 65-128 rows
            ladder = {{0, 16}, {1, 32}, {2, 64}, {11, 256}, {21, 512}, {12, 1024}, {22, 2048}, {13, 16384}};
            multi = 14;
        }
        else if (family == family_lev_linear_k || family == family_lev_affine_k) {
            for (int shape = 0; shape <= 8; ++shape)
                ladder.push_back({shape == 3 ? 23 : shape, shapes_k[shape].group * shapes_k[shape].rows_per_lane});
            ladder.back().second = 4096; // up to 4 stripes on one wave
            ladder.push_back({10, chained_rows_k});
            multi = 20;
        }
        else if (slot == family_nw_affine_k * 2 + symbols_u8_k) {
            ladder = {{0, 16}, {1, 32}, {2, 64}, {23, 128}, {4, 192}, {15, 256}, {16, 384}, {17, 16384},
                      {18, chained_rows_k}};
            multi = 19;
        }
        else if (slot == family_sw_linear_k * 2 + symbols_u8_k) { // local scores: no chained groups
            ladder = {{0, 16}, {1, 32}, {2, 64}, {23, 128}, {26, 160}, {4, 192}, {5, 256}, {6, 384}, {7, 512},
                      {17, 6144}};
            multi = 18;
        }
        else if (slot == family_sw_affine_k * 2 + symbols_u8_k) {
            ladder = {{0, 16}, {1, 32}, {2, 64}, {23, 128}, {4, 192}, {5, 256}, {16, 384}, {17, 6144}};
            multi = 18;
        }
        else {
            for (int shape = 0; shape <= 7; ++shape)
                ladder.push_back({shape == 3 ? 23 : shape, shapes_k[shape].group * shapes_k[shape].rows_per_lane});
            ladder.back().second = 2048;
            bool const local = family == family_sw_linear_k || family == family_sw_affine_k;
            if (local) multi = 9; // local scores need every stripe's maximum: no chained groups
            else {
                ladder.push_back({9, chained_rows_k});
                multi = 20;
            }
        }
        for (size_t i = 0; i < ladder.size(); ++i) {
            table.order[slot][i] = ladder[i].first;
            table.limits[slot][ladder[i].first] = ladder[i].second;
        }
        table.used[slot] = (int32_t)ladder.size();
        table.multi_pass_shape[slot] = multi;
    }
    return table;
}

/** Guards the active shape table between tuning calls and the batches that copy it. */
static std::mutex &tuning_mutex() {
    static std::mutex mutex;
    return mutex;
}

static shape_table_t &active_shape_table() {
    static shape_table_t table = default_shape_table();
    return table;
}

/** Rough latency of one work item of a bin (its bucket is the eighth-octave of lane-row-steps), heaviest first. */
static double bin_item_cost(uint32_t bin) {
    int const variant = (int)(bin / column_buckets_k);
    int const bucket = (int)(bin % column_buckets_k);
    int const family = variant / shapes_count_k / 2;
    double const per_cell = (family == family_lev_affine_k || family == family_nw_affine_k ||
                             family == family_sw_affine_k) ? 1.8 : 1.0;
    // Chained pairs go first: their groups form a pipeline whose fill is the batch's critical path.
    double const chained = shapes_k[variant % shapes_count_k].chained ? 1e30 : 1.0;
    return std::pow(2.0, bucket / (double)bucket_steps_per_octave_k) * per_cell * chained;
}

/** Whole-workgroup items (chained, then cooperative) come before one-wave items, each part heaviest first. */
static bool whole_workgroup_bin(uint32_t bin) {
    return shapes_k[(bin / column_buckets_k) % shapes_count_k].cooperative;
}

static std::vector<uint32_t> const &bin_priority() {
    static std::vector<uint32_t> priority = [] {
        std::vector<uint32_t> order(bins_count_k);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [](uint32_t a, uint32_t b) {
            if (whole_workgroup_bin(a) != whole_workgroup_bin(b)) return whole_workgroup_bin(a);
            return bin_item_cost(a) > bin_item_cost(b);
        });
        return order;
    }();
    return priority;
}

/** Priority position of the first one-wave bin. */
static uint32_t wave_bins_from() {
    static uint32_t const position = [] {
        std::vector<uint32_t> const &order = bin_priority();
        uint32_t p = 0;
        while (p < bins_count_k && whole_workgroup_bin(order[p])) ++p;
        return p;
    }();
    return position;
}

#pragma endregion

#pragma region Tape staging

/** `bytes_hint` is `offsets[count]` when the caller knows it (0: read it back from device offsets, which only a
 *  UTF-8 tape (sizing its runes) or host data behind device offsets needs; otherwise the kernels read it). */
template <typename offset_type_>
static sz_status_t stage_tape(sz_cptr_t data, offset_type_ const *offsets, uint64_t count, bool runes,
                              uint64_t bytes_hint, staged_tape_t &tape, hipStream_t stream,
                              char const **error_message) {
    tape.count = count;
    bool const device_offsets = is_device_pointer(offsets);
    bool const device_data = is_device_pointer(data);
    bool const need_bytes = runes || !device_data;
    // Offsets as u64 on the device.
    if (count == 0 || !offsets) {
        SZS_TRY_HIP(tape.offsets_copy.reserve(16), "Failed to allocate tape offsets");
        SZS_TRY_HIP(hipMemsetAsync(tape.offsets_copy.data, 0, 16, stream), "hipMemsetAsync failed");
        tape.offsets = tape.offsets_copy.as<uint64_t>();
        tape.bytes = 0;
    }
    else if (sizeof(offset_type_) == 8 && device_offsets) {
        tape.offsets = reinterpret_cast<uint64_t const *>(offsets);
        uint64_t last = bytes_hint ? bytes_hint : unknown_bytes_k;
        if (!bytes_hint && need_bytes) {
            SZS_TRY_HIP(hipMemcpyAsync(&last, offsets + count, 8, hipMemcpyDeviceToHost, stream), "offset readback");
            SZS_TRY_HIP(hipStreamSynchronize(stream), "offset readback");
        }
        tape.bytes = last;
    }
    else if (device_offsets) { // u32 offsets on the device
        SZS_TRY_HIP(tape.offsets_copy.reserve((count + 1) * 8), "Failed to allocate tape offsets");
        SZS_TRY_HIP(szs_amd_launch_widen_offsets(reinterpret_cast<uint32_t const *>(offsets),
                                                 tape.offsets_copy.as<uint64_t>(), count + 1, stream),
                    "offset widening");
        tape.offsets = tape.offsets_copy.as<uint64_t>();
        uint64_t last = bytes_hint ? bytes_hint : unknown_bytes_k;
        if (!bytes_hint && need_bytes) {
            uint32_t narrow = 0;
            SZS_TRY_HIP(hipMemcpyAsync(&narrow, offsets + count, 4, hipMemcpyDeviceToHost, stream), "offset readback");
            SZS_TRY_HIP(hipStreamSynchronize(stream), "offset readback");
            last = narrow;
        }
        tape.bytes = last;
    }
    else {
        // Widened into the tape's own host vector, which outlives the copy (the tape is reused only after the
        // batch's work has completed).
        std::vector<uint64_t> &wide = tape.host_offsets;
        wide.resize(count + 1);
        for (uint64_t i = 0; i <= count; ++i) wide[i] = (uint64_t)offsets[i]; // checked by `check_host_offsets`
        SZS_TRY_HIP(tape.offsets_copy.reserve((count + 1) * 8), "Failed to allocate tape offsets");
        SZS_TRY_HIP(hipMemcpyAsync(tape.offsets_copy.data, wide.data(), (count + 1) * 8, hipMemcpyHostToDevice,
                                   stream), "tape offsets upload");
        tape.offsets = tape.offsets_copy.as<uint64_t>();
        tape.bytes = wide[count];
    }
    // Data.
    if (tape.bytes == 0 || device_data) tape.data = reinterpret_cast<uint8_t const *>(data);
    else {
        SZS_TRY_HIP(tape.data_copy.reserve(tape.bytes + 16), "Failed to allocate the tape in HBM");
        SZS_TRY_HIP(hipMemcpyAsync(tape.data_copy.data, data, tape.bytes, hipMemcpyHostToDevice, stream),
                    "tape upload");
        tape.data = tape.data_copy.as<uint8_t>();
    }
    if (!tape.data) { // empty strings only: any valid device address works, nothing is read
        SZS_TRY_HIP(tape.data_copy.reserve(16), "Failed to allocate the tape in HBM");
        tape.data = tape.data_copy.as<uint8_t>();
    }
    if (runes) {
        SZS_TRY_HIP(tape.runes_buffer.reserve(std::max<uint64_t>(tape.bytes, 1) * 4), "Failed to allocate runes");
        SZS_TRY_HIP(tape.rune_counts_buffer.reserve(std::max<uint64_t>(count, 1) * 4), "Failed to allocate runes");
        tape.runes = tape.runes_buffer.as<uint32_t>();
        tape.rune_counts = tape.rune_counts_buffer.as<uint32_t>();
    }
    return ok(error_message);
}

template <typename offset_type_>
static bool offsets_never_decrease(offset_type_ const *offsets, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i)
        if (offsets[i + 1] < offsets[i]) return false;
    return true;
}

sz_status_t check_host_offsets(segment_input_t const &input, char const **error_message) {
    auto check = [&](auto const *tape) {
        if (!tape || !tape->count || !tape->offsets || is_device_pointer(tape->offsets)) return true;
        return offsets_never_decrease(tape->offsets, tape->count);
    };
    if (check(input.queries) && check(input.queries32) && check(input.candidates) && check(input.candidates32))
        return ok(error_message);
    return fail(sz_unexpected_dimensions_k, error_message, "Tape offsets must not decrease");
}

sz_status_t stage_bytes_tape(sz_sequence_u64tape_t const *input, staged_tape_t &tape, hipStream_t stream,
                             char const **error_message) {
    return stage_tape(input->data, input->offsets, input->count, false, 0, tape, stream, error_message);
}

sz_status_t stage_bytes_tape(sz_sequence_u32tape_t const *input, staged_tape_t &tape, hipStream_t stream,
                             char const **error_message) {
    return stage_tape(input->data, input->offsets, input->count, false, 0, tape, stream, error_message);
}

static sz_status_t stage_input(segment_input_t const &input, bool queries, bool runes, staged_tape_t &tape,
                               hipStream_t stream, char const **error_message) {
    if (queries) {
        if (input.queries)
            return stage_tape(input.queries->data, input.queries->offsets, input.queries->count, runes,
                              input.queries_bytes, tape, stream, error_message);
        return stage_tape(input.queries32->data, input.queries32->offsets, input.queries32->count, runes,
                          input.queries_bytes, tape, stream, error_message);
    }
    if (input.candidates)
        return stage_tape(input.candidates->data, input.candidates->offsets, input.candidates->count, runes,
                          input.candidates_bytes, tape, stream, error_message);
    return stage_tape(input.candidates32->data, input.candidates32->offsets, input.candidates32->count, runes,
                      input.candidates_bytes, tape, stream, error_message);
}

#pragma endregion

void batch_t::reset() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr, graph = nullptr;
    for (auto *tape : tapes) {
        tape->reset_view();
        spare_tapes.push_back(tape);
    }
    tapes.clear();
    for (auto *buffer : result_buffers) spare_results.push_back(buffer);
    result_buffers.clear();
    inputs.clear(), host_results.clear(), results_elements.clear(), result_rows.clear(), result_columns.clear();
    device_results.clear(), materialize_args.clear(), overflowed.clear(), pinned_offsets.clear();
    per_pair_tasks.clear(), per_pair_variants.clear();
    chained_pairs = 0, chain_pitch = 0, decode_tapes_count = 0, decode_strings = 0, wide_pairs = 0, wide_pitch = 0;
    pinned_results = nullptr, pinned_bytes = 0, tasks_count = 0, cells = 0, core_ops = 0;
    scratch_capacity = 0, waves = 0, lds_bytes = 0, last_kernel_ms = 0, materialized = false;
    params_table = nullptr, segment_table = nullptr, materialize_table = nullptr, decode_table = nullptr;
    shape_table = nullptr;
}

size_t batch_t::device_bytes() const {
    size_t total = 0;
    for (device_buffer_t const *buffer : {&tables, &tasks, &bin_of_task, &order, &shape_of_task, &identity, &histogram,
                                          &cursor, &task_start, &item_start, &priority, &position_of_bin, &queue_head,
                                          &scan_blocks, &scratch, &stats, &segment_overflow, &results_block,
                                          &chain_rows, &chain_progress, &wide_rows})
        total += buffer->capacity;
    for (std::vector<staged_tape_t *> const *list : {&tapes, &spare_tapes})
        for (staged_tape_t const *tape : *list)
            total += tape->data_copy.capacity + tape->offsets_copy.capacity + tape->runes_buffer.capacity +
                     tape->rune_counts_buffer.capacity;
    for (std::vector<device_buffer_t *> const *list : {&result_buffers, &spare_results})
        for (device_buffer_t const *buffer : *list) total += buffer->capacity;
    return total;
}

void batch_t::release_device_memory() {
    reset();
    for (device_buffer_t *buffer : {&tables, &tasks, &bin_of_task, &order, &shape_of_task, &identity, &histogram,
                                    &cursor, &task_start, &item_start, &priority, &position_of_bin, &queue_head,
                                    &scan_blocks,
                                    &scratch, &stats, &segment_overflow, &results_block, &chain_rows,
                                    &chain_progress, &wide_rows})
        buffer->release();
    for (staged_tape_t *tape : spare_tapes)
        for (device_buffer_t *buffer : {&tape->data_copy, &tape->offsets_copy, &tape->runes_buffer,
                                        &tape->rune_counts_buffer})
            buffer->release();
    for (device_buffer_t *buffer : spare_results) buffer->release();
    static_tables_ready = false;
}

batch_t::~batch_t() {
    reset();
    if (kernel_begin) (void)hipEventDestroy(kernel_begin);
    if (kernel_end) (void)hipEventDestroy(kernel_end);
    for (auto *tape : spare_tapes) delete tape;
    for (auto *buffer : spare_results) delete buffer;
}

/** A staged tape of the batch: an earlier call's (buffers kept) when there is one, else a new one. */
static staged_tape_t *next_tape(batch_t &b) {
    staged_tape_t *tape = nullptr;
    if (!b.spare_tapes.empty()) tape = b.spare_tapes.back(), b.spare_tapes.pop_back();
    else tape = new (std::nothrow) staged_tape_t();
    if (tape) b.tapes.push_back(tape);
    return tape;
}

static device_buffer_t *next_result_buffer(batch_t &b) {
    device_buffer_t *buffer = nullptr;
    if (!b.spare_results.empty()) buffer = b.spare_results.back(), b.spare_results.pop_back();
    else buffer = new (std::nothrow) device_buffer_t();
    if (buffer) b.result_buffers.push_back(buffer);
    return buffer;
}

static bins_view_t bins_view(batch_t &b) {
    bins_view_t bins;
    bins.histogram = b.histogram.as<uint32_t>();
    bins.cursor = b.cursor.as<uint32_t>();
    bins.task_start = b.task_start.as<uint32_t>();
    bins.item_start = b.item_start.as<uint32_t>();
    bins.priority = b.priority.as<uint32_t>();
    bins.order = b.order.as<uint32_t>();
    bins.queue_head = b.queue_head.as<uint32_t>();
    bins.shape_of_task = b.shape_of_task.as<uint32_t>();
    bins.scan_blocks = b.scan_blocks.as<uint32_t>();
    bins.wave_bins_from = wave_bins_from();
    return bins;
}

/** Device work up to (excluding) scoring: clear counters, transcode UTF-8, materialize, bin. With `reuse`, the
 *  runes, tasks and bin histogram that the staging's own pass left on the device are used as they are. */
static sz_status_t enqueue_front(batch_t &b, bool bin, char const **error_message, bool reuse = false) {
    hipStream_t const stream = b.stream;
    uint32_t const segments_count = (uint32_t)b.inputs.size();
    // Every counter the run starts from zero, in one launch: bin cursors, the work queue, chained pairs' progress
    // words, and (when materializing) the histogram, the stats and the segment flags.
    fill_list_t fills {};
    auto fill = [&](device_buffer_t &buffer, uint32_t count) {
        fills.data[fills.entries] = buffer.as<uint32_t>(), fills.count[fills.entries] = count;
        fills.value[fills.entries++] = 0;
    };
    fill(b.cursor, bins_count_k);
    fill(b.queue_head, 1);
    if (b.chained_pairs) fill(b.chain_progress, b.chained_pairs * (uint32_t)chain_groups_k);
    if (!reuse) {
        fill(b.histogram, bins_count_k);
        fill(b.stats, sizeof(prepare_stats_t) / 4);
        fill(b.segment_overflow, segments_count);
    }
    SZS_TRY_HIP(szs_amd_launch_fill_many(&fills, stream), "fill");
    if (!reuse) {
        SZS_TRY_HIP(szs_amd_launch_utf8_decode(b.decode_table, b.decode_tapes_count,
                                               b.decode_strings, b.stats.as<prepare_stats_t>(),
                                               b.segment_overflow.as<uint32_t>(), stream),
                    "UTF-8 decoding");
        SZS_TRY_HIP(szs_amd_launch_materialize(b.materialize_table, segments_count,
                                               b.tasks_count, b.shape_table, b.tasks.as<task_t>(),
                                               b.bin_of_task.as<uint32_t>(), b.histogram.as<uint32_t>(),
                                               b.segment_table, b.stats.as<prepare_stats_t>(),
                                               b.segment_overflow.as<uint32_t>(), stream),
                    "materialization");
    }
    if (bin) {
        bins_view_t const bins = bins_view(b);
        SZS_TRY_HIP(szs_amd_launch_scan_bins(&bins, stream), "bin scan");
        SZS_TRY_HIP(szs_amd_launch_scatter_bins(&bins, b.bin_of_task.as<uint32_t>(),
                                                b.position_of_bin.as<uint32_t>(), (uint32_t)b.tasks_count, stream),
                    "bin scatter");
    }
    return ok(error_message);
}

static chain_view_t chain_view(batch_t &b) {
    chain_view_t chain;
    chain.rows = b.chain_rows.as<int32_t>();
    chain.progress = b.chain_progress.as<uint32_t>();
    chain.pitch = b.chain_pitch;
    chain.timeout = &b.stats.as<prepare_stats_t>()->chain_timeout;
    chain.wide_rows = b.wide_rows.as<int64_t>();
    chain.wide_pitch = b.wide_pitch;
    return chain;
}

/** Scoring launches; HIP events bracket them unless they are being captured into a graph. */
static sz_status_t enqueue_scoring(batch_t &b, bool timed, char const **error_message) {
    hipStream_t const stream = b.stream;
    chain_view_t const chain = chain_view(b); // progress words were cleared with the front's counters
    if (timed) SZS_TRY_HIP(hipEventRecord(b.kernel_begin, stream), "event");
    if (b.mode == szs_batch_per_pair_k) {
        for (size_t i = 0; i < b.per_pair_tasks.size(); ++i)
            SZS_TRY_HIP(szs_amd_launch_single(b.tasks.as<task_t>(), b.segment_table,
                                              b.identity.as<uint32_t>(), b.per_pair_tasks[i], b.per_pair_variants[i],
                                              b.scratch.as<int32_t>(), b.scratch_capacity, b.lds_bytes, &chain,
                                              stream),
                        "per-pair launch");
    }
    else {
        bins_view_t const bins = bins_view(b);
        SZS_TRY_HIP(szs_amd_launch_persistent(b.tasks.as<task_t>(), b.segment_table, &bins,
                                              b.scratch.as<int32_t>(), b.scratch_capacity, b.waves, b.lds_bytes,
                                              &chain, stream),
                    "persistent scoring launch");
    }
    if (timed) SZS_TRY_HIP(hipEventRecord(b.kernel_end, stream), "event");
    return ok(error_message);
}

static sz_status_t enqueue_results(batch_t &b, char const **error_message) {
    for (size_t s = 0; s < b.inputs.size(); ++s) {
        if (!b.host_results[s] || !b.results_elements[s]) continue;
        if (b.pinned_offsets[s] != SIZE_MAX) continue; // copied together below
        if (b.inputs[s].mode == segment_pairs_k)
            SZS_TRY_HIP(hipMemcpyAsync(b.host_results[s], b.device_results[s], b.results_elements[s] * 8,
                                       hipMemcpyDeviceToHost, b.stream),
                        "results download");
        else
            SZS_TRY_HIP(hipMemcpy2DAsync(b.host_results[s], b.inputs[s].results_row_stride * 8, b.device_results[s],
                                         b.inputs[s].results_row_stride * 8, b.result_columns[s] * 8,
                                         b.result_rows[s], hipMemcpyDeviceToHost, b.stream),
                        "results download");
    }
    if (b.pinned_results && b.results_block.capacity) // every staged segment sits in one device block
        SZS_TRY_HIP(hipMemcpyAsync(b.pinned_results, b.results_block.data, b.pinned_bytes, hipMemcpyDeviceToHost,
                                   b.stream),
                    "results download");
    return ok(error_message);
}

sz_status_t batch_prepare(device_scope_t *scope, std::vector<segment_input_t> const &inputs, szs_batch_mode_t mode,
                          batch_t **out, char const **error_message, batch_options_t const &options) {
    batch_t *batch = new (std::nothrow) batch_t();
    if (!batch) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a batch");
    sz_status_t const status = batch_stage(*batch, scope, inputs, mode, error_message, options);
    if (status != sz_success_k) {
        delete batch;
        return status;
    }
    *out = batch;
    return ok(error_message);
}

sz_status_t batch_stage(batch_t &b, device_scope_t *scope, std::vector<segment_input_t> const &inputs,
                        szs_batch_mode_t mode, char const **error_message, batch_options_t const &options) {
    b.reset();
    // Host offsets are checked before any device work: a decreasing offset would send the sweeps outside the tape
    // (offsets in HBM are checked by the materialization kernel instead).
    for (segment_input_t const &input : inputs) {
        sz_status_t const checked = check_host_offsets(input, error_message);
        if (checked != sz_success_k) return checked;
    }
    int device = 0;
    hipStream_t stream = nullptr;
    sz_status_t status = resolve_gpu(scope, &device, &stream, error_message);
    if (status != sz_success_k) return status;
    if (options.stream) stream = options.stream;
    SZS_TRY_HIP(order_after_default_stream(stream), "Ordering after the default stream");
    b.device = device;
    b.stream = stream;
    b.mode = mode;
    b.inputs = inputs;
    size_t const segments_count = inputs.size();
    b.host_results.assign(segments_count, nullptr);
    b.results_elements.assign(segments_count, 0);
    b.result_rows.assign(segments_count, 0);
    b.result_columns.assign(segments_count, 0);
    b.device_results.assign(segments_count, nullptr);
    b.pinned_offsets.assign(segments_count, SIZE_MAX);
    b.materialize_args.resize(sizeof(materialize_args_t) * std::max<size_t>(segments_count, 1));
    auto *args = reinterpret_cast<materialize_args_t *>(b.materialize_args.data());

    // Pair-list results headed for host memory are staged together: one device block, one copy into the
    // caller's page-locked staging, then a host copy per segment when the batch is joined.
    size_t staged_bytes = 0;
    if (options.pinned_results)
        for (size_t s = 0; s < segments_count; ++s) {
            segment_input_t const &input = inputs[s];
            uint64_t const count = input.queries ? input.queries->count : input.queries32 ? input.queries32->count : 0;
            if (input.mode != segment_pairs_k || !count || !input.results || is_device_pointer(input.results))
                continue;
            if (staged_bytes + count * 8 > options.pinned_capacity) break;
            b.pinned_offsets[s] = staged_bytes;
            staged_bytes += count * 8;
        }
    if (staged_bytes) {
        SZS_TRY_HIP(b.results_block.reserve(staged_bytes), "Failed to allocate results in HBM");
        b.pinned_results = options.pinned_results;
        b.pinned_bytes = staged_bytes;
    }

    std::vector<engine_params_t> params(segments_count);
    std::vector<segment_view_t> views(segments_count);
    uint64_t tasks_total = 0;
    uint32_t lds_bytes = 0;
    // The ladders this batch is binned with: one copy, so that the LDS profile sizes below and the device's shape
    // choice agree even if a caller installs another ladder meanwhile.
    shape_table_t const table = [] {
        std::lock_guard<std::mutex> lock(szs_amd::tuning_mutex());
        return active_shape_table();
    }();
    // Profile bytes per lane and class a slot's ladder needs: its widest DP rows per lane in whole dwords.
    auto profile_row_bytes = [&table](int slot) {
        int rows = 8;
        auto consider = [&](int shape) {
            if (!shapes_k[shape].wide && !shapes_k[shape].bit_parallel) rows = std::max(rows, shapes_k[shape].rows_per_lane);
        };
        for (int i = 0; i < table.used[slot]; ++i) consider(table.order[slot][i]);
        consider(table.multi_pass_shape[slot]);
        return (uint32_t)((rows + 3) / 4 * 4);
    };
    for (size_t s = 0; s < segments_count; ++s) {
        segment_input_t const &input = inputs[s];
        engine_t const *engine = input.engine;
        bool const runes = engine->kind == szs_engine_levenshtein_utf8_k;
        staged_tape_t *queries = next_tape(b);
        if (!queries) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a tape");
        queries->segment = (uint32_t)s;
        status = stage_input(input, true, runes, *queries, stream, error_message);
        if (status != sz_success_k) return status;
        staged_tape_t *candidates = queries;
        if (input.mode != segment_symmetric_k) {
            candidates = next_tape(b);
            if (!candidates) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a tape");
            candidates->segment = (uint32_t)s;
            status = stage_input(input, false, runes, *candidates, stream, error_message);
            if (status != sz_success_k) return status;
        }
        uint64_t const q = queries->count, c = candidates->count;
        uint64_t tasks = 0, elements = 0;
        if (input.mode == segment_pairs_k) {
            if (q != c) return fail(sz_unexpected_dimensions_k, error_message, "Pair tapes must hold equal counts");
            tasks = q, elements = q;
            b.result_rows[s] = q, b.result_columns[s] = 1;
        }
        else {
            if (input.results_row_stride < c)
                return fail(sz_unexpected_dimensions_k, error_message, "Result row stride is below the column count");
            tasks = input.mode == segment_cross_k ? q * c : q * (q + 1) / 2;
            elements = q && c ? (q - 1) * input.results_row_stride + c : 0;
            b.result_rows[s] = q, b.result_columns[s] = c;
        }
        b.results_elements[s] = elements;
        if (!input.results && elements) return fail(sz_unexpected_dimensions_k, error_message, "Results are null");
        if (b.pinned_offsets[s] != SIZE_MAX) {
            b.host_results[s] = input.results;
            b.device_results[s] = reinterpret_cast<int64_t *>(b.results_block.as<char>() + b.pinned_offsets[s]);
        }
        else if (elements && !is_device_pointer(input.results)) {
            device_buffer_t *buffer = next_result_buffer(b);
            if (!buffer) return fail(sz_bad_alloc_k, error_message, "Failed to allocate results");
            SZS_TRY_HIP(buffer->reserve(elements * 8), "Failed to allocate results in HBM");
            b.host_results[s] = input.results;
            b.device_results[s] = buffer->as<int64_t>();
            if (input.mode != segment_pairs_k) // cells outside the matrix stay as the caller left them
                SZS_TRY_HIP(hipMemcpy2DAsync(buffer->data, input.results_row_stride * 8, input.results,
                                             input.results_row_stride * 8, c * 8, q, hipMemcpyHostToDevice, stream),
                            "results upload");
        }
        else b.device_results[s] = reinterpret_cast<int64_t *>(input.results);

        params[s] = engine->params;
        views[s].params = nullptr; // patched after upload
        views[s].query_symbols = runes ? (void const *)queries->runes : (void const *)queries->data;
        views[s].candidate_symbols = runes ? (void const *)candidates->runes : (void const *)candidates->data;
        views[s].results = b.device_results[s];
        views[s].symbols = runes ? symbols_u32_k : symbols_u8_k;
        views[s].family = engine->params.family;

        materialize_args_t &a = args[s];
        std::memset(&a, 0, sizeof(a));
        a.query_offsets = queries->offsets;
        a.candidate_offsets = input.mode == segment_symmetric_k ? nullptr : candidates->offsets;
        a.query_rune_counts = queries->rune_counts;
        a.candidate_rune_counts = input.mode == segment_symmetric_k ? nullptr : candidates->rune_counts;
        a.queries_count = q;
        a.candidates_count = c;
        a.query_bytes = queries->bytes;
        a.candidate_bytes = candidates->bytes;
        a.results_row_stride = input.results_row_stride;
        a.tasks_count = tasks;
        a.task_base = (uint32_t)tasks_total;
        a.segment = (uint32_t)s;
        a.family = engine->params.family;
        a.runes = runes;
        a.symbols = views[s].symbols;
        bool const local_family = engine->params.family == family_sw_linear_k ||
                                  engine->params.family == family_sw_affine_k;
        int32_t const fast_bias = local_family ? engine->params.open : engine->params.open + engine->params.extend;
        if (engine->params.family >= family_nw_linear_k && engine->params.shift != fast_bias)
            a.symbols = class_unshifted_k;
        a.mode = input.mode;
        a.open = engine->params.open;
        a.extend = engine->params.extend;
        a.unit_costs = engine->params.match == 0 && engine->params.mismatch == 1 && engine->params.open == 1 &&
                       engine->params.extend == 1;
        a.cell_growth = std::max(1, engine->substitution_magnitude + 6 * engine->gap_magnitude);
        a.packed_rows = a.symbols == symbols_u8_k ? engine->packed_rows : 0;
        a.latency_stripes = options.latency_stripes;
        tasks_total += tasks;
        if (tasks_total >= 0xFFFFFFF0ull)
            return fail(sz_overflow_risk_k, error_message, "A batch holds at most 2^32 - 16 scored cells");
        if (engine->params.family >= family_nw_linear_k) { // per-wave profile: classes x 64 lanes x rows per lane
            int const slot = engine->params.family * 2 + (a.symbols == class_unshifted_k ? 1 : 0);
            lds_bytes = std::max<uint32_t>(lds_bytes, (uint32_t)engine->params.dense_count * 64u * profile_row_bytes(slot));
        }
        if (a.unit_costs && !runes && engine->params.family == family_lev_linear_k) // Myers tables + match table
            lds_bytes = std::max<uint32_t>(lds_bytes, myers_lds_bytes_k - inject_ring_bytes_k);
    }
    b.tasks_count = tasks_total;
    // Dynamic LDS per wave: the NW/SW profile or the Myers symbol tables, whichever is larger, then the ring.
    lds_bytes = std::max<uint32_t>(lds_bytes, 512u);
    b.lds_bytes = ((lds_bytes + 15u) & ~15u) + inject_ring_bytes_k;

    // Every UTF-8 tape is transcoded by one launch.
    std::vector<utf8_tape_t> decode;
    for (auto *tape : b.tapes)
        if (tape && tape->runes && tape->count) {
            utf8_tape_t entry;
            entry.bytes = tape->data;
            entry.offsets = tape->offsets;
            entry.runes = tape->runes;
            entry.rune_counts = tape->rune_counts;
            entry.string_base = b.decode_strings;
            entry.capacity = std::max<uint64_t>(tape->bytes, 1);
            entry.segment = tape->segment;
            entry.reserved = 0;
            decode.push_back(entry);
            b.decode_strings += tape->count;
        }
    b.decode_tapes_count = (uint32_t)decode.size();
    // The per-call device tables (engine parameters, segment views, materialization arguments, UTF-8 tapes, the
    // shape table) go up in one copy from a host image the batch keeps until it is staged again.
    size_t const table_count = std::max<size_t>(segments_count, 1);
    auto aligned = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    size_t const at_params = 0;
    size_t const at_views = aligned(at_params + sizeof(engine_params_t) * table_count);
    size_t const at_args = aligned(at_views + sizeof(segment_view_t) * table_count);
    size_t const at_decode = aligned(at_args + sizeof(materialize_args_t) * table_count);
    size_t const at_shapes = aligned(at_decode + sizeof(utf8_tape_t) * std::max<size_t>(decode.size(), 1));
    size_t const table_bytes = at_shapes + sizeof(shape_table_t);
    SZS_TRY_HIP(b.tables.reserve(table_bytes), "allocation");
    char *device_tables = b.tables.as<char>();
    b.params_table = reinterpret_cast<engine_params_t *>(device_tables + at_params);
    b.segment_table = reinterpret_cast<segment_view_t *>(device_tables + at_views);
    b.materialize_table = reinterpret_cast<materialize_args_t *>(device_tables + at_args);
    b.decode_table = reinterpret_cast<utf8_tape_t *>(device_tables + at_decode);
    b.shape_table = reinterpret_cast<shape_table_t *>(device_tables + at_shapes);
    for (size_t s = 0; s < segments_count; ++s) views[s].params = b.params_table + s;
    b.table_image.assign(table_bytes, 0);
    char *image = b.table_image.data();
    std::memcpy(image + at_params, params.data(), sizeof(engine_params_t) * segments_count);
    std::memcpy(image + at_views, views.data(), sizeof(segment_view_t) * segments_count);
    std::memcpy(image + at_args, args, sizeof(materialize_args_t) * segments_count);
    if (!decode.empty()) std::memcpy(image + at_decode, decode.data(), sizeof(utf8_tape_t) * decode.size());
    std::memcpy(image + at_shapes, &table, sizeof(shape_table_t));
    SZS_TRY_HIP(hipMemcpyAsync(device_tables, image, table_bytes, hipMemcpyHostToDevice, stream), "upload");
    if (!b.static_tables_ready) { // constant tables: once per batch object
        std::vector<uint32_t> const &priority = bin_priority();
        static std::vector<uint32_t> const position = [&] {
            std::vector<uint32_t> inverse(bins_count_k);
            for (uint32_t p = 0; p < bins_count_k; ++p) inverse[priority[p]] = p;
            return inverse;
        }();
        SZS_TRY_HIP(b.priority.reserve(bins_count_k * 4), "allocation");
        SZS_TRY_HIP(hipMemcpyAsync(b.priority.data, priority.data(), bins_count_k * 4, hipMemcpyHostToDevice, stream),
                    "upload");
        SZS_TRY_HIP(b.position_of_bin.reserve(bins_count_k * 4), "allocation");
        SZS_TRY_HIP(hipMemcpyAsync(b.position_of_bin.data, position.data(), bins_count_k * 4, hipMemcpyHostToDevice,
                                   stream), "upload");
        b.static_tables_ready = true;
    }
    uint64_t const slots = std::max<uint64_t>(tasks_total, 1);
    SZS_TRY_HIP(b.tasks.reserve(slots * sizeof(task_t)), "Failed to allocate tasks");
    SZS_TRY_HIP(b.bin_of_task.reserve(slots * 4), "Failed to allocate tasks");
    SZS_TRY_HIP(b.order.reserve(slots * 4), "Failed to allocate tasks");
    SZS_TRY_HIP(b.shape_of_task.reserve(slots * 4), "Failed to allocate tasks");
    SZS_TRY_HIP(b.histogram.reserve(bins_count_k * 4), "allocation");
    SZS_TRY_HIP(b.cursor.reserve(bins_count_k * 4), "allocation");
    SZS_TRY_HIP(b.task_start.reserve(bins_count_k * 4), "allocation");
    SZS_TRY_HIP(b.item_start.reserve((bins_count_k + 1) * 4), "allocation");
    SZS_TRY_HIP(b.queue_head.reserve(64), "allocation");
    SZS_TRY_HIP(b.scan_blocks.reserve(scan_blocks_k * 2 * 4), "allocation");
    SZS_TRY_HIP(b.stats.reserve(sizeof(prepare_stats_t)), "allocation");
    SZS_TRY_HIP(b.segment_overflow.reserve(table_count * 4), "allocation");
    if (!b.kernel_begin) SZS_TRY_HIP(hipEventCreate(&b.kernel_begin), "event");
    if (!b.kernel_end) SZS_TRY_HIP(hipEventCreate(&b.kernel_end), "event");

    // Materialize once: the stats size the scratch rows, carry the reach check and count the cells.
    status = enqueue_front(b, false, error_message);
    if (status != sz_success_k) return status;
    prepare_stats_t stats;
    std::memset(&stats, 0, sizeof(stats));
    b.overflowed.assign(segments_count, 0);
    SZS_TRY_HIP(hipMemcpyAsync(&stats, b.stats.data, sizeof(stats), hipMemcpyDeviceToHost, stream), "readback");
    std::vector<task_t> tasks;
    std::vector<uint32_t> bins;
    if (mode == szs_batch_per_pair_k && tasks_total) { // the per-pair baseline launches task by task
        tasks.resize(tasks_total);
        bins.resize(tasks_total);
        SZS_TRY_HIP(hipMemcpyAsync(tasks.data(), b.tasks.data, tasks_total * sizeof(task_t), hipMemcpyDeviceToHost,
                                   stream), "task readback");
        SZS_TRY_HIP(hipMemcpyAsync(bins.data(), b.bin_of_task.data, tasks_total * 4, hipMemcpyDeviceToHost, stream),
                    "task readback");
    }
    SZS_TRY_HIP(hipStreamSynchronize(stream), "materialization");
    if (stats.overflow || stats.bad_tape) {
        if (!options.defer_overflow && stats.bad_tape)
            return fail(sz_unexpected_dimensions_k, error_message,
                        "A tape's offsets decrease or run past its byte count");
        if (!options.defer_overflow)
            return fail(sz_overflow_risk_k, error_message, "A string holds 2^32 or more symbols");
        if (segments_count) {
            SZS_TRY_HIP(hipMemcpyAsync(b.overflowed.data(), b.segment_overflow.data, segments_count * 4,
                                       hipMemcpyDeviceToHost, stream), "readback");
            SZS_TRY_HIP(hipStreamSynchronize(stream), "readback");
        }
    }
    b.cells = stats.cells;
    b.core_ops = stats.core_ops_64 / 64;
    for (uint64_t t = 0; t < tasks.size(); ++t) {
        if (bins[t] == ~0u) continue;
        b.per_pair_tasks.push_back((uint32_t)t);
        b.per_pair_variants.push_back((int)(bins[t] / column_buckets_k));
    }
    uint64_t const widest_multi_pass = stats.widest_multi_pass;
    b.scratch_capacity = (uint32_t)(widest_multi_pass + 128);
    int compute_units = 0, blocks_per_cu = 0;
    SZS_TRY_HIP(hipDeviceGetAttribute(&compute_units, hipDeviceAttributeMultiprocessorCount, device), "attributes");
    SZS_TRY_HIP(szs_amd_persistent_occupancy(&blocks_per_cu, b.lds_bytes), "occupancy");
    blocks_per_cu = std::max(1, std::min(blocks_per_cu, 8));
    b.waves = mode == szs_batch_per_pair_k ? 1u : (uint32_t)(compute_units * blocks_per_cu); // workgroups
    // A batch of few pairs needs few waves: a one-wave item holds at least one pair, and the queue hands out every
    // item in order whatever the grid (a chained group only waits for groups drawn before it), so such a batch
    // launches one workgroup per 4 pairs instead of filling the chip: an idle workgroup still starts, draws once
    // from the queue's single counter and exits, and those draws serialize.
    b.waves = (uint32_t)std::min<uint64_t>(b.waves, std::max<uint64_t>(1, (tasks_total + waves_per_group_k - 1) /
                                                                            waves_per_group_k));
    // Every resident workgroup holds stripe-boundary rows as wide as the widest multi-stripe pair: a pair of
    // millions of columns would ask for tens of GB. Past a limit (the caller's, else 4 GiB or a quarter of the free
    // HBM) the persistent launch runs fewer workgroups instead; the work queue does not depend on their number.
    uint64_t const group_scratch_bytes = waves_per_group_k * 2ull * b.scratch_capacity * 4;
    uint64_t scratch_limit = options.max_scratch_bytes;
    if (!scratch_limit && widest_multi_pass && b.waves * group_scratch_bytes > (256ull << 20)) {
        size_t free_bytes = 0, total_bytes = 0;
        scratch_limit = 4ull << 30;
        if (hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess)
            scratch_limit = std::min<uint64_t>(scratch_limit, std::max<uint64_t>(free_bytes / 4, group_scratch_bytes));
        else (void)hipGetLastError();
    }
    if (widest_multi_pass && scratch_limit && b.waves * group_scratch_bytes > scratch_limit)
        b.waves = (uint32_t)std::max<uint64_t>(1, scratch_limit / group_scratch_bytes);
    uint64_t const scratch_groups = widest_multi_pass ? b.waves : 1;
    SZS_TRY_HIP(b.scratch.reserve(scratch_groups * waves_per_group_k * 2ull * b.scratch_capacity * 4 + 256),
                "Failed to allocate scratch");
    // Chained pairs: one bottom row (scores and gaps) between consecutive groups, and a progress word per group.
    b.chained_pairs = stats.chained_pairs;
    b.chain_pitch = (stats.chained_columns + 1 + 63) / 64 * 64;
    if (b.chained_pairs) {
        SZS_TRY_HIP(b.chain_rows.reserve((size_t)b.chained_pairs * (chain_groups_k - 1) * 2 * b.chain_pitch * 4),
                    "Failed to allocate the chain rows");
        SZS_TRY_HIP(b.chain_progress.reserve((size_t)b.chained_pairs * chain_groups_k * 4),
                    "Failed to allocate the chain progress words");
    }
    // Pairs of the int64 shape: each keeps its stripes' bottom row and vertical gaps in its own rows.
    b.wide_pairs = stats.wide_pairs;
    b.wide_pitch = ((uint64_t)stats.wide_columns + 1 + 63) / 64 * 64;
    if (b.wide_pairs)
        SZS_TRY_HIP(b.wide_rows.reserve((size_t)b.wide_pairs * 2 * b.wide_pitch * 8), "Failed to allocate the int64 rows");
    if (mode == szs_batch_per_pair_k) {
        std::vector<uint32_t> identity(slots);
        std::iota(identity.begin(), identity.end(), 0u);
        SZS_TRY_HIP(b.identity.reserve(slots * 4), "allocation");
        SZS_TRY_HIP(hipMemcpy(b.identity.data, identity.data(), slots * 4, hipMemcpyHostToDevice), "upload");
    }
    if (mode == szs_batch_graph_k) {
        SZS_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
        sz_status_t front = enqueue_front(b, true, error_message);
        sz_status_t scoring = front == sz_success_k ? enqueue_scoring(b, false, error_message) : front;
        hipGraph_t graph = nullptr;
        hipError_t const ended = hipStreamEndCapture(stream, &graph);
        if (scoring != sz_success_k) {
            if (graph) (void)hipGraphDestroy(graph);
            return scoring;
        }
        SZS_TRY_HIP(ended, "graph capture");
        b.graph = graph;
        SZS_TRY_HIP(hipGraphInstantiate(&b.graph_exec, graph, nullptr, nullptr, 0), "graph instantiation");
    }
    b.materialized = mode != szs_batch_graph_k;
    return ok(error_message);
}

sz_status_t batch_launch(batch_t *batch, char const **error_message, bool ordered) {
    batch_t &b = *batch;
    SZS_TRY_HIP(hipSetDevice(b.device), "hipSetDevice failed");
    if (!ordered) SZS_TRY_HIP(order_after_default_stream(b.stream), "Ordering after the default stream");
    if (b.mode == szs_batch_graph_k) { // the events then time the whole replay
        SZS_TRY_HIP(hipEventRecord(b.kernel_begin, b.stream), "event");
        SZS_TRY_HIP(hipGraphLaunch(b.graph_exec, b.stream), "graph launch");
        SZS_TRY_HIP(hipEventRecord(b.kernel_end, b.stream), "event");
    }
    else {
        // The first launch after staging scores what the staging pass materialized; later ones redo it all.
        sz_status_t status = enqueue_front(b, b.mode != szs_batch_per_pair_k, error_message, b.materialized);
        b.materialized = false;
        if (status != sz_success_k) return status;
        status = enqueue_scoring(b, true, error_message);
        if (status != sz_success_k) return status;
    }
    return enqueue_results(b, error_message);
}

sz_status_t batch_wait(batch_t *batch, char const **error_message) {
    batch_t &b = *batch;
    SZS_TRY_HIP(hipSetDevice(b.device), "hipSetDevice failed");
    SZS_TRY_HIP(hipStreamSynchronize(b.stream), "scoring failed on the device");
    if (b.chained_pairs) { // a chained group that gave up waiting leaves its pair's score undefined
        uint32_t timeout = 0;
        SZS_TRY_HIP(hipMemcpy(&timeout, &b.stats.as<prepare_stats_t>()->chain_timeout, 4, hipMemcpyDeviceToHost),
                    "readback");
        if (timeout)
            return fail(sz_status_unknown_k, error_message,
                        "A chained pair's workgroup stopped waiting for its predecessor; its score is undefined");
    }
    for (size_t s = 0; s < b.inputs.size(); ++s)
        if (b.pinned_offsets[s] != SIZE_MAX)
            std::memcpy(b.host_results[s], b.pinned_results + b.pinned_offsets[s], b.results_elements[s] * 8);
    float ms = 0;
    if (hipEventElapsedTime(&ms, b.kernel_begin, b.kernel_end) == hipSuccess) b.last_kernel_ms = ms;
    else (void)hipGetLastError();
    return ok(error_message);
}

} // namespace szs_amd

#pragma region Tuning hooks


extern "C" {

int szs_tuning_slots(void) { return szs_amd::shape_slots_k; }
int szs_tuning_shapes(void) { return szs_amd::shapes_count_k; }

sz_status_t szs_tuning_shape(int index, int *group, int *rows_per_lane, int *cooperative, int *bit_parallel,
                             int *packed, int *chained) {
    if (index < 0 || index >= szs_amd::shapes_count_k || !group || !rows_per_lane || !cooperative || !bit_parallel ||
        !packed || !chained)
        return sz_unexpected_dimensions_k;
    szs_amd::shape_t const shape = szs_amd::shapes_k[index];
    *group = shape.group, *rows_per_lane = shape.rows_per_lane;
    *cooperative = shape.cooperative, *bit_parallel = shape.bit_parallel, *packed = shape.packed;
    *chained = shape.chained;
    return sz_success_k;
}

int szs_tuning_shape_wide(int index) {
    return index >= 0 && index < szs_amd::shapes_count_k && szs_amd::shapes_k[index].wide;
}

int szs_tuning_packed_rows(void const *engine) {
    return engine ? static_cast<szs_amd::engine_t const *>(engine)->packed_rows : 0;
}

sz_status_t szs_tuning_get_ladder(int slot, int *shapes, int *max_rows, int capacity, int *count, int *beyond_shape) {
    if (slot < 0 || slot >= szs_amd::shape_slots_k || !count || !beyond_shape) return sz_unexpected_dimensions_k;
    std::lock_guard<std::mutex> lock(szs_amd::tuning_mutex());
    szs_amd::shape_table_t const &table = szs_amd::active_shape_table();
    *count = table.used[slot];
    *beyond_shape = table.multi_pass_shape[slot];
    for (int i = 0; i < table.used[slot] && i < capacity; ++i) {
        shapes[i] = table.order[slot][i];
        max_rows[i] = table.limits[slot][table.order[slot][i]];
    }
    return sz_success_k;
}

/** Validates that every shape can score the slot's family and that single-stripe shapes stay within capacity. */
sz_status_t szs_tuning_set_ladder(int slot, int const *shapes, int const *max_rows, int count, int beyond_shape,
                                  char const **error_message) {
    using namespace szs_amd;
    if (slot < 0 || slot >= shape_slots_k || count < 0 || count > shapes_count_k || (count && (!shapes || !max_rows)))
        return fail(sz_unexpected_dimensions_k, error_message, "Invalid tuning slot or ladder size");
    bool const unit = slot == unit_bytes_slot_k || slot == unit_runes_slot_k;
    int const family = unit ? family_lev_linear_k : slot / 2;
    bool const uniform = family == family_lev_linear_k || family == family_lev_affine_k;
    // Class-cost families never score runes: their second slot holds the engines whose biased costs leave int8,
    // which run the plain recurrences (no packed 16-bit form).
    bool const plain = !uniform && slot % 2 == 1;
    auto allowed = [&](int shape) {
        if (shape < 0 || shape >= shapes_count_k) return false;
        shape_t const s = shapes_k[shape];
        if (s.wide) return true; // int64 cells score every family, bytes or runes
        if (s.bit_parallel) return unit;
        // Chained groups hand over bottom rows only: global families, whose score is the last cell.
        if (s.chained && (unit || family == family_sw_linear_k || family == family_sw_affine_k)) return false;
        if (s.packed)
            return !uniform && !plain; // class-cost families, when the engine's range bound admits the rows
        if (!uniform && s.rows_per_lane > 12) return false; // LDS profiles hold at most 12 rows per lane
        return true;
    };
    if (!allowed(beyond_shape) || shapes_k[beyond_shape].group != 64)
        return fail(sz_unexpected_dimensions_k, error_message, "The shape for tall pairs must be a 64-lane shape");
    int previous = 0;
    for (int i = 0; i < count; ++i) {
        if (!allowed(shapes[i])) return fail(sz_unexpected_dimensions_k, error_message, "Shape not valid for this slot");
        shape_t const s = shapes_k[shapes[i]];
        if (max_rows[i] <= previous) return fail(sz_unexpected_dimensions_k, error_message, "Row limits must increase");
        if (s.group < 64 && max_rows[i] > s.group * s.rows_per_lane)
            return fail(sz_unexpected_dimensions_k, error_message, "A sub-wave shape holds a single stripe only");
        for (int j = 0; j < i; ++j)
            if (shapes[j] == shapes[i]) return fail(sz_unexpected_dimensions_k, error_message, "Shapes must be distinct");
        previous = max_rows[i];
    }
    std::lock_guard<std::mutex> lock(szs_amd::tuning_mutex());
    shape_table_t &table = active_shape_table();
    for (int shape = 0; shape < shapes_count_k; ++shape) table.limits[slot][shape] = -1;
    for (int i = 0; i < count; ++i) table.order[slot][i] = shapes[i], table.limits[slot][shapes[i]] = max_rows[i];
    table.used[slot] = count;
    table.multi_pass_shape[slot] = beyond_shape;
    return ok(error_message);
}

} // extern "C"

#pragma endregion
