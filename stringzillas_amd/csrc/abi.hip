/**
 *  @file abi.hip
 *  @brief The `szs_*` C ABI: the reference's symbols (include/stringzillas/stringzillas.h) and this library's
 *         pair, batch and serving extensions.
 *
 *  Reference-shaped calls score a cross-product into `results[q * stride + c]` (symmetric lower triangle plus
 *  mirror when `candidates == NULL`, as serial.hpp:2545-2580) by staging a one-segment batch.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "szs_runtime.hpp"

using namespace szs_amd;

namespace {

constexpr int version_major_k = 5, version_minor_k = 3, version_patch_k = 2; // the reference's VERSION

bool gpu_present() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return count > 0;
}

/** Default scope: device 0, created lazily (c/stringzillas/stringzillas.cuh:261-281 does the same). */
device_scope_t *as_scope(szs_device_scope_t scope) { return reinterpret_cast<device_scope_t *>(scope); }

/** `sz_sequence_t` members gathered into a host tape. */
struct gathered_tape_t {
    std::vector<char> data;
    std::vector<uint64_t> offsets;
    sz_sequence_u64tape_t view;
    void gather(sz_sequence_t const *sequence) {
        offsets.assign(sequence->count + 1, 0);
        size_t total = 0;
        for (size_t i = 0; i < sequence->count; ++i) total += sequence->get_length(sequence->handle, i);
        data.resize(total + 1);
        size_t at = 0;
        for (size_t i = 0; i < sequence->count; ++i) {
            size_t const length = sequence->get_length(sequence->handle, i);
            if (length) std::memcpy(data.data() + at, sequence->get_start(sequence->handle, i), length);
            at += length;
            offsets[i + 1] = at;
        }
        view.data = data.data();
        view.offsets = offsets.data();
        view.count = sequence->count;
    }
};

/** One call on the scope's workspace: its device buffers persist between calls and only grow (the reference keeps
 *  such scratch per engine, include/stringzillas/similarities/serial.hpp:3030-3040). */
sz_status_t run_segments(szs_device_scope_t scope, std::vector<segment_input_t> const &inputs,
                         char const **error_message) {
    device_scope_t *s = as_scope(scope);
    return guarded(error_message, [&]() {
        std::lock_guard<std::mutex> lock(s->call_mutex);
        if (!s->workspace) s->workspace = new batch_t();
        batch_t &batch = *s->workspace;
        sz_status_t status = batch_stage(batch, s, inputs, szs_batch_persistent_k, error_message);
        if (status == sz_success_k) status = batch_launch(&batch, error_message, true);
        if (status == sz_success_k) status = batch_wait(&batch, error_message);
        if (status != sz_success_k) (void)hipStreamSynchronize(batch.stream); // nothing of this call is in flight
        return status;
    });
}

template <typename tape_type_>
sz_status_t cross_call(int kind, void *engine_handle, szs_device_scope_t scope, tape_type_ const *queries,
                       tape_type_ const *candidates, void *results, sz_size_t stride, char const **error_message) {
    if (!engine_handle) return fail(sz_unexpected_dimensions_k, error_message, "Engine must be initialized");
    if (!queries) return fail(sz_unexpected_dimensions_k, error_message, "Query texts cannot be null");
    engine_t *engine = reinterpret_cast<engine_t *>(engine_handle);
    if (engine->kind != kind) return fail(sz_unexpected_dimensions_k, error_message, "Engine kind mismatch");
    std::lock_guard<std::mutex> lock(engine->mutex);
    segment_input_t input;
    input.engine = engine;
    input.mode = candidates ? segment_cross_k : segment_symmetric_k;
    if constexpr (sizeof(*queries->offsets) == 8) {
        input.queries = reinterpret_cast<sz_sequence_u64tape_t const *>(queries);
        input.candidates = reinterpret_cast<sz_sequence_u64tape_t const *>(candidates);
    }
    else {
        input.queries32 = reinterpret_cast<sz_sequence_u32tape_t const *>(queries);
        input.candidates32 = reinterpret_cast<sz_sequence_u32tape_t const *>(candidates);
    }
    input.results = results;
    input.results_row_stride = stride;
    if (!scope) return fail(sz_unexpected_dimensions_k, error_message, "Device must be initialized");
    return run_segments(scope, {input}, error_message);
}

sz_status_t sequence_call(int kind, void *engine, szs_device_scope_t scope, sz_sequence_t const *queries,
                          sz_sequence_t const *candidates, void *results, sz_size_t stride,
                          char const **error_message) {
    if (!queries) return fail(sz_unexpected_dimensions_k, error_message, "Query texts cannot be null");
    return guarded(error_message, [&]() {
        gathered_tape_t q, c;
        q.gather(queries);
        if (candidates) c.gather(candidates);
        return cross_call(kind, engine, scope, &q.view, candidates ? &c.view : nullptr, results, stride,
                          error_message);
    });
}

sz_status_t pairs_call(int kind, void *engine_handle, szs_device_scope_t scope, sz_sequence_u64tape_t const *firsts,
                       sz_sequence_u64tape_t const *seconds, void *results, char const **error_message) {
    if (!engine_handle) return fail(sz_unexpected_dimensions_k, error_message, "Engine must be initialized");
    if (!firsts || !seconds) return fail(sz_unexpected_dimensions_k, error_message, "Pair tapes cannot be null");
    engine_t *engine = reinterpret_cast<engine_t *>(engine_handle);
    if (engine->kind != kind) return fail(sz_unexpected_dimensions_k, error_message, "Engine kind mismatch");
    std::lock_guard<std::mutex> lock(engine->mutex);
    segment_input_t input;
    input.engine = engine;
    input.mode = segment_pairs_k;
    input.queries = firsts;
    input.candidates = seconds;
    input.results = results;
    if (!scope) return fail(sz_unexpected_dimensions_k, error_message, "Device must be initialized");
    return run_segments(scope, {input}, error_message);
}

sz_status_t batch_init(szs_device_scope_t device, szs_pairs_segment_t const *segments, sz_size_t segments_count,
                       szs_batch_mode_t mode, szs_batch_t *batch, char const **error_message) {
    std::vector<segment_input_t> inputs(segments_count);
    for (size_t s = 0; s < segments_count; ++s) {
        szs_pairs_segment_t const &segment = segments[s];
        if (!segment.engine || !segment.firsts || !segment.seconds)
            return fail(sz_unexpected_dimensions_k, error_message, "Segment fields must not be null");
        engine_t *engine = reinterpret_cast<engine_t *>(segment.engine);
        if (engine->kind != (int)segment.kind)
            return fail(sz_unexpected_dimensions_k, error_message, "Segment kind does not match its engine");
        inputs[s].engine = engine;
        inputs[s].mode = segment_pairs_k;
        inputs[s].queries = segment.firsts;
        inputs[s].candidates = segment.seconds;
        inputs[s].results = segment.results;
        inputs[s].queries_bytes = segment.firsts_bytes;
        inputs[s].candidates_bytes = segment.seconds_bytes;
    }
    batch_t *prepared = nullptr;
    sz_status_t const status = batch_prepare(as_scope(device), inputs, mode, &prepared, error_message);
    if (status != sz_success_k) return status;
    *batch = prepared;
    return ok(error_message);
}

} // namespace

extern "C" {

#pragma region Version and capabilities

int szs_version_major(void) { return version_major_k; }
int szs_version_minor(void) { return version_minor_k; }
int szs_version_patch(void) { return version_patch_k; }
sz_capability_t szs_capabilities_comptime(void) { return sz_cap_rocm_k; }
sz_capability_t szs_capabilities_runtime(void) { return gpu_present() ? sz_cap_rocm_k : sz_caps_none_k; }
sz_capability_t szs_capabilities(void) { return (sz_capability_t)(szs_capabilities_comptime() & szs_capabilities_runtime()); }

static void *unified_allocate(sz_size_t size, void *) { return szs_unified_alloc(size); }
static void unified_free(void *ptr, sz_size_t size, void *) { szs_unified_free(ptr, size); }

sz_status_t sz_memory_allocator_init_unified(sz_memory_allocator_t *alloc, char const **error_message) {
    if (!alloc) return fail(sz_unexpected_dimensions_k, error_message, "Allocator must not be null");
    alloc->allocate = unified_allocate;
    alloc->free = unified_free;
    alloc->handle = nullptr;
    return ok(error_message);
}

void *szs_unified_alloc(sz_size_t size_bytes) {
    if (!gpu_present()) return std::malloc(size_bytes ? size_bytes : 1);
    void *ptr = nullptr;
    if (hipMallocManaged(&ptr, size_bytes ? size_bytes : 1, hipMemAttachGlobal) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return ptr;
}

void szs_unified_free(void *ptr, sz_size_t) {
    if (!ptr) return;
    if (gpu_present() && is_device_pointer(ptr)) (void)hipFree(ptr);
    else std::free(ptr);
}

#pragma endregion

#pragma region Device scopes

sz_status_t szs_device_scope_init_default(szs_device_scope_t *scope, char const **error_message) {
    if (!scope) return fail(sz_unexpected_dimensions_k, error_message, "Scope output pointer is null");
    device_scope_t *s = new (std::nothrow) device_scope_t();
    if (!s) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a device scope");
    s->kind = scope_default_k;
    s->gpu_device = 0;
    *scope = s;
    return ok(error_message);
}

sz_status_t szs_device_scope_init_cpu_cores(sz_size_t cpu_cores, szs_device_scope_t *scope,
                                            char const **error_message) {
    if (!scope) return fail(sz_unexpected_dimensions_k, error_message, "Scope output pointer is null");
    device_scope_t *s = new (std::nothrow) device_scope_t();
    if (!s) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a device scope");
    s->kind = scope_cpu_k;
    s->cpu_cores = cpu_cores ? cpu_cores : std::thread::hardware_concurrency();
    *scope = s;
    return ok(error_message);
}

sz_status_t szs_device_scope_init_gpu_device(sz_size_t gpu_device, szs_device_scope_t *scope,
                                             char const **error_message) {
    if (!scope) return fail(sz_unexpected_dimensions_k, error_message, "Scope output pointer is null");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || (sz_size_t)count <= gpu_device) {
        (void)hipGetLastError();
        return fail(sz_missing_gpu_k, error_message, "Requested GPU device is not available");
    }
    device_scope_t *s = new (std::nothrow) device_scope_t();
    if (!s) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a device scope");
    s->kind = scope_gpu_k;
    s->gpu_device = (int)gpu_device;
    *scope = s;
    return ok(error_message);
}

sz_status_t szs_device_scope_get_cpu_cores(szs_device_scope_t scope, sz_size_t *cpu_cores, char const **error_message) {
    device_scope_t *s = as_scope(scope);
    if (!s || !cpu_cores) return fail(sz_unexpected_dimensions_k, error_message, "Scope must be initialized");
    if (s->kind != scope_cpu_k) return fail(sz_device_code_mismatch_k, error_message, "Scope does not target CPU cores");
    *cpu_cores = s->cpu_cores;
    return ok(error_message);
}

sz_status_t szs_device_scope_get_gpu_device(szs_device_scope_t scope, sz_size_t *gpu_device,
                                            char const **error_message) {
    device_scope_t *s = as_scope(scope);
    if (!s || !gpu_device) return fail(sz_unexpected_dimensions_k, error_message, "Scope must be initialized");
    if (s->kind != scope_gpu_k) return fail(sz_device_code_mismatch_k, error_message, "Scope does not target a GPU");
    *gpu_device = (sz_size_t)s->gpu_device;
    return ok(error_message);
}

sz_status_t szs_device_scope_get_capabilities(szs_device_scope_t scope, sz_capability_t *capabilities,
                                              char const **error_message) {
    device_scope_t *s = as_scope(scope);
    if (!s || !capabilities) return fail(sz_unexpected_dimensions_k, error_message, "Scope must be initialized");
    *capabilities = s->kind == scope_cpu_k ? (sz_capability_t)(sz_cap_serial_k | sz_cap_parallel_k)
                                           : szs_capabilities_runtime();
    return ok(error_message);
}

void szs_device_scope_free(szs_device_scope_t scope) {
    device_scope_t *s = as_scope(scope);
    if (!s) return;
    if (s->stream) {
        (void)hipSetDevice(s->gpu_device);
        (void)hipStreamDestroy(s->stream);
    }
    delete s;
}

#pragma endregion

#pragma region Engines

sz_status_t szs_levenshtein_distances_init(sz_error_cost_t match, sz_error_cost_t mismatch, sz_error_cost_t open,
                                           sz_error_cost_t extend, sz_memory_allocator_t const *,
                                           sz_capability_t, szs_levenshtein_distances_t *engine,
                                           char const **error_message) {
    return make_engine(szs_engine_levenshtein_k, match, mismatch, open, extend, nullptr, nullptr, engine,
                       error_message);
}
sz_status_t szs_levenshtein_distances(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                      sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                      sz_size_t *results, sz_size_t stride, char const **error_message) {
    return sequence_call(szs_engine_levenshtein_k, engine, device, queries, candidates, results, stride,
                         error_message);
}
sz_status_t szs_levenshtein_distances_u32tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                              sz_sequence_u32tape_t const *queries,
                                              sz_sequence_u32tape_t const *candidates, sz_size_t *results,
                                              sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_levenshtein_k, engine, device, queries, candidates, results, stride, error_message);
}
sz_status_t szs_levenshtein_distances_u64tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                              sz_sequence_u64tape_t const *queries,
                                              sz_sequence_u64tape_t const *candidates, sz_size_t *results,
                                              sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_levenshtein_k, engine, device, queries, candidates, results, stride, error_message);
}
void szs_levenshtein_distances_free(szs_levenshtein_distances_t engine) { delete reinterpret_cast<engine_t *>(engine); }

sz_status_t szs_levenshtein_distances_utf8_init(sz_error_cost_t match, sz_error_cost_t mismatch,
                                                sz_error_cost_t open, sz_error_cost_t extend,
                                                sz_memory_allocator_t const *, sz_capability_t,
                                                szs_levenshtein_distances_utf8_t *engine,
                                                char const **error_message) {
    return make_engine(szs_engine_levenshtein_utf8_k, match, mismatch, open, extend, nullptr, nullptr, engine,
                       error_message);
}
sz_status_t szs_levenshtein_distances_utf8(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                           sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                           sz_size_t *results, sz_size_t stride, char const **error_message) {
    return sequence_call(szs_engine_levenshtein_utf8_k, engine, device, queries, candidates, results, stride,
                         error_message);
}
sz_status_t szs_levenshtein_distances_utf8_u32tape(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                                   sz_sequence_u32tape_t const *queries,
                                                   sz_sequence_u32tape_t const *candidates, sz_size_t *results,
                                                   sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_levenshtein_utf8_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
sz_status_t szs_levenshtein_distances_utf8_u64tape(szs_levenshtein_distances_utf8_t engine, szs_device_scope_t device,
                                                   sz_sequence_u64tape_t const *queries,
                                                   sz_sequence_u64tape_t const *candidates, sz_size_t *results,
                                                   sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_levenshtein_utf8_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
void szs_levenshtein_distances_utf8_free(szs_levenshtein_distances_utf8_t engine) {
    delete reinterpret_cast<engine_t *>(engine);
}

sz_status_t szs_needleman_wunsch_scores_init(sz_u8_t const *byte_to_class,
                                             sz_error_cost_t const *class_substitution_costs, sz_error_cost_t open,
                                             sz_error_cost_t extend, sz_memory_allocator_t const *, sz_capability_t,
                                             szs_needleman_wunsch_scores_t *engine, char const **error_message) {
    return make_engine(szs_engine_needleman_wunsch_k, 0, 0, open, extend, byte_to_class,
                       reinterpret_cast<int8_t const *>(class_substitution_costs), engine, error_message);
}
sz_status_t szs_needleman_wunsch_scores(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                        sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                        sz_ssize_t *results, sz_size_t stride, char const **error_message) {
    return sequence_call(szs_engine_needleman_wunsch_k, engine, device, queries, candidates, results, stride,
                         error_message);
}
sz_status_t szs_needleman_wunsch_scores_u32tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                sz_sequence_u32tape_t const *queries,
                                                sz_sequence_u32tape_t const *candidates, sz_ssize_t *results,
                                                sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_needleman_wunsch_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
sz_status_t szs_needleman_wunsch_scores_u64tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                sz_sequence_u64tape_t const *queries,
                                                sz_sequence_u64tape_t const *candidates, sz_ssize_t *results,
                                                sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_needleman_wunsch_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
void szs_needleman_wunsch_scores_free(szs_needleman_wunsch_scores_t engine) {
    delete reinterpret_cast<engine_t *>(engine);
}

sz_status_t szs_smith_waterman_scores_init(sz_u8_t const *byte_to_class,
                                           sz_error_cost_t const *class_substitution_costs, sz_error_cost_t open,
                                           sz_error_cost_t extend, sz_memory_allocator_t const *, sz_capability_t,
                                           szs_smith_waterman_scores_t *engine, char const **error_message) {
    return make_engine(szs_engine_smith_waterman_k, 0, 0, open, extend, byte_to_class,
                       reinterpret_cast<int8_t const *>(class_substitution_costs), engine, error_message);
}
sz_status_t szs_smith_waterman_scores(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                      sz_sequence_t const *queries, sz_sequence_t const *candidates,
                                      sz_ssize_t *results, sz_size_t stride, char const **error_message) {
    return sequence_call(szs_engine_smith_waterman_k, engine, device, queries, candidates, results, stride,
                         error_message);
}
sz_status_t szs_smith_waterman_scores_u32tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                              sz_sequence_u32tape_t const *queries,
                                              sz_sequence_u32tape_t const *candidates, sz_ssize_t *results,
                                              sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_smith_waterman_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
sz_status_t szs_smith_waterman_scores_u64tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                              sz_sequence_u64tape_t const *queries,
                                              sz_sequence_u64tape_t const *candidates, sz_ssize_t *results,
                                              sz_size_t stride, char const **error_message) {
    return cross_call(szs_engine_smith_waterman_k, engine, device, queries, candidates, results, stride,
                      error_message);
}
void szs_smith_waterman_scores_free(szs_smith_waterman_scores_t engine) { delete reinterpret_cast<engine_t *>(engine); }

#pragma endregion

#pragma region Pairs and batches

sz_status_t szs_levenshtein_distances_pairs_u64tape(szs_levenshtein_distances_t engine, szs_device_scope_t device,
                                                    sz_sequence_u64tape_t const *firsts,
                                                    sz_sequence_u64tape_t const *seconds, sz_size_t *results,
                                                    char const **error_message) {
    return pairs_call(szs_engine_levenshtein_k, engine, device, firsts, seconds, results, error_message);
}
sz_status_t szs_levenshtein_distances_utf8_pairs_u64tape(szs_levenshtein_distances_utf8_t engine,
                                                         szs_device_scope_t device,
                                                         sz_sequence_u64tape_t const *firsts,
                                                         sz_sequence_u64tape_t const *seconds, sz_size_t *results,
                                                         char const **error_message) {
    return pairs_call(szs_engine_levenshtein_utf8_k, engine, device, firsts, seconds, results, error_message);
}
sz_status_t szs_needleman_wunsch_scores_pairs_u64tape(szs_needleman_wunsch_scores_t engine, szs_device_scope_t device,
                                                      sz_sequence_u64tape_t const *firsts,
                                                      sz_sequence_u64tape_t const *seconds, sz_ssize_t *results,
                                                      char const **error_message) {
    return pairs_call(szs_engine_needleman_wunsch_k, engine, device, firsts, seconds, results, error_message);
}
sz_status_t szs_smith_waterman_scores_pairs_u64tape(szs_smith_waterman_scores_t engine, szs_device_scope_t device,
                                                    sz_sequence_u64tape_t const *firsts,
                                                    sz_sequence_u64tape_t const *seconds, sz_ssize_t *results,
                                                    char const **error_message) {
    return pairs_call(szs_engine_smith_waterman_k, engine, device, firsts, seconds, results, error_message);
}

sz_status_t szs_batch_init(szs_device_scope_t device, szs_pairs_segment_t const *segments, sz_size_t segments_count,
                           szs_batch_mode_t mode, szs_batch_t *batch, char const **error_message) {
    if (!batch || (!segments && segments_count))
        return fail(sz_unexpected_dimensions_k, error_message, "Batch arguments must not be null");
    if (!device) return fail(sz_unexpected_dimensions_k, error_message, "Device must be initialized");
    return guarded(error_message, [&]() { return batch_init(device, segments, segments_count, mode, batch, error_message); });
}

sz_status_t szs_batch_launch(szs_batch_t batch, char const **error_message) {
    if (!batch) return fail(sz_unexpected_dimensions_k, error_message, "Batch must be initialized");
    return guarded(error_message, [&]() { return batch_launch(reinterpret_cast<batch_t *>(batch), error_message); });
}

sz_status_t szs_batch_wait(szs_batch_t batch, char const **error_message) {
    if (!batch) return fail(sz_unexpected_dimensions_k, error_message, "Batch must be initialized");
    return guarded(error_message, [&]() { return batch_wait(reinterpret_cast<batch_t *>(batch), error_message); });
}

sz_status_t szs_batch_run(szs_batch_t batch, char const **error_message) {
    sz_status_t const status = szs_batch_launch(batch, error_message);
    if (status != sz_success_k) return status;
    return szs_batch_wait(batch, error_message);
}

sz_size_t szs_batch_cells(szs_batch_t batch) { return batch ? reinterpret_cast<batch_t *>(batch)->cells : 0; }

sz_size_t szs_batch_core_ops(szs_batch_t batch) { return batch ? reinterpret_cast<batch_t *>(batch)->core_ops : 0; }

sz_size_t szs_batch_device_bytes(szs_batch_t batch) {
    return batch ? reinterpret_cast<batch_t *>(batch)->device_bytes() : 0;
}

double szs_batch_last_kernel_ms(szs_batch_t batch) {
    return batch ? reinterpret_cast<batch_t *>(batch)->last_kernel_ms : 0;
}

void szs_batch_free(szs_batch_t batch) { delete reinterpret_cast<batch_t *>(batch); }

#pragma endregion

} // extern "C"
