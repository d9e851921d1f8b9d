/**
 *  @file noalloc.hip
 *  @brief `szs_batch_run_noalloc`: the mixed-batch call for callers that cannot allocate device memory.
 *
 *  No reference counterpart. It scores what `szs_batch_init` (persistent mode) + `szs_batch_run` + `szs_batch_free`
 *  score, with the same kernels, but every device buffer the call needs (tasks, bins, counters, rune tapes,
 *  stripe-boundary rows, chain rows) is carved from `scratch`, one static device array per GPU, so the call itself
 *  never reaches `hipMalloc`/`hipFree`. The array is reserved when the library's code object is loaded on a GPU
 *  and held until the process ends.
 *
 *  Host code never uses `scratch` by name: the first call on a device, with that device current, asks HIP for the
 *  array's address on it (`hipGetSymbolAddress`) and keeps it; every later use (the pool the buffers are carved
 *  from, hence every kernel argument) goes through that address. Calls on one GPU share the array and would race,
 *  so they are serialized per GPU; the array keeps the last call's contents, and everything a launch needs at
 *  zero is cleared by fill kernels before the launch reads it.
 */
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "szs_runtime.hpp"

using namespace szs_amd;

namespace {

/** Bytes of the static scratch array (per GPU); must equal `SZS_NOALLOC_SCRATCH_BYTES` in the public header. */
constexpr size_t noalloc_scratch_bytes_k = size_t(256) << 20;
static_assert(noalloc_scratch_bytes_k == SZS_NOALLOC_SCRATCH_BYTES, "header and library agree on the size");

static __device__ __align__(256) uint8_t scratch[noalloc_scratch_bytes_k];

/** One GPU's view of the array: its address there, the pool over it, and the batch object staged into it. */
struct noalloc_device_t {
    std::mutex mutex; /**< one call at a time per GPU: they share the array */
    void *scratch_address = nullptr;
    arena_t pool;
    std::unique_ptr<batch_t> batch;
};

std::mutex &devices_mutex() {
    static std::mutex mutex;
    return mutex;
}

/** Never destroyed: a GPU's state lives as long as its copy of the array, until the process ends (destroying its
 *  events from a static destructor could run after the HIP runtime has shut down). */
std::map<int, noalloc_device_t *> &devices() {
    static auto *map = new std::map<int, noalloc_device_t *>();
    return *map;
}

/** Set-up, once per device with `device` current: the array's address there, and the batch's events. */
sz_status_t device_state(int device, noalloc_device_t **out, char const **error_message) {
    std::lock_guard<std::mutex> lock(devices_mutex());
    noalloc_device_t *&slot = devices()[device];
    if (!slot) {
        auto state = std::make_unique<noalloc_device_t>();
        hipError_t error = hipSetDevice(device);
        if (error == hipSuccess) error = hipGetSymbolAddress(&state->scratch_address, HIP_SYMBOL(scratch));
        if (error != hipSuccess || !state->scratch_address)
            return from_hip(error == hipSuccess ? hipErrorInvalidSymbol : error, error_message,
                            "Locating the static scratch array on the device failed");
        state->pool.adopt(state->scratch_address, noalloc_scratch_bytes_k);
        state->pool.exclusive = true;
        state->batch = std::make_unique<batch_t>();
        error = hipEventCreate(&state->batch->kernel_begin);
        if (error == hipSuccess) error = hipEventCreate(&state->batch->kernel_end);
        if (error != hipSuccess) return from_hip(error, error_message, "Creating the call's events failed");
        slot = state.release();
    }
    *out = slot;
    return ok(error_message);
}

/** Restores the calling thread's pool when the call ends. */
struct pool_scope_t {
    arena_t *previous;
    explicit pool_scope_t(arena_t *pool) : previous(thread_arena()) { thread_arena() = pool; }
    ~pool_scope_t() { thread_arena() = previous; }
};

} // namespace

extern "C" {

sz_status_t szs_batch_run_noalloc(szs_device_scope_t device, szs_pairs_segment_t const *segments,
                                  sz_size_t segments_count, char const **error_message) {
    if (!device || (!segments && segments_count))
        return fail(sz_unexpected_dimensions_k, error_message, "Batch arguments must not be null");
    device_scope_t *scope = reinterpret_cast<device_scope_t *>(device);
    return guarded(error_message, [&]() -> sz_status_t {
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
        int gpu = 0;
        hipStream_t stream = nullptr;
        sz_status_t status = resolve_gpu(scope, &gpu, &stream, error_message);
        if (status != sz_success_k) return status;
        // Staging host tapes or results would need device buffers of their own: everything must already be there.
        for (size_t s = 0; s < segments_count; ++s) {
            szs_pairs_segment_t const &segment = segments[s];
            bool const empty = segment.firsts->count == 0;
            bool const resident = empty || (is_device_pointer(segment.firsts->data) &&
                                            is_device_pointer(segment.firsts->offsets) &&
                                            is_device_pointer(segment.seconds->data) &&
                                            is_device_pointer(segment.seconds->offsets) &&
                                            is_device_pointer(segment.results));
            if (!resident)
                return fail(sz_device_memory_mismatch_k, error_message,
                            "The allocation-free call takes tapes and results in device memory only");
        }
        noalloc_device_t *state = nullptr;
        status = device_state(gpu, &state, error_message);
        if (status != sz_success_k) return status;
        std::lock_guard<std::mutex> lock(state->mutex);
        pool_scope_t const pool(&state->pool);
        batch_t &batch = *state->batch;
        batch.release_device_memory(); // the whole array is free again: every call lays it out the same way
        batch_options_t options;
        options.max_scratch_bytes = noalloc_scratch_bytes_k / 2;
        status = batch_stage(batch, scope, inputs, szs_batch_persistent_k, error_message, options);
        if (status == sz_bad_alloc_k)
            return fail(sz_bad_alloc_k, error_message,
                        "The input needs more device memory than the allocation-free call's static scratch "
                        "(SZS_NOALLOC_SCRATCH_BYTES per GPU) holds");
        if (status == sz_success_k) status = batch_launch(&batch, error_message, true);
        if (status == sz_success_k) status = batch_wait(&batch, error_message);
        if (status != sz_success_k) (void)hipStreamSynchronize(batch.stream);
        batch.reset(); // forget the caller's pointers; the buffers stay in the array until the next call
        return status;
    });
}

} // extern "C"
