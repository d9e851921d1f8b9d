/**
 *  @file runtime.hip
 *  @brief Host runtime of the MI355X similarity engine: device scopes, engines, tape staging, batch preparation
 *         (materialize, bin, persistent launch, graph capture, per-pair baseline) and the `szs_*` C ABI.
 *
 *  The C ABI mirrors include/stringzillas/stringzillas.h of the reference (names, signatures, statuses); the
 *  engine selection mirrors c/stringzillas/levenshtein.cuh:96-179 and needleman_wunsch.cuh:86-164 (`open == extend`
 *  picks the linear recurrence). Everything scores on the GPU; there is no CPU scoring path in this library.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/stringzillas_amd.h"
#include "szs_device.hpp"
#include "szs_runtime.hpp"

using namespace szs_amd;

extern "C" {
hipError_t szs_amd_launch_fill(uint32_t *data, uint32_t count, uint32_t value, hipStream_t stream);
hipError_t szs_amd_launch_widen_offsets(uint32_t const *narrow, uint64_t *wide, uint64_t count, hipStream_t stream);
}

namespace szs_amd {

#pragma region Status helpers

sz_status_t fail(sz_status_t status, char const **error_message, char const *text) {
    if (error_message) *error_message = text;
    return status;
}

sz_status_t ok(char const **error_message) {
    if (error_message) *error_message = nullptr;
    return sz_success_k;
}

sz_status_t from_hip(hipError_t error, char const **error_message, char const *context) {
    if (error == hipSuccess) return ok(error_message);
    (void)hipGetLastError();
    if (error_message) *error_message = context;
    if (error == hipErrorOutOfMemory || error == hipErrorMemoryAllocation) return sz_bad_alloc_k;
    if (error == hipErrorNoDevice || error == hipErrorInvalidDevice) return sz_missing_gpu_k;
    return sz_status_unknown_k;
}

#pragma endregion

#pragma region Device buffers

hipError_t arena_t::init(size_t bytes) {
    capacity = (bytes + 4095) & ~size_t(4095);
    void *pointer = nullptr;
    hipError_t const error = hipMalloc(&pointer, capacity);
    if (error != hipSuccess) {
        (void)hipGetLastError();
        capacity = 0;
        return error;
    }
    base = static_cast<char *>(pointer);
    free_blocks.clear();
    free_blocks[0] = capacity;
    return hipSuccess;
}

void arena_t::adopt(void *memory, size_t bytes) {
    base = static_cast<char *>(memory);
    capacity = bytes & ~size_t(4095);
    owned = false;
    free_blocks.clear();
    if (capacity) free_blocks[0] = capacity;
}

void *arena_t::allocate(size_t bytes) {
    std::lock_guard<std::mutex> lock(mutex);
    for (auto block = free_blocks.begin(); block != free_blocks.end(); ++block) {
        if (block->second < bytes) continue;
        size_t const offset = block->first, left = block->second - bytes;
        free_blocks.erase(block);
        if (left) free_blocks[offset + bytes] = left;
        return base + offset;
    }
    return nullptr;
}

void arena_t::deallocate(void *pointer, size_t bytes) {
    std::lock_guard<std::mutex> lock(mutex);
    size_t offset = static_cast<char *>(pointer) - base;
    auto next = free_blocks.lower_bound(offset);
    if (next != free_blocks.end() && offset + bytes == next->first) { // merge with the following block
        bytes += next->second;
        next = free_blocks.erase(next);
    }
    if (next != free_blocks.begin()) { // merge with the preceding block
        auto previous = std::prev(next);
        if (previous->first + previous->second == offset) {
            previous->second += bytes;
            return;
        }
    }
    free_blocks[offset] = bytes;
}

arena_t::~arena_t() {
    if (base && owned) (void)hipFree(base);
}

arena_t *&thread_arena() {
    static thread_local arena_t *arena = nullptr;
    return arena;
}

void device_buffer_t::release() {
    if (data && arena) arena->deallocate(data, capacity);
    else if (data) (void)hipFree(data);
    data = nullptr;
    arena = nullptr;
    capacity = 0;
}

hipError_t device_buffer_t::reserve(size_t bytes) {
    if (bytes <= capacity) return hipSuccess;
    release();
    size_t const rounded = (bytes + 4095) & ~size_t(4095);
    if (arena_t *pool = thread_arena()) {
        data = pool->allocate(rounded);
        if (data) {
            arena = pool;
            capacity = rounded;
            return hipSuccess;
        }
        if (pool->exclusive) return hipErrorOutOfMemory;
    }
    hipError_t const error = hipMalloc(&data, rounded);
    if (error != hipSuccess) {
        data = nullptr;
        (void)hipGetLastError();
        return error;
    }
    capacity = rounded;
    return hipSuccess;
}

hipError_t order_after_default_stream(hipStream_t stream) {
    int device = 0;
    hipError_t error = hipGetDevice(&device);
    if (error != hipSuccess) return error;
    static thread_local std::map<int, hipEvent_t> events; // one per device, kept for the thread's life
    hipEvent_t &event = events[device];
    if (!event) {
        error = hipEventCreateWithFlags(&event, hipEventDisableTiming);
        if (error != hipSuccess) {
            event = nullptr;
            return error;
        }
    }
    error = hipEventRecord(event, nullptr);
    if (error == hipSuccess) error = hipStreamWaitEvent(stream, event, 0);
    return error;
}

bool is_device_pointer(void const *pointer) {
    if (!pointer) return false;
    hipPointerAttribute_t attributes;
    hipError_t const error = hipPointerGetAttributes(&attributes, pointer);
    if (error != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attributes.type == hipMemoryTypeDevice || attributes.type == hipMemoryTypeManaged ||
           attributes.type == hipMemoryTypeUnified;
}

#pragma endregion

#pragma region Scopes

device_scope_t::~device_scope_t() {
    if (!workspace) return;
    (void)hipSetDevice(gpu_device);
    delete workspace;
}

sz_status_t resolve_gpu(device_scope_t *scope, int *device, hipStream_t *stream, char const **error_message) {
    if (!scope) return fail(sz_device_code_mismatch_k, error_message, "Device scope must be initialized");
    if (scope->kind == scope_cpu_k)
        return fail(sz_device_code_mismatch_k, error_message, "GPU engines cannot run on a CPU-cores device scope");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(sz_missing_gpu_k, error_message, "No AMD GPU is visible to this process");
    }
    if (scope->gpu_device >= count) return fail(sz_missing_gpu_k, error_message, "GPU device index is out of range");
    std::lock_guard<std::mutex> lock(scope->mutex);
    hipError_t error = hipSetDevice(scope->gpu_device);
    if (error != hipSuccess) return from_hip(error, error_message, "hipSetDevice failed");
    if (!scope->stream) {
        error = hipStreamCreateWithFlags(&scope->stream, hipStreamNonBlocking);
        if (error != hipSuccess) return from_hip(error, error_message, "hipStreamCreate failed");
    }
    *device = scope->gpu_device;
    *stream = scope->stream;
    return ok(error_message);
}

#pragma endregion

#pragma region Engines

/** Maps `byte_to_class` onto the dense range of classes it actually uses, so LDS profiles stay small. */
static sz_status_t densify(engine_t *engine, uint8_t const *byte_to_class, int8_t const *class_costs,
                           char const **error_message) {
    int dense_of_class[32];
    int class_of_dense[32];
    for (int c = 0; c < 32; ++c) dense_of_class[c] = -1;
    int dense = 0;
    for (int b = 0; b < 256; ++b) {
        uint8_t const c = byte_to_class[b];
        if (c >= 32)
            return fail(sz_unexpected_dimensions_k, error_message, "byte_to_class values must be below 32");
        if (dense_of_class[c] < 0) class_of_dense[dense] = c, dense_of_class[c] = dense++;
        engine->params.byte_to_dense[b] = (uint8_t)dense_of_class[c];
    }
    engine->params.dense_count = dense;
    std::memset(engine->params.dense_costs, 0, sizeof(engine->params.dense_costs));
    for (int x = 0; x < dense; ++x)
        for (int y = 0; y < dense; ++y)
            engine->params.dense_costs[x * max_dense_classes_k + y] =
                class_costs[class_of_dense[x] * 32 + class_of_dense[y]];
    int magnitude = 0;
    for (int i = 0; i < 1024; ++i) magnitude = std::max(magnitude, std::abs((int)class_costs[i]));
    engine->substitution_magnitude = magnitude; // over all 1024 entries, as serial.hpp:180-186
    return ok(error_message);
}

/**
 *  The packed 16-bit sweep (packed_sweep.hpp) stores cells as u16 around a base per lane, re-chosen every 64
 *  columns, in [0, 0x7BFF] (finite non-negative halves, compared by a float maximum), so it needs every value a
 *  lane holds during a chunk within 15871 of that base. For a maximizing global family with o, e <= 0 (the Gotoh
 *  recurrences of serial.hpp:854-964; linear is o = e):
 *  - C(i,j) - C(i-1,j) >= o, since V(i,j) >= C(i-1,j) + o;
 *  - C(i,j) - C(i-1,j) <= Dp = max(0, s_max - min(o, e)) + max(0, o - e): the diagonal term gains at most
 *    s_max - min(o, e) over C(i-1,j) (which is at least C(i-1,j-1) + min(o, e)), V(i,j) <= C(i-1,j), and H(i,j)
 *    gains at most Dp over H(i-1,j) by induction along the row (the border row adds o - e when o > e);
 *  - the same holds along rows, so adjacent cells of Ĉ = C - e(i+j) differ by at most
 *    delta = max(|o|, Dp) + |e|;
 *  - gap tracks stay within max(|o|, Dp) + |e| + |o| of their cell, and every candidate sum within
 *    2 delta + |s_min| + |o| + |e| + |o - e|.
 *  The base is the lane's first row at its current column when the chunk starts; the cells the lane holds during
 *  the chunk (its R rows and the row above, from its lower half's column to 64 columns on) lie within a Manhattan
 *  distance of R + 63 of it; R + 70 is used. Profile bytes `cost - (o + e)` must also lie in [0, 127] to read as
 *  unsigned bytes.
 *  Local gaps (serial.hpp:762-841 linear, 977-1079 affine; o, e <= 0) keep X = H + o instead. There
 *  H(i,j) - H(i-1,j) >= o (the gap opened from above, or the zero floor), and the upper bound of the global
 *  induction holds as well, since the floor only lowers it and the border row is 0: the same limits apply (their
 *  delta adds |e| for the anti-diagonal domain, which the local form does not need). The profile bytes hold
 *  `cost - packed_bias` with packed_bias = min(o, s_min), and the substitution term subtracts o - packed_bias
 *  again (modulo 2^16, so only the corrected sum must be in range). The zero floor is the lowest value of every
 *  cell, and the gap tracks lie between the cell beside them plus o and the cell they end in.
 */
int packed_rows_limit(engine_params_t const &params) {
    bool const local = params.family == family_sw_linear_k || params.family == family_sw_affine_k;
    if (params.family != family_nw_linear_k && params.family != family_nw_affine_k && !local) return 0;
    int64_t const o = params.open, e = params.extend;
    if (o > 0 || e > 0 || params.shift != (local ? o : o + e) || params.dense_count == 0) return 0;
    int64_t s_max = INT64_MIN, s_min = INT64_MAX;
    for (int x = 0; x < params.dense_count; ++x)
        for (int y = 0; y < params.dense_count; ++y) {
            int64_t const cost = params.dense_costs[x * max_dense_classes_k + y];
            if (cost - params.packed_bias < 0 || cost - params.packed_bias > 127) return 0;
            s_max = std::max(s_max, cost), s_min = std::min(s_min, cost);
        }
    int64_t const rise = std::max<int64_t>(0, s_max - std::min(o, e)) + std::max<int64_t>(0, o - e);
    int64_t const step = std::max<int64_t>(-o, rise);
    int64_t const delta = step + (-e);
    int64_t const slack = 2 * delta + std::abs(s_min) + (-o) + (-e) + std::abs(o - e) + 64;
    for (int rows_per_lane = 8; rows_per_lane >= 2; rows_per_lane -= 2)
        if (delta * (rows_per_lane + 70) + slack <= 15871) return rows_per_lane;
    return 0;
}

/**
 *  The same bound for one base per wave, taken at lane 32 (the middle of the stripe) when a chunk starts: the
 *  cells any lane holds during the chunk lie within 32R + 1 rows and 129 columns of that cell (lanes run two
 *  columns apart, plus the chunk's 64 steps), so R rows per lane need delta * (32R + 140) + slack <= 15871. The
 *  sweep then moves no value between lane bases.
 */
int packed_uniform_rows_limit(engine_params_t const &params, int packed_rows) {
    if (packed_rows == 0) return 0;
    int64_t const o = params.open, e = params.extend;
    int64_t s_max = INT64_MIN, s_min = INT64_MAX;
    for (int x = 0; x < params.dense_count; ++x)
        for (int y = 0; y < params.dense_count; ++y) {
            int64_t const cost = params.dense_costs[x * max_dense_classes_k + y];
            s_max = std::max(s_max, cost), s_min = std::min(s_min, cost);
        }
    int64_t const rise = std::max<int64_t>(0, s_max - std::min(o, e)) + std::max<int64_t>(0, o - e);
    int64_t const delta = std::max<int64_t>(-o, rise) + (-e);
    int64_t const slack = 2 * delta + std::abs(s_min) + (-o) + (-e) + std::abs(o - e) + 64;
    for (int rows_per_lane = packed_rows; rows_per_lane >= 2; rows_per_lane -= 2)
        if (delta * (32 * rows_per_lane + 140) + slack <= 15871) return rows_per_lane;
    return 0;
}

sz_status_t make_engine(int kind, int match, int mismatch, int open, int extend, uint8_t const *byte_to_class,
                        int8_t const *class_costs, void **handle, char const **error_message) {
    if (!handle) return fail(sz_unexpected_dimensions_k, error_message, "Engine output pointer is null");
    engine_t *engine = new (std::nothrow) engine_t();
    if (!engine) return fail(sz_bad_alloc_k, error_message, "Failed to allocate the engine");
    std::memset(&engine->params, 0, sizeof(engine->params));
    engine->kind = kind;
    bool const linear = open == extend;
    engine->params.open = open;
    engine->params.extend = extend;
    engine->params.match = match;
    engine->params.mismatch = mismatch;
    engine->gap_magnitude = std::max(std::abs(open), std::abs(extend));
    if (kind == szs_engine_levenshtein_k || kind == szs_engine_levenshtein_utf8_k) {
        if (match < 0 || mismatch < 0 || open < 0 || extend < 0) {
            delete engine;
            return fail(sz_overflow_risk_k, error_message,
                        "Levenshtein costs must be non-negative: distances are kept in unsigned cells");
        }
        engine->params.family = linear ? family_lev_linear_k : family_lev_affine_k;
        engine->params.dense_count = 0;
        engine->substitution_magnitude = std::max(std::abs(match), std::abs(mismatch));
    }
    else {
        if (!byte_to_class || !class_costs) {
            delete engine;
            return fail(sz_unexpected_dimensions_k, error_message, "Substitution tables must not be null");
        }
        sz_status_t const status = densify(engine, byte_to_class, class_costs, error_message);
        if (status != sz_success_k) {
            delete engine;
            return status;
        }
        if (kind == szs_engine_needleman_wunsch_k) engine->params.family = linear ? family_nw_linear_k : family_nw_affine_k;
        else engine->params.family = linear ? family_sw_linear_k : family_sw_affine_k;
    }
    // The sweep's cost bias: o + e (affine global, anti-diagonal domain), 2g (linear global), o (local, scores
    // shifted by the open cost). The LDS profile holds `cost - bias` as int8; engines with a used cost that would
    // not fit fall back to the plain recurrences (bias 0).
    int const family = engine->params.family;
    bool const local = family == family_sw_linear_k || family == family_sw_affine_k;
    engine->params.shift = local ? open : open + extend; // 2g when open == extend
    for (int x = 0; x < engine->params.dense_count; ++x)
        for (int y = 0; y < engine->params.dense_count; ++y) {
            int const biased = engine->params.dense_costs[x * max_dense_classes_k + y] - engine->params.shift;
            if (biased < -128 || biased > 127) engine->params.shift = 0;
        }
    engine->params.packed_bias = engine->params.shift;
    if (local && engine->params.shift == open)
        for (int x = 0; x < engine->params.dense_count; ++x)
            for (int y = 0; y < engine->params.dense_count; ++y)
                engine->params.packed_bias = std::min<int32_t>(engine->params.packed_bias,
                                                               engine->params.dense_costs[x * max_dense_classes_k + y]);
    engine->packed_rows = packed_rows_limit(engine->params);
    engine->params.packed_uniform_rows = packed_uniform_rows_limit(engine->params, engine->packed_rows);
    *handle = engine;
    return ok(error_message);
}

#pragma endregion

} // namespace szs_amd
