/**
 *  @file tape_file.hip
 *  @brief On-disk string tapes, mapped and streamed into HBM.
 *
 *  File layout (little-endian, every section 64-byte aligned):
 *    [0, 8)    magic "SZSTAPE1"
 *    [8, 16)   u64 count of strings
 *    [16, 24)  u64 bytes of string data (= offsets[count])
 *    [24, 64)  reserved, zero
 *    [64, ...) u64 offsets[count + 1], offsets[0] = 0, non-decreasing (Arrow LargeBinary offsets)
 *    then, at the next multiple of 64: u8 data[bytes]
 *  The offsets and data sections are exactly an Arrow `large_binary` array's buffers, so Arrow data is written
 *  and read without re-encoding.
 *
 *  `szs_tape_load` maps the file read-only, validates the header and the offsets, and copies both sections into
 *  device memory through two page-locked staging buffers (one chunk copies while the next is filled), so a tape
 *  larger than host memory streams through. The device copy is an `sz_sequence_u64tape_t` of HBM pointers that
 *  every scoring entry point and the server take in place. No reference counterpart (the reference reads tapes
 *  from host memory only).
 */
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "szs_runtime.hpp"

using namespace szs_amd;

namespace {

char const tape_magic_k[8] = {'S', 'Z', 'S', 'T', 'A', 'P', 'E', '1'};
constexpr size_t tape_header_bytes_k = 64;
constexpr size_t tape_stage_bytes_k = size_t(32) << 20;

size_t align64(size_t value) { return (value + 63) & ~size_t(63); }

struct device_tape_t {
    device_buffer_t data, offsets;
    uint64_t count = 0, bytes = 0;
    int device = 0;
};

struct mapping_t {
    void *base = MAP_FAILED;
    size_t size = 0;
    int fd = -1;
    ~mapping_t() {
        if (base != MAP_FAILED) munmap(base, size);
        if (fd >= 0) close(fd);
    }
};

/** Copies host bytes (possibly pageable, e.g. a file mapping) to the device through two pinned buffers. */
sz_status_t stream_to_device(void *device_destination, void const *source, size_t bytes, hipStream_t stream,
                             char const **error_message) {
    if (!bytes) return ok(error_message);
    void *stages[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    sz_status_t status = sz_success_k;
    size_t const chunk = std::min(bytes, tape_stage_bytes_k);
    for (int s = 0; s < 2 && status == sz_success_k; ++s) {
        if (hipHostMalloc(&stages[s], chunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&done[s], hipEventDisableTiming) != hipSuccess)
            status = from_hip(hipErrorOutOfMemory, error_message, "Failed to allocate page-locked staging");
    }
    for (size_t at = 0, index = 0; status == sz_success_k && at < bytes; at += chunk, ++index) {
        int const s = (int)(index % 2);
        size_t const length = std::min(chunk, bytes - at);
        if (index >= 2 && hipEventSynchronize(done[s]) != hipSuccess) // the copy that used this stage is done
            status = from_hip(hipErrorUnknown, error_message, "Tape upload failed");
        if (status != sz_success_k) break;
        std::memcpy(stages[s], static_cast<char const *>(source) + at, length);
        hipError_t error = hipMemcpyAsync(static_cast<char *>(device_destination) + at, stages[s], length,
                                          hipMemcpyHostToDevice, stream);
        if (error == hipSuccess) error = hipEventRecord(done[s], stream);
        if (error != hipSuccess) status = from_hip(error, error_message, "Tape upload failed");
    }
    if (status == sz_success_k) status = from_hip(hipStreamSynchronize(stream), error_message, "Tape upload failed");
    for (int s = 0; s < 2; ++s) {
        if (done[s]) (void)hipEventDestroy(done[s]);
        if (stages[s]) (void)hipHostFree(stages[s]);
    }
    return status;
}

} // namespace

extern "C" {

sz_status_t szs_tape_save(char const *path, sz_sequence_u64tape_t const *tape, char const **error_message) {
    if (!path || !tape || (tape->count && (!tape->offsets || !tape->data)))
        return fail(sz_unexpected_dimensions_k, error_message, "Tape arguments must not be null");
    if (tape->count && (is_device_pointer(tape->offsets) || is_device_pointer(tape->data)))
        return fail(sz_device_memory_mismatch_k, error_message, "Saving takes a host tape");
    uint64_t const count = tape->count;
    uint64_t const bytes = count ? tape->offsets[count] : 0;
    uint64_t const first = count ? tape->offsets[0] : 0;
    FILE *file = std::fopen(path, "wb");
    if (!file) return fail(sz_unexpected_dimensions_k, error_message, "Cannot open the tape file for writing");
    char header[tape_header_bytes_k] = {};
    std::memcpy(header, tape_magic_k, 8);
    uint64_t const stored_bytes = bytes - first;
    std::memcpy(header + 8, &count, 8);
    std::memcpy(header + 16, &stored_bytes, 8);
    bool good = std::fwrite(header, 1, sizeof(header), file) == sizeof(header);
    // Offsets rebased to start at 0, so a slice of a larger tape saves as a standalone file.
    for (uint64_t i = 0; good && i <= count; ++i) {
        uint64_t const offset = count ? tape->offsets[i] - first : 0;
        good = std::fwrite(&offset, 8, 1, file) == 1;
    }
    size_t const offsets_end = tape_header_bytes_k + (count + 1) * 8;
    static char const zeros[64] = {};
    good = good && std::fwrite(zeros, 1, align64(offsets_end) - offsets_end, file) == align64(offsets_end) - offsets_end;
    if (good && stored_bytes) good = std::fwrite(tape->data + first, 1, stored_bytes, file) == stored_bytes;
    good = (std::fclose(file) == 0) && good;
    if (!good) return fail(sz_status_unknown_k, error_message, "Writing the tape file failed");
    return ok(error_message);
}

sz_status_t szs_tape_load(szs_device_scope_t device, char const *path, szs_device_tape_t *handle,
                          sz_sequence_u64tape_t *view, char const **error_message) {
    if (!device || !path || !handle || !view)
        return fail(sz_unexpected_dimensions_k, error_message, "Tape arguments must not be null");
    device_scope_t *scope = reinterpret_cast<device_scope_t *>(device);
    int gpu = 0;
    hipStream_t stream = nullptr;
    sz_status_t status = resolve_gpu(scope, &gpu, &stream, error_message);
    if (status != sz_success_k) return status;

    mapping_t map;
    map.fd = open(path, O_RDONLY);
    if (map.fd < 0) return fail(sz_unexpected_dimensions_k, error_message, "Cannot open the tape file");
    struct stat info;
    if (fstat(map.fd, &info) != 0 || (size_t)info.st_size < tape_header_bytes_k + 8)
        return fail(sz_unexpected_dimensions_k, error_message, "The tape file is too short");
    map.size = (size_t)info.st_size;
    map.base = mmap(nullptr, map.size, PROT_READ, MAP_PRIVATE, map.fd, 0);
    if (map.base == MAP_FAILED) return fail(sz_bad_alloc_k, error_message, "Cannot map the tape file");
    char const *file = static_cast<char const *>(map.base);
    uint64_t count = 0, bytes = 0;
    std::memcpy(&count, file + 8, 8);
    std::memcpy(&bytes, file + 16, 8);
    if (std::memcmp(file, tape_magic_k, 8) != 0)
        return fail(sz_unexpected_dimensions_k, error_message, "Not a tape file (bad magic)");
    if (count > (map.size - tape_header_bytes_k) / 8 - 1)
        return fail(sz_unexpected_dimensions_k, error_message, "The tape file's count exceeds its size");
    size_t const data_start = align64(tape_header_bytes_k + (count + 1) * 8);
    if (data_start > map.size || bytes > map.size - data_start)
        return fail(sz_unexpected_dimensions_k, error_message, "The tape file's byte count exceeds its size");
    uint64_t const *offsets = reinterpret_cast<uint64_t const *>(file + tape_header_bytes_k);
    if (offsets[0] != 0 || offsets[count] != bytes)
        return fail(sz_unexpected_dimensions_k, error_message, "The tape file's offsets do not span its data");
    for (uint64_t i = 0; i < count; ++i)
        if (offsets[i + 1] < offsets[i])
            return fail(sz_unexpected_dimensions_k, error_message, "The tape file's offsets decrease");

    device_tape_t *tape = new (std::nothrow) device_tape_t();
    if (!tape) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a tape");
    tape->count = count, tape->bytes = bytes, tape->device = gpu;
    hipError_t error = tape->offsets.reserve((count + 1) * 8);
    if (error == hipSuccess) error = tape->data.reserve(std::max<uint64_t>(bytes, 1) + 16);
    if (error != hipSuccess) {
        delete tape;
        return from_hip(error, error_message, "Failed to allocate the tape in HBM");
    }
    status = stream_to_device(tape->offsets.data, offsets, (count + 1) * 8, stream, error_message);
    if (status == sz_success_k) status = stream_to_device(tape->data.data, file + data_start, bytes, stream, error_message);
    if (status != sz_success_k) {
        delete tape;
        return status;
    }
    view->data = tape->data.as<char>();
    view->offsets = tape->offsets.as<sz_u64_t>();
    view->count = count;
    *handle = tape;
    return ok(error_message);
}

sz_size_t szs_tape_bytes(szs_device_tape_t handle) {
    return handle ? (sz_size_t) reinterpret_cast<device_tape_t *>(handle)->bytes : 0;
}

void szs_tape_free(szs_device_tape_t handle) { delete reinterpret_cast<device_tape_t *>(handle); }

} // extern "C"
