/**
 *  @file server.hip
 *  @brief Continuous-batching request loop over the batch engine.
 *
 *  Requests wait in a host FIFO until their HBM footprint fits the budget. A dispatcher thread repeatedly takes
 *  every admissible request, stages them as one multi-segment round (one persistent, length-binned launch), and
 *  queues it on one of `max_inflight_rounds` streams, so a round formed from requests that arrived during earlier
 *  launches runs while those are still on the device. Completed rounds release their budget and wake waiters.
 *  Rounds draw their device buffers from a pool reserved once at init, so the loop never calls `hipMalloc` or
 *  `hipFree` (which waits for every stream of the device) while other rounds run.
 *  Results do not depend on which round a request joins: each pair is scored independently.
 *  No reference counterpart (the reference has no serving layer).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "szs_runtime.hpp"

using namespace szs_amd;

namespace {

struct request_t {
    szs_ticket_t ticket = 0;
    segment_input_t input;
    sz_sequence_u64tape_t firsts, seconds; /**< the caller's tapes (host or HBM), kept alive by the caller until wait */
    size_t footprint = 0;
};

struct round_t {
    batch_t *batch = nullptr;
    hipStream_t stream = nullptr;
    std::vector<szs_ticket_t> tickets;
    size_t footprint = 0;
};

enum ticket_state_t { ticket_pending_k = 0, ticket_done_k = 1, ticket_failed_k = 2 };

struct server_t {
    device_scope_t *scope = nullptr;
    int device = 0;
    size_t budget = 0;
    size_t in_use = 0;
    size_t max_inflight = 2;
    std::vector<hipStream_t> streams;
    std::vector<bool> stream_busy;
    std::mutex mutex;
    std::condition_variable wake_dispatcher, wake_waiters;
    std::deque<request_t> pending;
    std::vector<round_t> inflight;
    std::unordered_map<szs_ticket_t, int> states;
    std::unordered_map<szs_ticket_t, char const *> errors;
    szs_ticket_t next_ticket = 1;
    bool stopping = false;
    std::thread dispatcher;
    arena_t arena; /**< every round's device buffers, so rounds never allocate from the driver */

    void run();
};

/** Bytes of a tape's string data: `offsets[count]`, read from HBM for device-resident tapes. */
sz_status_t tape_bytes(sz_sequence_u64tape_t const *tape, bool device, uint64_t *bytes, char const **error_message) {
    *bytes = 0;
    if (!tape->count) return ok(error_message);
    if (!device) {
        *bytes = tape->offsets[tape->count];
        return ok(error_message);
    }
    return from_hip(hipMemcpy(bytes, tape->offsets + tape->count, 8, hipMemcpyDeviceToHost), error_message,
                    "Reading the tape length failed");
}

/**
 *  HBM a request occupies while staged: host tapes are copied (data and u64 offsets), device tapes are used in
 *  place; UTF-8 requests add the decoded runes; every request adds its tasks, bin bookkeeping and results.
 */
size_t footprint_of(engine_t const *engine, uint64_t count, uint64_t bytes, bool staged) {
    size_t total = count * (sizeof(task_t) + 4 * 3 + 8);
    if (staged) total += bytes + 16 * (count + 1);
    if (engine->kind == szs_engine_levenshtein_utf8_k) total += bytes * 4 + count * 8;
    return total + 64 * 1024;
}

void server_t::run() {
    if (hipSetDevice(device) != hipSuccess) (void)hipGetLastError();
    thread_arena() = arena.base ? &arena : nullptr;
    std::unique_lock<std::mutex> lock(mutex);
    while (true) {
        // Retire finished rounds.
        for (size_t r = 0; r < inflight.size();) {
            round_t &round = inflight[r];
            if (hipStreamQuery(round.stream) == hipErrorNotReady) {
                ++r;
                continue;
            }
            char const *error = nullptr;
            sz_status_t const status = batch_wait(round.batch, &error);
            for (szs_ticket_t t : round.tickets) {
                states[t] = status == sz_success_k ? ticket_done_k : ticket_failed_k;
                if (status != sz_success_k) errors[t] = error;
            }
            delete round.batch;
            in_use -= round.footprint;
            for (size_t s = 0; s < streams.size(); ++s)
                if (streams[s] == round.stream) stream_busy[s] = false;
            inflight.erase(inflight.begin() + r);
            wake_waiters.notify_all();
        }
        // Admit a new round if a stream is free and something fits.
        int free_stream = -1;
        for (size_t s = 0; s < streams.size(); ++s)
            if (!stream_busy[s]) {
                free_stream = (int)s;
                break;
            }
        std::vector<request_t> admitted;
        size_t admitted_bytes = 0;
        if (free_stream >= 0) {
            while (!pending.empty() && in_use + admitted_bytes + pending.front().footprint <= budget) {
                admitted_bytes += pending.front().footprint;
                admitted.push_back(std::move(pending.front()));
                pending.pop_front();
            }
        }
        if (!admitted.empty()) {
            in_use += admitted_bytes;
            stream_busy[free_stream] = true;
            hipStream_t const stream = streams[free_stream];
            lock.unlock();
            std::vector<segment_input_t> inputs;
            inputs.reserve(admitted.size());
            for (auto &request : admitted) {
                request.input.queries = &request.firsts;
                request.input.candidates = &request.seconds;
                inputs.push_back(request.input);
            }
            batch_t *batch = nullptr;
            char const *error = nullptr;
            sz_status_t status = batch_prepare(scope, inputs, szs_batch_persistent_k, &batch, &error, stream);
            if (status == sz_success_k) status = batch_launch(batch, &error);
            lock.lock();
            if (status != sz_success_k) {
                for (auto &request : admitted) states[request.ticket] = ticket_failed_k, errors[request.ticket] = error;
                delete batch;
                in_use -= admitted_bytes;
                stream_busy[free_stream] = false;
                wake_waiters.notify_all();
            }
            else {
                round_t round;
                round.batch = batch;
                round.stream = stream;
                round.footprint = admitted_bytes;
                for (auto &request : admitted) round.tickets.push_back(request.ticket);
                inflight.push_back(std::move(round));
            }
            continue;
        }
        if (stopping && pending.empty() && inflight.empty()) break;
        // Sleep until a submission arrives, or poll in-flight rounds shortly.
        if (inflight.empty()) wake_dispatcher.wait(lock, [&] { return stopping || !pending.empty(); });
        else wake_dispatcher.wait_for(lock, std::chrono::microseconds(50));
    }
}

} // namespace

extern "C" {

sz_status_t szs_server_init(szs_device_scope_t device, sz_size_t hbm_budget_bytes, sz_size_t max_inflight_rounds,
                            szs_server_t *out, char const **error_message) {
    if (!out || !device) return fail(sz_unexpected_dimensions_k, error_message, "Server arguments must not be null");
    device_scope_t *scope = reinterpret_cast<device_scope_t *>(device);
    int gpu = 0;
    hipStream_t unused = nullptr;
    sz_status_t const status = resolve_gpu(scope, &gpu, &unused, error_message);
    if (status != sz_success_k) return status;
    server_t *server = new (std::nothrow) server_t();
    if (!server) return fail(sz_bad_alloc_k, error_message, "Failed to allocate a server");
    server->scope = scope;
    server->device = gpu;
    server->budget = hbm_budget_bytes;
    server->max_inflight = max_inflight_rounds ? max_inflight_rounds : 2;
    // The pool holds the admitted footprints plus room for each in-flight round's boundary-row scratch; a round
    // that still does not fit falls back to plain allocations.
    size_t const margin = std::max<size_t>(size_t(256) << 20, hbm_budget_bytes / 8);
    hipError_t const pooled = server->arena.init(hbm_budget_bytes + margin);
    if (pooled != hipSuccess) {
        delete server;
        return from_hip(pooled, error_message, "Failed to reserve the server's HBM pool");
    }
    for (size_t s = 0; s < server->max_inflight; ++s) {
        hipStream_t stream = nullptr;
        hipError_t const error = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        if (error != hipSuccess) {
            for (auto st : server->streams) (void)hipStreamDestroy(st);
            delete server;
            return from_hip(error, error_message, "hipStreamCreate failed");
        }
        server->streams.push_back(stream);
        server->stream_busy.push_back(false);
    }
    server->dispatcher = std::thread([server] { server->run(); });
    *out = server;
    return ok(error_message);
}

sz_status_t szs_server_submit(szs_server_t handle, szs_pairs_segment_t const *request, szs_ticket_t *ticket,
                              char const **error_message) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server || !request || !ticket || !request->engine || !request->firsts || !request->seconds)
        return fail(sz_unexpected_dimensions_k, error_message, "Request fields must not be null");
    engine_t *engine = reinterpret_cast<engine_t *>(request->engine);
    if (engine->kind != (int)request->kind)
        return fail(sz_unexpected_dimensions_k, error_message, "Request kind does not match its engine");
    if (request->firsts->count != request->seconds->count)
        return fail(sz_unexpected_dimensions_k, error_message, "Pair tapes must hold equal counts");
    // Device-resident tapes are scored in place; anything else is staged into HBM with the round.
    bool const device_firsts = is_device_pointer(request->firsts->offsets);
    bool const device_seconds = is_device_pointer(request->seconds->offsets);
    bool const staged_firsts = !device_firsts || !is_device_pointer(request->firsts->data);
    bool const staged_seconds = !device_seconds || !is_device_pointer(request->seconds->data);
    uint64_t firsts_bytes = 0, seconds_bytes = 0;
    sz_status_t status = tape_bytes(request->firsts, device_firsts, &firsts_bytes, error_message);
    if (status == sz_success_k) status = tape_bytes(request->seconds, device_seconds, &seconds_bytes, error_message);
    if (status != sz_success_k) return status;
    request_t r;
    r.firsts = *request->firsts;
    r.seconds = *request->seconds;
    r.input.engine = engine;
    r.input.mode = segment_pairs_k;
    r.input.results = request->results;
    r.footprint = footprint_of(engine, request->firsts->count, firsts_bytes, staged_firsts) +
                  footprint_of(engine, 0, seconds_bytes, staged_seconds);
    if (r.footprint > server->budget)
        return fail(sz_bad_alloc_k, error_message, "Request exceeds the server's HBM budget");
    std::lock_guard<std::mutex> lock(server->mutex);
    r.ticket = server->next_ticket++;
    server->states[r.ticket] = ticket_pending_k;
    *ticket = r.ticket;
    server->pending.push_back(std::move(r));
    server->wake_dispatcher.notify_one();
    return ok(error_message);
}

sz_status_t szs_server_wait(szs_server_t handle, szs_ticket_t ticket, char const **error_message) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return fail(sz_unexpected_dimensions_k, error_message, "Server must be initialized");
    std::unique_lock<std::mutex> lock(server->mutex);
    auto found = server->states.find(ticket);
    if (found == server->states.end()) return fail(sz_unexpected_dimensions_k, error_message, "Unknown ticket");
    server->wake_waiters.wait(lock, [&] { return server->states[ticket] != ticket_pending_k; });
    int const state = server->states[ticket];
    char const *error = server->errors.count(ticket) ? server->errors[ticket] : nullptr;
    server->states.erase(ticket);
    server->errors.erase(ticket);
    if (state == ticket_failed_k) return fail(sz_status_unknown_k, error_message, error ? error : "Request failed");
    return ok(error_message);
}

int szs_server_poll(szs_server_t handle, szs_ticket_t ticket) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return 0;
    std::lock_guard<std::mutex> lock(server->mutex);
    auto found = server->states.find(ticket);
    return found != server->states.end() && found->second != ticket_pending_k;
}

sz_size_t szs_server_bytes_in_use(szs_server_t handle) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return 0;
    std::lock_guard<std::mutex> lock(server->mutex);
    return server->in_use;
}

void szs_server_free(szs_server_t handle) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return;
    {
        std::lock_guard<std::mutex> lock(server->mutex);
        server->stopping = true;
        server->wake_dispatcher.notify_all();
    }
    if (server->dispatcher.joinable()) server->dispatcher.join();
    for (auto stream : server->streams) (void)hipStreamDestroy(stream);
    delete server;
}

} // extern "C"
