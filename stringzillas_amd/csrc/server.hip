/**
 *  @file server.hip
 *  @brief Continuous-batching request loop over the batch engine.
 *
 *  Requests wait in a host FIFO until their HBM footprint fits the budget. A dispatcher thread repeatedly takes
 *  every admissible request, stages them as one multi-segment round (one materialization, one persistent,
 *  length-binned launch), and queues it on one of `max_inflight_rounds` streams, so a round formed from requests
 *  that arrived during earlier launches runs while those are still on the device. Completed rounds release their
 *  budget, stamp their tickets with the completion time and wake waiters.
 *  Rounds draw their device buffers from a pool reserved once at init, so the loop never calls `hipMalloc` or
 *  `hipFree` (which waits for every stream of the device) while other rounds run, and results come back through
 *  page-locked staging owned by each stream, so queuing a round never blocks on an earlier one.
 *  Results do not depend on which round a request joins: each pair is scored independently.
 *  No reference counterpart (the reference has no serving layer).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <ctime>
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
    size_t stream = 0; /**< index into the server's streams */
    std::vector<szs_ticket_t> tickets;
    size_t footprint = 0;
};

enum ticket_state_t { ticket_pending_k = 0, ticket_done_k = 1, ticket_failed_k = 2 };

struct ticket_t {
    int state = ticket_pending_k;
    sz_status_t status = sz_success_k;
    char const *error = nullptr;
    double completed = 0; /**< CLOCK_MONOTONIC seconds when the server saw the round finish */
};

/** Page-locked host memory for one stream's results, grown when a round needs more. */
struct pinned_t {
    char *data = nullptr;
    size_t capacity = 0;
    hipError_t reserve(size_t bytes) {
        if (bytes <= capacity) return hipSuccess;
        if (data) (void)hipHostFree(data);
        data = nullptr, capacity = 0;
        size_t const rounded = std::max<size_t>((bytes + (1u << 20) - 1) & ~size_t((1u << 20) - 1), size_t(8) << 20);
        void *pointer = nullptr;
        hipError_t const error = hipHostMalloc(&pointer, rounded, hipHostMallocDefault);
        if (error != hipSuccess) {
            (void)hipGetLastError();
            return error;
        }
        data = static_cast<char *>(pointer), capacity = rounded;
        return hipSuccess;
    }
    ~pinned_t() {
        if (data) (void)hipHostFree(data);
    }
};

double monotonic_seconds() {
    timespec now;
    clock_gettime(CLOCK_MONOTONIC, &now);
    return (double)now.tv_sec + (double)now.tv_nsec * 1e-9;
}

struct server_t {
    device_scope_t *scope = nullptr;
    int device = 0;
    size_t budget = 0;
    size_t in_use = 0;
    size_t max_inflight = 2;
    std::vector<hipStream_t> streams;
    std::vector<bool> stream_busy;
    std::vector<pinned_t> pinned;
    std::mutex mutex;
    std::condition_variable wake_dispatcher, wake_waiters;
    std::deque<request_t> pending;
    std::vector<round_t> inflight;
    std::unordered_map<szs_ticket_t, ticket_t> tickets;
    szs_ticket_t next_ticket = 1;
    bool stopping = false;
    std::thread dispatcher;
    arena_t arena; /**< every round's device buffers, so rounds never allocate from the driver */
    size_t margin = 0; /**< pool bytes beyond the budget: the rounds' stripe-boundary scratch */
    szs_server_stats_t stats {};

    void run();
    void retire(round_t &round, sz_status_t status, char const *error);
};

/** Bytes of a tape's string data: `offsets[count]`, read from HBM for device-resident tapes without a hint. */
sz_status_t tape_bytes(sz_sequence_u64tape_t const *tape, bool device, sz_size_t hint, uint64_t *bytes,
                       char const **error_message) {
    *bytes = 0;
    if (!tape->count) return ok(error_message);
    if (hint) {
        *bytes = hint;
        return ok(error_message);
    }
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

/** Marks a round's tickets finished (called with the mutex held) and releases its budget and stream. */
void server_t::retire(round_t &round, sz_status_t status, char const *error) {
    double const now = monotonic_seconds();
    for (size_t i = 0; i < round.tickets.size(); ++i) {
        ticket_t &ticket = tickets[round.tickets[i]];
        ticket.completed = now;
        uint32_t const flags = status == sz_success_k && round.batch && i < round.batch->overflowed.size()
                                   ? round.batch->overflowed[i]
                                   : 0u;
        if (status != sz_success_k || flags) {
            ticket.state = ticket_failed_k;
            if (status != sz_success_k) ticket.status = status, ticket.error = error;
            else if (flags & segment_bad_tape_k)
                ticket.status = sz_unexpected_dimensions_k,
                ticket.error = "A tape's offsets decrease or run past its byte count";
            else ticket.status = sz_overflow_risk_k, ticket.error = "A string holds 2^32 or more symbols";
        }
        else ticket.state = ticket_done_k;
    }
    if (round.batch && status == sz_success_k) {
        stats.cells += round.batch->cells;
        stats.device_seconds += round.batch->last_kernel_ms * 1e-3;
    }
    delete round.batch;
    round.batch = nullptr;
    in_use -= round.footprint;
    stream_busy[round.stream] = false;
    wake_waiters.notify_all();
}

void server_t::run() {
    if (hipSetDevice(device) != hipSuccess) (void)hipGetLastError();
    thread_arena() = arena.base ? &arena : nullptr;
    std::unique_lock<std::mutex> lock(mutex);
    while (true) {
        // Retire finished rounds.
        for (size_t r = 0; r < inflight.size();) {
            round_t &round = inflight[r];
            if (hipStreamQuery(streams[round.stream]) == hipErrorNotReady) {
                ++r;
                continue;
            }
            char const *error = nullptr;
            sz_status_t const status = batch_wait(round.batch, &error);
            retire(round, status, error);
            inflight.erase(inflight.begin() + r);
        }
        // Admit a new round if a stream is free and something fits.
        size_t free_stream = streams.size();
        for (size_t s = 0; s < streams.size(); ++s)
            if (!stream_busy[s]) {
                free_stream = s;
                break;
            }
        std::vector<request_t> admitted;
        size_t admitted_bytes = 0, admitted_results = 0;
        if (free_stream < streams.size()) {
            while (!pending.empty() && in_use + admitted_bytes + pending.front().footprint <= budget) {
                admitted_bytes += pending.front().footprint;
                admitted_results += pending.front().firsts.count * 8;
                admitted.push_back(std::move(pending.front()));
                pending.pop_front();
            }
        }
        if (!admitted.empty()) {
            in_use += admitted_bytes;
            stream_busy[free_stream] = true;
            round_t round;
            round.stream = free_stream;
            round.footprint = admitted_bytes;
            for (auto &request : admitted) round.tickets.push_back(request.ticket);
            lock.unlock();
            double const started = monotonic_seconds();
            std::vector<segment_input_t> inputs;
            inputs.reserve(admitted.size());
            for (auto &request : admitted) {
                request.input.queries = &request.firsts;
                request.input.candidates = &request.seconds;
                inputs.push_back(request.input);
            }
            batch_options_t options;
            options.stream = streams[free_stream];
            options.defer_overflow = true;
            // A round's latency is its longest item: pairs taller than 16 stripes of a one-wave shape (8,192 rows of
            // packed NW, 65,536 of Myers words) take the whole workgroup.
            options.latency_stripes = 16;
            // Stripe-boundary scratch stays within this round's share of the pool's margin over the budget.
            options.max_scratch_bytes = std::max<size_t>(size_t(16) << 20, margin / streams.size());
            if (pinned[free_stream].reserve(admitted_results) == hipSuccess) {
                options.pinned_results = pinned[free_stream].data;
                options.pinned_capacity = pinned[free_stream].capacity;
            }
            char const *error = nullptr;
            sz_status_t const status = guarded(&error, [&]() {
                sz_status_t prepared = batch_prepare(scope, inputs, szs_batch_persistent_k, &round.batch, &error,
                                                     options);
                return prepared == sz_success_k ? batch_launch(round.batch, &error, true) : prepared;
            });
            double const prepared = monotonic_seconds();
            lock.lock();
            stats.rounds += 1;
            stats.requests += admitted.size();
            stats.prepare_seconds += prepared - started;
            stats.largest_round = std::max<sz_u64_t>(stats.largest_round, admitted.size());
            if (status != sz_success_k) retire(round, status, error);
            else inflight.push_back(std::move(round));
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
    server->margin = margin;
    hipError_t const pooled = server->arena.init(hbm_budget_bytes + margin);
    if (pooled != hipSuccess) {
        delete server;
        return from_hip(pooled, error_message, "Failed to reserve the server's HBM pool");
    }
    server->pinned = std::vector<pinned_t>(server->max_inflight);
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
    sz_status_t status =
        tape_bytes(request->firsts, device_firsts, request->firsts_bytes, &firsts_bytes, error_message);
    if (status == sz_success_k)
        status = tape_bytes(request->seconds, device_seconds, request->seconds_bytes, &seconds_bytes, error_message);
    if (status != sz_success_k) return status;
    request_t r;
    r.firsts = *request->firsts;
    r.seconds = *request->seconds;
    r.input.queries = &r.firsts;
    r.input.candidates = &r.seconds;
    status = check_host_offsets(r.input, error_message); // rejected now rather than failing its round later
    if (status != sz_success_k) return status;
    r.input.engine = engine;
    r.input.mode = segment_pairs_k;
    r.input.results = request->results;
    r.input.queries_bytes = firsts_bytes;
    r.input.candidates_bytes = seconds_bytes;
    r.footprint = footprint_of(engine, request->firsts->count, firsts_bytes, staged_firsts) +
                  footprint_of(engine, 0, seconds_bytes, staged_seconds);
    if (r.footprint > server->budget)
        return fail(sz_bad_alloc_k, error_message, "Request exceeds the server's HBM budget");
    return guarded(error_message, [&]() {
        std::lock_guard<std::mutex> lock(server->mutex);
        server->tickets[server->next_ticket] = ticket_t();
        r.ticket = server->next_ticket++;
        server->pending.push_back(std::move(r));
        *ticket = server->pending.back().ticket;
        server->wake_dispatcher.notify_one();
        return ok(error_message);
    });
}

sz_status_t szs_server_wait_timed(szs_server_t handle, szs_ticket_t ticket, double *completed_seconds,
                                  char const **error_message) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return fail(sz_unexpected_dimensions_k, error_message, "Server must be initialized");
    std::unique_lock<std::mutex> lock(server->mutex);
    auto found = server->tickets.find(ticket);
    if (found == server->tickets.end()) return fail(sz_unexpected_dimensions_k, error_message, "Unknown ticket");
    server->wake_waiters.wait(lock, [&] { return server->tickets[ticket].state != ticket_pending_k; });
    ticket_t const finished = server->tickets[ticket];
    server->tickets.erase(ticket);
    if (completed_seconds) *completed_seconds = finished.completed;
    if (finished.state == ticket_failed_k)
        return fail(finished.status == sz_success_k ? sz_status_unknown_k : finished.status, error_message,
                    finished.error ? finished.error : "Request failed");
    return ok(error_message);
}

sz_status_t szs_server_wait(szs_server_t handle, szs_ticket_t ticket, char const **error_message) {
    return szs_server_wait_timed(handle, ticket, nullptr, error_message);
}

int szs_server_poll(szs_server_t handle, szs_ticket_t ticket) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return 0;
    std::lock_guard<std::mutex> lock(server->mutex);
    auto found = server->tickets.find(ticket);
    return found != server->tickets.end() && found->second.state != ticket_pending_k;
}

sz_size_t szs_server_bytes_in_use(szs_server_t handle) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server) return 0;
    std::lock_guard<std::mutex> lock(server->mutex);
    return server->in_use;
}

sz_status_t szs_server_get_stats(szs_server_t handle, szs_server_stats_t *stats) {
    server_t *server = reinterpret_cast<server_t *>(handle);
    if (!server || !stats) return sz_unexpected_dimensions_k;
    std::lock_guard<std::mutex> lock(server->mutex);
    *stats = server->stats;
    stats->pending = server->pending.size();
    stats->inflight = server->inflight.size();
    return sz_success_k;
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
