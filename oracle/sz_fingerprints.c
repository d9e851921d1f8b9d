/**
 *  @file oracle/sz_fingerprints.c
 *  @brief TEST INFRASTRUCTURE ONLY: CPU oracle restating the reference's rolling MinHash / Count-Min fingerprints.
 *
 *  Only the test-suite, the smoke check and the CPU-baseline leg of the benchmark use this file, as the
 *  checker. It restates, step for step, the serial `floating_rolling_hashers<sz_cap_serial_k, 64>` of
 *  include/stringzillas/fingerprints/serial.hpp:876-1054 (f64 state, `std::fma` push and roll, Barrett-style
 *  reduction with both fix-ups) and the dimension layout that `szs_fingerprints_init` builds
 *  (c/stringzillas/fingerprints.cuh:27-140). Every intermediate stays an integer below 2^52, so the f64 steps
 *  compute exact residues; the reference's fallback `basic_rolling_hashers` (serial.hpp:497-870, a fused
 *  one-reduction roll) therefore yields the same numbers, which the tests check against both compiled
 *  reference paths.
 */
#include "sz_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <string.h>

/** `splitmix64`, serial.hpp:38-44: the first output of Vigna's generator seeded with @p state. */
uint64_t szo_splitmix64(uint64_t state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/**
 *  One dimension's parameters: `floating_rolling_hasher<f64_t>(window, alphabet, dim, seed)`, serial.hpp:370-431.
 *  The multiplier is `256 + splitmix64(seed + dim) % 384`, the modulo `base - splitmix64(splitmix64(seed + dim))
 *  % 2^20`; the negative discarding multiplier is `-(multiplier^(window - 1) mod modulo)`, built by the same
 *  `fmod` loop; the discarding multiplier is its product with the multiplier, reduced into [0, modulo).
 */
static void szo_seed_hasher(uint64_t window, uint64_t dim, uint64_t seed, szo_hasher_t *hasher) {
    uint64_t const mixed = szo_splitmix64(seed + dim);
    double const multiplier = (double)(256ull + mixed % 384ull);
    double const modulo = (double)(4503599626977ull /* default_modulo_base_k, serial.hpp:369 */ - szo_splitmix64(mixed) % (1ull << 20));
    double negative = 1.0;
    for (uint64_t i = 0; i + 1 < window; ++i) negative = fmod(negative * multiplier, modulo);
    negative = -negative;
    double discarding = fmod(negative * multiplier, modulo) + modulo;
    if (discarding >= modulo) discarding -= modulo;
    hasher->multiplier = multiplier;
    hasher->modulo = modulo;
    hasher->inverse_modulo = 1.0 / modulo;
    hasher->negative_discarding_multiplier = negative;
    hasher->discarding_multiplier = discarding;
    hasher->window_width = window;
}

/**
 *  Per-dimension hashers in output order, as `szs_fingerprints_init` lays them out
 *  (c/stringzillas/fingerprints.cuh:39-47, 117-140): the default widths are {3, 4, 5, 7, 9, 11, 15, 31}; when
 *  every width gets the same number of dimensions and that number is a multiple of 64, dimension `d` belongs to
 *  the 64-wide slice `d / 64`, whose width is `widths[(d / 64) % count]` (the sliced engines); otherwise
 *  dimension `d` uses `widths[d % count]` (the per-dimension fallback). Either way dimension `d` is seeded
 *  with `seed + d` (serial.hpp:544-553 and 913-925).
 *  @return 1 when the sliced layout applies, 0 for the fallback.
 */
int szo_fingerprint_hashers(uint64_t dimensions, uint64_t const *widths, uint64_t widths_count, uint64_t seed,
                            szo_hasher_t *hashers) {
    static uint64_t const default_widths[] = {3, 4, 5, 7, 9, 11, 15, 31};
    if (!widths || widths_count == 0) widths = default_widths, widths_count = 8;
    uint64_t const per_width_min = dimensions / widths_count;
    uint64_t const per_width_max = (dimensions + widths_count - 1) / widths_count;
    int const sliced = per_width_min == per_width_max && per_width_min % 64 == 0;
    for (uint64_t d = 0; d < dimensions; ++d) {
        uint64_t const window = sliced ? widths[(d / 64) % widths_count] : widths[d % widths_count];
        szo_seed_hasher(window, d, seed, &hashers[d]);
    }
    return sliced;
}

/** `barrett_mod`, serial.hpp:1046-1054: floor-quotient reduction with both fix-ups. */
static inline double szo_barrett(double x, double modulo, double inverse) {
    double const q = floor(x * inverse);
    double result = x - q * modulo;
    result += modulo * (result < 0.0);
    result -= modulo * (result >= modulo);
    return result;
}

/**
 *  Fingerprint of one text: `floating_rolling_hashers::fingerprint` + `fingerprint_chunk`, serial.hpp:934-1032,
 *  applied per dimension. Terms are `byte + 1`. The first `window` bytes are pushed (`fma(state, multiplier,
 *  term)` then reduced); the state after them is the first window hash, which starts the minimum with count 1;
 *  every later byte first discards the oldest term (`fma(negative_discarding, old, state)`, reduced), then pushes
 *  the new one. The count is reset on a strictly smaller hash and incremented on every hash equal to the new
 *  minimum. Output: the minimum's low 32 bits and its count, or 0xFFFFFFFF and 0 for a text shorter than the
 *  window (serial.hpp:937-941, 1021-1028).
 */
void szo_fingerprint(szo_hasher_t const *hashers, uint64_t dimensions, uint8_t const *text, size_t length,
                     uint32_t *min_hashes, uint32_t *min_counts) {
    for (uint64_t d = 0; d < dimensions; ++d) {
        szo_hasher_t const *h = &hashers[d];
        size_t const window = (size_t)h->window_width;
        if (length < window) {
            min_hashes[d] = 0xFFFFFFFFu;
            min_counts[d] = 0;
            continue;
        }
        double state = 0.0, minimum = DBL_MAX;
        uint32_t count = 0;
        size_t at = 0;
        for (; at < window; ++at) {
            state = fma(state, h->multiplier, (double)text[at] + 1.0);
            state = szo_barrett(state, h->modulo, h->inverse_modulo);
        }
        minimum = state < minimum ? state : minimum;
        count = 1;
        for (; at < length; ++at) {
            double const new_term = (double)text[at] + 1.0;
            double const old_term = (double)text[at - window] + 1.0;
            state = fma(h->negative_discarding_multiplier, old_term, state);
            state = szo_barrett(state, h->modulo, h->inverse_modulo);
            state = fma(state, h->multiplier, new_term);
            state = szo_barrett(state, h->modulo, h->inverse_modulo);
            count *= state >= minimum;
            count += state <= minimum;
            minimum = state < minimum ? state : minimum;
        }
        min_hashes[d] = (uint32_t)((uint64_t)minimum & 0xFFFFFFFFull);
        min_counts[d] = count;
    }
}

typedef struct {
    szo_hasher_t const *hashers;
    uint64_t dimensions;
    uint8_t const *data;
    uint64_t const *offsets;
    uint64_t count;
    uint32_t *min_hashes, *min_counts;
    uint64_t hashes_stride, counts_stride;
    uint64_t begin, step;
} szo_fingerprint_job_t;

static void *szo_fingerprint_worker(void *arg) {
    szo_fingerprint_job_t const *job = (szo_fingerprint_job_t const *)arg;
    for (uint64_t i = job->begin; i < job->count; i += job->step)
        szo_fingerprint(job->hashers, job->dimensions, job->data + job->offsets[i],
                        (size_t)(job->offsets[i + 1] - job->offsets[i]),
                        (uint32_t *)((char *)job->min_hashes + i * job->hashes_stride),
                        (uint32_t *)((char *)job->min_counts + i * job->counts_stride));
    return NULL;
}

int szo_fingerprints(szo_hasher_t const *hashers, uint64_t dimensions, uint8_t const *data, uint64_t const *offsets,
                     uint64_t count, uint32_t *min_hashes, uint64_t hashes_stride, uint32_t *min_counts,
                     uint64_t counts_stride, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    szo_fingerprint_job_t jobs[256];
    pthread_t handles[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t].hashers = hashers, jobs[t].dimensions = dimensions, jobs[t].data = data, jobs[t].offsets = offsets;
        jobs[t].count = count, jobs[t].min_hashes = min_hashes, jobs[t].min_counts = min_counts;
        jobs[t].hashes_stride = hashes_stride, jobs[t].counts_stride = counts_stride;
        jobs[t].begin = (uint64_t)t, jobs[t].step = (uint64_t)threads;
    }
    if (threads == 1) {
        szo_fingerprint_worker(&jobs[0]);
        return 0;
    }
    for (int t = 0; t < threads; ++t) pthread_create(&handles[t], NULL, szo_fingerprint_worker, &jobs[t]);
    for (int t = 0; t < threads; ++t) pthread_join(handles[t], NULL);
    return 0;
}
