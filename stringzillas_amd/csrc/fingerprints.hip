/**
 *  @file fingerprints.hip
 *  @brief Rolling MinHash + Count-Min fingerprints on gfx950, behind the reference's `szs_fingerprints_*` ABI
 *         (include/stringzillas/stringzillas.h:436-536, c/stringzillas/fingerprints.cuh:27-190).
 *
 *  Semantics (include/stringzillas/fingerprints/serial.hpp:364-431, 934-1032): dimension `d` hashes every window
 *  of `w` bytes of a text as the Rabin-Karp polynomial `sum((byte + 1) * multiplier^(w - 1 - k)) mod modulo`,
 *  with `multiplier = 256 + splitmix64(seed + d) % 384` and `modulo = 4503599626977 - splitmix64(splitmix64(seed
 *  + d)) % 2^20`; it keeps the smallest hash and how many windows reach it. The reference evaluates this in f64
 *  with every intermediate an integer below 2^52, so its results are the exact residues; this file computes the
 *  same residues.
 *
 *  Device layout. Dimensions are grouped by window width (the reference's layout gives dimension `d` the width
 *  `widths[(d / 64) % count]` or `widths[d % count]`), and each group is cut into runs of `64 * slots` dimensions:
 *  one wavefront hashes one text for one run, lane `l` owning dimensions `l, l + 64, ...` of the run (its
 *  "slots"). The text and the window width are then uniform across the wave: each 256-byte chunk is read once
 *  (4 bytes per lane), and every byte reaches all lanes through `v_readlane` into a scalar register, so the
 *  per-byte cost is the per-dimension arithmetic alone:
 *      x = fma(state, multiplier, new + 1); x = fma(discard, old + 1, x)          (exact: x < 2^52)
 *      r = x - floor(x * inverse_low) * modulo;  state = r - floor((r + 1) * inverse_low) * modulo
 *  where `discard = (-multiplier^w) mod modulo` and `inverse_low` is 1/modulo rounded down by two ulps: the first
 *  quotient is never above the true one and at most one below it, and the second is exactly the missing one
 *  (the proof is at `reduce`). The minimum and count update follow serial.hpp:1012-1016.
 *
 *  Texts with few dimensions per width (the reference's per-dimension layout, e.g. 64 dimensions over 8 widths)
 *  go to a second kernel instead, where each lane walks its own text and updates 8 dimensions of one width, whose
 *  hashers sit in scalar registers.
 *
 *  Work units: a text of at most `segment` bytes is one unit; a longer one is cut into units of `segment` window
 *  starts. Each such unit pushes its own first window, its (minimum, count) pairs go to a device scratch array,
 *  and a merge kernel folds them in order (equal minima add their counts), as the reference's multi-threaded
 *  path does (serial.hpp:817-860).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "szs_runtime.hpp"

using namespace szs_amd;

namespace {

constexpr uint32_t wave_k = 64;
constexpr uint32_t waves_per_block_k = 4;
constexpr uint64_t wave_segment_bytes_k = 16384; // window starts per segment of a long text, wave kernel
constexpr uint32_t chunk_bytes_k = wave_k * 4;
constexpr uint32_t no_dimension_k = 0xFFFFFFFFu;
constexpr uint64_t max_blocks_k = 1u << 18;
constexpr int lane_slots_k = 8;                  // dimensions per lane in the lane kernel
constexpr uint64_t direct_k = ~0ull;             // a unit covering a whole text writes its rows directly

/** One lane-slot: the hasher of one dimension (or a harmless dummy where a run or group is shorter). */
struct fp_hasher_t {
    double multiplier, modulo, inverse, discard;
};

/** Wave kernel: a run of up to `64 * slots` dimensions sharing one window width; lane-slots start at `first`. */
struct fp_run_t {
    uint64_t width;
    uint32_t first, reserved;
};

/** A text, or one segment of a long text: window starts `[start, start + segment)`. */
struct fp_unit_t {
    uint64_t text, start;
    uint64_t partial; /**< row of the partial (minimum, count) arrays, or `direct_k` */
};

struct fp_long_text_t {
    uint64_t text, first_partial, partials;
};

struct fp_launch_t {
    uint8_t const *data;
    uint64_t const *offsets;
    fp_hasher_t const *hashers; /**< wave kernel: [runs][slots][64]; lane kernel: [groups][8] */
    uint32_t const *dims;       /**< same shape: output dimension of each lane-slot */
    fp_run_t const *runs;       /**< wave kernel: runs; lane kernel: one entry per group (its width) */
    uint32_t runs_count, dimensions;
    fp_unit_t const *units;     /**< null: unit `u` is text `u`, whole */
    uint64_t units_count;
    uint64_t segment;           /**< window starts per unit */
    uint32_t *hashes, *counts;  /**< rows at these byte strides */
    uint64_t hashes_stride, counts_stride;
    double *partial_minimums;   /**< [partial][position]: position = lane-slot index in `hashers` */
    uint32_t *partial_counts;
    uint32_t positions;         /**< lane-slots of the layout */
};

struct fp_merge_t {
    fp_long_text_t const *texts;
    uint64_t texts_count;
    uint32_t positions;
    uint32_t const *dims; /**< output dimension of each position (or none) */
    double const *partial_minimums;
    uint32_t const *partial_counts;
    uint32_t *hashes, *counts;
    uint64_t hashes_stride, counts_stride;
};

__device__ __forceinline__ uint32_t low_word(double value) { return (uint32_t)(uint64_t)value; }

/**
 *  Canonical residue of an integer-valued `0 <= x < 2^52`, in six VALU operations and no compare.
 *  `inverse_low` lies 2^-53 to 2^-51 (relative) below 1/modulo. Then `x * inverse_low` rounds strictly below
 *  x/modulo and within 2^-40 of it, so the first quotient is the true one or one less and `r` lands in
 *  [0, 2 * modulo). For an integer `r` there, `(r + 1) * inverse_low` (one fma, one rounding) is below 1 exactly
 *  when `r < modulo` and below 2 always, so the second quotient is exactly `r >= modulo`. Every product
 *  `q * modulo` and difference is an integer below 2^53, hence exact.
 */
__device__ __forceinline__ double reduce(double x, double modulo, double inverse_low) {
    double const r = fma(-floor(x * inverse_low), modulo, x);
    return fma(-floor(fma(r, inverse_low, inverse_low)), modulo, r);
}

/** A value every lane holds alike, moved into scalar registers. */
__device__ __forceinline__ double uniform(double value) {
    uint64_t const bits = __builtin_bit_cast(uint64_t, value);
    uint32_t const low = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)bits);
    uint32_t const high = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(bits >> 32));
    return __builtin_bit_cast(double, ((uint64_t)high << 32) | low);
}

template <int slots_>
struct fp_lanes_t {
    double multiplier[slots_], modulo[slots_], inverse[slots_], discard[slots_];
    double state[slots_], minimum[slots_];
    uint32_t count[slots_];

    __device__ __forceinline__ void reset() {
#pragma unroll
        for (int s = 0; s < slots_; ++s) state[s] = 0.0, minimum[s] = DBL_MAX, count[s] = 0;
    }

    __device__ __forceinline__ void push(double term) {
#pragma unroll
        for (int s = 0; s < slots_; ++s) state[s] = reduce(fma(state[s], multiplier[s], term), modulo[s], inverse[s]);
    }

    __device__ __forceinline__ void first_window() {
#pragma unroll
        for (int s = 0; s < slots_; ++s) minimum[s] = state[s], count[s] = 1;
    }

    /**
     *  One window step of every slot. A new minimum or a tie gets rarer with every window (about 1/j at the j-th),
     *  so the hot path only compares; each slot's minimum and count update runs behind its own branch, taken
     *  only when some lane of the wave hits in that slot.
     */
    __device__ __forceinline__ void roll(double fresh, double stale) {
#pragma unroll
        for (int s = 0; s < slots_; ++s) { // all slots first, so their independent chains interleave
            double const x = fma(discard[s], stale, fma(state[s], multiplier[s], fresh));
            state[s] = reduce(x, modulo[s], inverse[s]);
        }
        // Wave-wide hit masks first (scalar registers): the chains above then stay in one block and interleave.
        uint64_t hits[slots_];
#pragma unroll
        for (int s = 0; s < slots_; ++s) hits[s] = __ballot(state[s] <= minimum[s]);
#pragma unroll
        for (int s = 0; s < slots_; ++s) {
            if (__builtin_expect(hits[s] != 0, 0) && state[s] <= minimum[s]) {
                count[s] = state[s] < minimum[s] ? 1u : count[s] + 1u;
                minimum[s] = state[s];
            }
        }
    }
};

/**
 *  Writes one lane-slot's result: to the caller's rows for a whole text, else to the partial arrays, where the
 *  lane-slots of one unit are adjacent (a lane's slots, or a slot's lanes, fill whole sectors).
 */
__device__ __forceinline__ void store(fp_launch_t const &a, uint64_t text, uint64_t partial, uint32_t position,
                                      uint32_t dim, bool any, double minimum, uint32_t count) {
    if (partial != direct_k) {
        a.partial_minimums[partial * a.positions + position] = any ? minimum : DBL_MAX;
        a.partial_counts[partial * a.positions + position] = any ? count : 0u;
        return;
    }
    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(a.hashes) + text * a.hashes_stride + dim * 4ull) =
        any ? low_word(minimum) : 0xFFFFFFFFu;
    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(a.counts) + text * a.counts_stride + dim * 4ull) =
        any ? count : 0u;
}

__device__ __forceinline__ fp_unit_t unit_at(fp_launch_t const &a, uint64_t index) {
    if (a.units) return a.units[index];
    return fp_unit_t {index, 0, direct_k};
}

/** Term of byte `index` (0..3) of a packed scalar word: `byte + 1` as f64. */
__device__ __forceinline__ double term_of(uint32_t word, uint32_t index) {
    return (double)(((word >> (8u * index)) & 0xFFu) + 1u);
}

/** Packs bytes `text[at + 4 * lane + b]` for `4 * lane + b < steps` (0 beyond). */
__device__ __forceinline__ uint32_t load_quad(uint8_t const *text, uint64_t at, uint32_t lane, uint32_t steps) {
    uint32_t word = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
        uint32_t const offset = lane * 4 + b;
        if (offset < steps) word |= (uint32_t)text[at + offset] << (8u * b);
    }
    return word;
}

/**
 *  Wave kernel: one wave per (unit, run of `64 * slots` dimensions of one width). The text is uniform across the
 *  wave: each 256-byte chunk is loaded once, 4 bytes per lane, and every byte reaches all lanes through
 *  `v_readlane`. Window starts `[start, end)`: pushes bytes `[start, start + w)`, then rolls bytes
 *  `[start + w, end + w - 1)` against the bytes `w` positions back.
 */
template <int slots_>
__global__ __launch_bounds__(wave_k *waves_per_block_k) void fingerprint_wave_kernel(fp_launch_t a) {
    uint32_t const lane = threadIdx.x % wave_k;
    // The wave index is uniform; saying so keeps the task, the text and every loop bound in scalar registers.
    uint32_t const wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / wave_k));
    uint64_t const waves = (uint64_t)gridDim.x * waves_per_block_k;
    uint64_t const tasks = a.units_count * a.runs_count;
    for (uint64_t task = (uint64_t)blockIdx.x * waves_per_block_k + wave; task < tasks; task += waves) {
        fp_unit_t const unit = unit_at(a, task / a.runs_count);
        fp_run_t const run = a.runs[task % a.runs_count];
        uint64_t const begin = a.offsets[unit.text];
        uint64_t const length = a.offsets[unit.text + 1] - begin;
        uint64_t const width = run.width;
        uint64_t const windows = length >= width ? length - width + 1 : 0;
        uint64_t const end = std::min<uint64_t>(windows, unit.start + a.segment);
        bool const any = end > unit.start;

        fp_lanes_t<slots_> h;
        uint32_t dim[slots_];
#pragma unroll
        for (int s = 0; s < slots_; ++s) {
            uint32_t const slot = run.first + (uint32_t)s * wave_k + lane;
            fp_hasher_t const hasher = a.hashers[slot];
            dim[s] = a.dims[slot];
            h.multiplier[s] = hasher.multiplier, h.modulo[s] = hasher.modulo;
            h.inverse[s] = hasher.inverse, h.discard[s] = hasher.discard;
        }
        h.reset();
        if (any) {
            uint8_t const *text = a.data + begin;
            uint64_t const start = unit.start;
            for (uint64_t chunk = start; chunk < start + width; chunk += chunk_bytes_k) { // the first window
                uint32_t const steps = (uint32_t)std::min<uint64_t>(chunk_bytes_k, start + width - chunk);
                uint32_t const fresh = load_quad(text, chunk, lane, steps);
                for (uint32_t k = 0; k < steps; ++k) {
                    uint32_t const word = (uint32_t)__builtin_amdgcn_readlane((int)fresh, (int)(k / 4));
                    h.push(term_of(word, k % 4));
                }
            }
            h.first_window();
            uint64_t const roll_end = end + width - 1;
            for (uint64_t chunk = start + width; chunk < roll_end; chunk += chunk_bytes_k) {
                uint32_t const steps = (uint32_t)std::min<uint64_t>(chunk_bytes_k, roll_end - chunk);
                uint32_t const fresh = load_quad(text, chunk, lane, steps);
                uint32_t const stale = load_quad(text, chunk - width, lane, steps);
                uint32_t const quads = steps / 4;
                for (uint32_t k = 0; k < quads; ++k) {
                    uint32_t const f = (uint32_t)__builtin_amdgcn_readlane((int)fresh, (int)k);
                    uint32_t const o = (uint32_t)__builtin_amdgcn_readlane((int)stale, (int)k);
                    h.roll(term_of(f, 0), term_of(o, 0));
                    h.roll(term_of(f, 1), term_of(o, 1));
                    h.roll(term_of(f, 2), term_of(o, 2));
                    h.roll(term_of(f, 3), term_of(o, 3));
                }
                for (uint32_t k = quads * 4; k < steps; ++k) {
                    uint32_t const f = (uint32_t)__builtin_amdgcn_readlane((int)fresh, (int)(k / 4));
                    uint32_t const o = (uint32_t)__builtin_amdgcn_readlane((int)stale, (int)(k / 4));
                    h.roll(term_of(f, k % 4), term_of(o, k % 4));
                }
            }
        }
#pragma unroll
        for (int s = 0; s < slots_; ++s)
            if (dim[s] != no_dimension_k)
                store(a, unit.text, unit.partial, run.first + (uint32_t)s * wave_k + lane, dim[s], any, h.minimum[s],
                      h.count[s]);
    }
}

/** `byte + 1` as f64, for a byte held per lane. */
__device__ __forceinline__ double lane_term(uint8_t byte) { return (double)((uint32_t)byte + 1u); }

/**
 *  Lane kernel: one wave per (64 units, group of up to 8 dimensions of one width). Each lane walks its own unit
 *  and updates 8 dimensions per byte; the group's hashers are uniform across the wave and sit in scalar
 *  registers. Suits fingerprints whose widths hold fewer than 64 dimensions each, where the wave kernel would
 *  leave lanes empty.
 */
__global__ __launch_bounds__(wave_k *waves_per_block_k) void fingerprint_lane_kernel(fp_launch_t a) {
    uint32_t const lane = threadIdx.x % wave_k;
    uint32_t const wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / wave_k));
    uint64_t const waves = (uint64_t)gridDim.x * waves_per_block_k;
    uint64_t const blocks = (a.units_count + wave_k - 1) / wave_k;
    uint64_t const tasks = blocks * a.runs_count;
    for (uint64_t task = (uint64_t)blockIdx.x * waves_per_block_k + wave; task < tasks; task += waves) {
        uint32_t const group = (uint32_t)(task % a.runs_count);
        uint64_t const index = (task / a.runs_count) * wave_k + lane;
        uint64_t const width = a.runs[group].width;
        fp_lanes_t<lane_slots_k> h;
        uint32_t dim[lane_slots_k];
#pragma unroll
        for (int s = 0; s < lane_slots_k; ++s) {
            fp_hasher_t const hasher = a.hashers[group * lane_slots_k + s];
            dim[s] = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.dims[group * lane_slots_k + s]);
            h.multiplier[s] = uniform(hasher.multiplier), h.modulo[s] = uniform(hasher.modulo);
            h.inverse[s] = uniform(hasher.inverse), h.discard[s] = uniform(hasher.discard);
        }
        h.reset();
        if (index >= a.units_count) continue;
        fp_unit_t const unit = unit_at(a, index);
        uint64_t const begin = a.offsets[unit.text];
        uint64_t const length = a.offsets[unit.text + 1] - begin;
        uint64_t const windows = length >= width ? length - width + 1 : 0;
        uint64_t const end = std::min<uint64_t>(windows, unit.start + a.segment);
        bool const any = end > unit.start;
        if (any) {
            uint8_t const *text = a.data + begin;
            for (uint64_t k = 0; k < width; ++k) h.push(lane_term(text[unit.start + k]));
            h.first_window();
            uint64_t at = unit.start + width;
            uint64_t const stop = end + width - 1;
            for (; at + 4 <= stop; at += 4) {
                uint8_t fresh[4], stale[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) fresh[b] = text[at + b], stale[b] = text[at + b - width];
#pragma unroll
                for (int b = 0; b < 4; ++b) h.roll(lane_term(fresh[b]), lane_term(stale[b]));
            }
            for (; at < stop; ++at) h.roll(lane_term(text[at]), lane_term(text[at - width]));
        }
#pragma unroll
        for (int s = 0; s < lane_slots_k; ++s)
            if (dim[s] != no_dimension_k)
                store(a, unit.text, unit.partial, group * lane_slots_k + s, dim[s], any, h.minimum[s], h.count[s]);
    }
}

/** One thread per (long text, lane-slot position): folds the segments, equal minima adding their counts. */
__global__ __launch_bounds__(256) void fingerprint_merge_kernel(fp_merge_t m) {
    uint64_t const index = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= m.texts_count * m.positions) return;
    fp_long_text_t const text = m.texts[index / m.positions];
    uint32_t const position = (uint32_t)(index % m.positions);
    uint32_t const dim = m.dims[position];
    if (dim == no_dimension_k) return;
    double best = DBL_MAX;
    uint32_t count = 0;
    for (uint64_t p = text.first_partial; p < text.first_partial + text.partials; ++p) {
        double const value = m.partial_minimums[p * m.positions + position];
        uint32_t const found = m.partial_counts[p * m.positions + position];
        if (value < best) best = value, count = found;
        else if (value == best) count += found;
    }
    bool const any = best != DBL_MAX;
    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(m.hashes) + text.text * m.hashes_stride + dim * 4ull) =
        any ? low_word(best) : 0xFFFFFFFFu;
    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(m.counts) + text.text * m.counts_stride + dim * 4ull) =
        any ? count : 0u;
}

#pragma region Host side

/** `splitmix64`, include/stringzillas/fingerprints/serial.hpp:38-44. */
uint64_t splitmix64(uint64_t state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t power_mod(uint64_t base, uint64_t exponent, uint64_t modulo) {
    unsigned __int128 result = 1 % modulo, factor = base % modulo;
    while (exponent) {
        if (exponent & 1) result = result * factor % modulo;
        factor = factor * factor % modulo;
        exponent >>= 1;
    }
    return (uint64_t)result;
}

struct fp_device_tables_t {
    device_buffer_t wave_hashers, wave_dims, wave_runs, lane_hashers, lane_dims, lane_groups;
};

enum fp_kernel_t { fp_kernel_auto_k = 0, fp_kernel_wave_k = 1, fp_kernel_lane_k = 2 };

struct fp_engine_t {
    uint64_t dimensions = 0;
    // Wave kernel layout: runs of `64 * slots` lane-slots, lane `l` of slot `s` holding the run's dimension `64s + l`.
    int slots = 1;
    std::vector<fp_hasher_t> wave_hashers;
    std::vector<uint32_t> wave_dims;
    std::vector<fp_run_t> wave_runs;
    // Lane kernel layout: groups of up to 8 dimensions of one width.
    std::vector<fp_hasher_t> lane_hashers;
    std::vector<uint32_t> lane_dims;
    std::vector<fp_run_t> lane_groups;
    size_t smallest_width_group = 0;
    int kernel = fp_kernel_auto_k;  /**< tuning: which kernel runs */
    uint64_t segment_bytes = 0;     /**< tuning: window starts per unit of the lane kernel (0: sized to the job) */
    std::map<int, std::unique_ptr<fp_device_tables_t>> tables;
    double last_kernel_ms = 0;
    std::mutex mutex; /**< one call at a time per engine */
};

fp_hasher_t seeded_hasher(uint32_t dim, uint64_t width, uint64_t seed) {
    uint64_t const mixed = splitmix64(seed + dim);
    uint64_t const multiplier = 256 + mixed % 384;
    uint64_t const modulo = 4503599626977ull - splitmix64(mixed) % (1ull << 20);
    double const inverse = 1.0 / (double)modulo;
    fp_hasher_t hasher;
    hasher.multiplier = (double)multiplier;
    hasher.modulo = (double)modulo;
    hasher.inverse = std::nextafter(std::nextafter(inverse, 0.0), 0.0);
    hasher.discard = (double)((modulo - power_mod(multiplier, width, modulo)) % modulo);
    return hasher;
}

/**
 *  Builds the per-dimension hashers in the reference's layout (c/stringzillas/fingerprints.cuh:39-47, 117-140),
 *  groups them by width in dimension order, and lays each group out for both kernels.
 */
sz_status_t build_engine(fp_engine_t &engine, uint64_t dimensions, uint64_t const *widths, uint64_t widths_count,
                         uint64_t seed, char const **error_message) {
    static uint64_t const default_widths[] = {3, 4, 5, 7, 9, 11, 15, 31};
    if (!widths || widths_count == 0) widths = default_widths, widths_count = 8;
    for (uint64_t i = 0; i < widths_count; ++i)
        if (widths[i] == 0) return fail(sz_unexpected_dimensions_k, error_message, "Window widths must be positive");
    if (dimensions >= no_dimension_k)
        return fail(sz_unexpected_dimensions_k, error_message, "Too many fingerprint dimensions");
    uint64_t const per_width_min = dimensions / widths_count;
    uint64_t const per_width_max = (dimensions + widths_count - 1) / widths_count;
    bool const sliced = per_width_min == per_width_max && per_width_min % 64 == 0;

    std::map<uint64_t, std::vector<uint32_t>> by_width; // width -> dimensions, ascending
    std::vector<uint64_t> width_order;
    for (uint64_t d = 0; d < dimensions; ++d) {
        uint64_t const width = sliced ? widths[(d / 64) % widths_count] : widths[d % widths_count];
        auto &list = by_width[width];
        if (list.empty()) width_order.push_back(width);
        list.push_back((uint32_t)d);
    }
    size_t largest = 0, smallest = dimensions ? SIZE_MAX : 0;
    for (auto const &entry : by_width)
        largest = std::max(largest, entry.second.size()), smallest = std::min(smallest, entry.second.size());
    engine.dimensions = dimensions;
    engine.smallest_width_group = smallest;
    engine.slots = largest > 128 ? 4 : largest > 64 ? 2 : 1;
    fp_hasher_t const idle {0.0, 1.0, 1.0, 0.0}; // harmless arithmetic for an empty lane-slot
    uint32_t const run_capacity = (uint32_t)engine.slots * wave_k;
    for (uint64_t width : width_order) {
        std::vector<uint32_t> const &list = by_width[width];
        for (size_t at = 0; at < list.size(); at += run_capacity) {
            engine.wave_runs.push_back(fp_run_t {width, (uint32_t)engine.wave_hashers.size(), 0});
            for (uint32_t slot = 0; slot < run_capacity; ++slot) {
                bool const used = at + slot < list.size();
                uint32_t const dim = used ? list[at + slot] : no_dimension_k;
                engine.wave_hashers.push_back(used ? seeded_hasher(dim, width, seed) : idle);
                engine.wave_dims.push_back(dim);
            }
        }
        for (size_t at = 0; at < list.size(); at += lane_slots_k) {
            engine.lane_groups.push_back(fp_run_t {width, (uint32_t)engine.lane_hashers.size(), 0});
            for (int slot = 0; slot < lane_slots_k; ++slot) {
                bool const used = at + slot < list.size();
                uint32_t const dim = used ? list[at + slot] : no_dimension_k;
                engine.lane_hashers.push_back(used ? seeded_hasher(dim, width, seed) : idle);
                engine.lane_dims.push_back(dim);
            }
        }
    }
    return ok(error_message);
}

template <typename value_type_>
hipError_t upload(device_buffer_t &buffer, std::vector<value_type_> const &values, hipStream_t stream) {
    hipError_t error = buffer.reserve(std::max<size_t>(values.size(), 1) * sizeof(value_type_));
    if (error == hipSuccess && !values.empty())
        error = hipMemcpyAsync(buffer.data, values.data(), values.size() * sizeof(value_type_), hipMemcpyHostToDevice,
                               stream);
    return error;
}

sz_status_t device_tables(fp_engine_t &engine, int device, hipStream_t stream, fp_device_tables_t **out,
                          char const **error_message) {
    auto found = engine.tables.find(device);
    if (found != engine.tables.end()) {
        *out = found->second.get();
        return ok(error_message);
    }
    auto tables = std::make_unique<fp_device_tables_t>();
    hipError_t error = upload(tables->wave_hashers, engine.wave_hashers, stream);
    if (error == hipSuccess) error = upload(tables->wave_dims, engine.wave_dims, stream);
    if (error == hipSuccess) error = upload(tables->wave_runs, engine.wave_runs, stream);
    if (error == hipSuccess) error = upload(tables->lane_hashers, engine.lane_hashers, stream);
    if (error == hipSuccess) error = upload(tables->lane_dims, engine.lane_dims, stream);
    if (error == hipSuccess) error = upload(tables->lane_groups, engine.lane_groups, stream);
    if (error == hipSuccess) error = hipStreamSynchronize(stream);
    if (error != hipSuccess) return from_hip(error, error_message, "Uploading the fingerprint hashers");
    *out = tables.get();
    engine.tables[device] = std::move(tables);
    return ok(error_message);
}

hipError_t launch_kernel(fp_engine_t const &engine, bool lanes, fp_launch_t const &args, hipStream_t stream) {
    uint64_t const waves = lanes ? (args.units_count + wave_k - 1) / wave_k * args.runs_count
                                 : args.units_count * args.runs_count;
    if (!waves) return hipSuccess;
    uint64_t const blocks = std::min<uint64_t>((waves + waves_per_block_k - 1) / waves_per_block_k, max_blocks_k);
    dim3 const grid((uint32_t)blocks), block(wave_k * waves_per_block_k);
    if (lanes) hipLaunchKernelGGL(fingerprint_lane_kernel, grid, block, 0, stream, args);
    else if (engine.slots == 1) hipLaunchKernelGGL(fingerprint_wave_kernel<1>, grid, block, 0, stream, args);
    else if (engine.slots == 2) hipLaunchKernelGGL(fingerprint_wave_kernel<2>, grid, block, 0, stream, args);
    else hipLaunchKernelGGL(fingerprint_wave_kernel<4>, grid, block, 0, stream, args);
    return hipGetLastError();
}

/** `sz_sequence_t` gathered into a host tape. */
struct gathered_t {
    std::vector<char> data;
    std::vector<uint64_t> offsets;
    sz_sequence_u64tape_t view {};
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
        view.data = data.data(), view.offsets = offsets.data(), view.count = sequence->count;
    }
};

/**
 *  Cuts the texts into units of at most `segment` window starts: a text of up to `segment` bytes is one unit that
 *  writes its rows directly; a longer one becomes `ceil(length / segment)` units writing partials, merged after.
 *  For the lane kernel the units are ordered longest first (a counting sort on 32-byte buckets), so the lanes of
 *  a wave walk units of similar length.
 */
void plan_units(std::vector<uint64_t> const &lengths, uint64_t segment, bool sort_by_length,
                std::vector<fp_unit_t> &units, std::vector<fp_long_text_t> &long_texts, bool &identity) {
    uint64_t const count = lengths.size();
    uint64_t partials = 0, shortest = UINT64_MAX, longest = 0;
    for (uint64_t i = 0; i < count; ++i) {
        shortest = std::min(shortest, lengths[i]), longest = std::max(longest, lengths[i]);
        if (lengths[i] > segment) {
            uint64_t const pieces = (lengths[i] + segment - 1) / segment;
            long_texts.push_back(fp_long_text_t {i, partials, pieces});
            partials += pieces;
        }
    }
    bool const spread = sort_by_length && longest - shortest > 64;
    identity = long_texts.empty() && !spread;
    if (identity) return;
    units.reserve(count + partials);
    for (uint64_t i = 0, next_long = 0; i < count; ++i) {
        if (lengths[i] <= segment) {
            units.push_back(fp_unit_t {i, 0, direct_k});
            continue;
        }
        fp_long_text_t const &entry = long_texts[next_long++];
        for (uint64_t piece = 0; piece < entry.partials; ++piece)
            units.push_back(fp_unit_t {i, piece * segment, entry.first_partial + piece});
    }
    if (!spread) return;
    auto steps_of = [&](fp_unit_t const &unit) {
        return std::min<uint64_t>(lengths[unit.text] - unit.start, segment) / 32;
    };
    uint64_t const buckets = segment / 32 + 2;
    std::vector<uint64_t> starts(buckets + 1, 0);
    for (auto const &unit : units) ++starts[buckets - 1 - steps_of(unit)];
    for (uint64_t b = 0, sum = 0; b <= buckets; ++b) {
        uint64_t const here = starts[b];
        starts[b] = sum, sum += here;
    }
    std::vector<fp_unit_t> ordered(units.size());
    for (auto const &unit : units) ordered[starts[buckets - 1 - steps_of(unit)]++] = unit;
    units.swap(ordered);
}

template <typename tape_type_>
sz_status_t fingerprint_call_unguarded(void *engine_handle, szs_device_scope_t device, tape_type_ const *texts,
                                       sz_u32_t *min_hashes, sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                       sz_size_t min_counts_stride, char const **error_message) {
    if (!engine_handle) return fail(sz_unexpected_dimensions_k, error_message, "Engine must be initialized");
    if (!device) return fail(sz_unexpected_dimensions_k, error_message, "Device must be initialized");
    if (!texts) return fail(sz_unexpected_dimensions_k, error_message, "Input texts cannot be null");
    fp_engine_t &engine = *reinterpret_cast<fp_engine_t *>(engine_handle);
    uint64_t const count = texts->count, dims = engine.dimensions;
    if (count == 0 || dims == 0) return ok(error_message);
    if (!min_hashes || !min_counts)
        return fail(sz_unexpected_dimensions_k, error_message, "Output min_hashes and min_counts cannot be null");
    if (min_hashes_stride < dims * 4 || min_counts_stride < dims * 4)
        return fail(sz_unexpected_dimensions_k, error_message, "Output strides are shorter than a fingerprint");
    std::lock_guard<std::mutex> lock(engine.mutex);

    int gpu = 0;
    hipStream_t stream = nullptr;
    sz_status_t status = resolve_gpu(reinterpret_cast<device_scope_t *>(device), &gpu, &stream, error_message);
    if (status != sz_success_k) return status;
    hipError_t const ordered = order_after_default_stream(stream);
    if (ordered != hipSuccess) return from_hip(ordered, error_message, "Ordering after the default stream");
    fp_device_tables_t *tables = nullptr;
    status = device_tables(engine, gpu, stream, &tables, error_message);
    if (status != sz_success_k) return status;

    staged_tape_t tape;
    status = stage_bytes_tape(texts, tape, stream, error_message);
    if (status != sz_success_k) return status;

#define SZS_FP_TRY(expr, context)                                                                                     \
    do {                                                                                                             \
        hipError_t const error_ = (expr);                                                                            \
        if (error_ != hipSuccess) return from_hip(error_, error_message, context);                                   \
    } while (0)
    // Text lengths on the host: from the caller's offsets, or read back when they live on the device.
    std::vector<uint64_t> lengths(count);
    std::vector<uint64_t> offsets;
    if (is_device_pointer(texts->offsets)) {
        offsets.resize(count + 1);
        SZS_FP_TRY(hipMemcpyAsync(offsets.data(), tape.offsets, (count + 1) * 8, hipMemcpyDeviceToHost, stream),
                   "Reading back text offsets");
        SZS_FP_TRY(hipStreamSynchronize(stream), "Reading back text offsets");
    }
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t const begin = offsets.empty() ? (uint64_t)texts->offsets[i] : offsets[i];
        uint64_t const end = offsets.empty() ? (uint64_t)texts->offsets[i + 1] : offsets[i + 1];
        if (end < begin) return fail(sz_unexpected_dimensions_k, error_message, "Text offsets must not decrease");
        lengths[i] = end - begin;
    }

    // The wave kernel fills its lanes when every width holds at least 64 dimensions; otherwise lanes walk texts.
    bool const lanes = engine.kernel == fp_kernel_lane_k ||
                       (engine.kernel == fp_kernel_auto_k && engine.smallest_width_group < wave_k);
    uint64_t segment = wave_segment_bytes_k;
    if (lanes) {
        uint64_t total = 0;
        for (uint64_t length : lengths) total += length;
        // About 32K waves of work, several per SIMD at the lane kernel's occupancy, so the tail stays short: on
        // 200 MB of 10 KB documents at 64 dimensions this picks 1 KB units, the fastest of 1-64 KB measured.
        uint64_t const target_units = std::max<uint64_t>(1, 32768ull * wave_k / engine.lane_groups.size());
        segment = 256;
        while (segment < 16384 && segment * target_units < total) segment *= 2;
    }
    if (engine.segment_bytes) segment = engine.segment_bytes;
    std::vector<fp_unit_t> units;
    std::vector<fp_long_text_t> long_texts;
    bool identity = true;
    plan_units(lengths, segment, lanes, units, long_texts, identity);
    uint64_t const partials = long_texts.empty() ? 0 : long_texts.back().first_partial + long_texts.back().partials;

    // Outputs: in place when both are device memory, otherwise through contiguous device rows.
    bool const device_outputs = is_device_pointer(min_hashes) && is_device_pointer(min_counts);
    device_buffer_t staged_hashes, staged_counts, partial_minimums, partial_counts, units_buffer, long_buffer;
    uint32_t *hashes = min_hashes, *counts = min_counts;
    uint64_t hashes_stride = min_hashes_stride, counts_stride = min_counts_stride;
    if (!device_outputs) {
        SZS_FP_TRY(staged_hashes.reserve(count * dims * 4), "Failed to allocate fingerprint outputs");
        SZS_FP_TRY(staged_counts.reserve(count * dims * 4), "Failed to allocate fingerprint outputs");
        hashes = staged_hashes.as<uint32_t>(), counts = staged_counts.as<uint32_t>();
        hashes_stride = counts_stride = dims * 4;
    }
    if (!identity) SZS_FP_TRY(upload(units_buffer, units, stream), "Uploading text units");
    if (partials) {
        uint64_t const positions = lanes ? engine.lane_dims.size() : engine.wave_dims.size();
        SZS_FP_TRY(partial_minimums.reserve(partials * positions * 8), "Failed to allocate segment partials");
        SZS_FP_TRY(partial_counts.reserve(partials * positions * 4), "Failed to allocate segment partials");
        SZS_FP_TRY(upload(long_buffer, long_texts, stream), "Uploading long texts");
    }

    hipEvent_t begin = nullptr, finish = nullptr;
    SZS_FP_TRY(hipEventCreate(&begin), "hipEventCreate failed");
    hipError_t error = hipEventCreate(&finish);
    if (error == hipSuccess) error = hipEventRecord(begin, stream);

    fp_launch_t args {};
    args.data = tape.data, args.offsets = tape.offsets;
    if (lanes) {
        args.hashers = tables->lane_hashers.as<fp_hasher_t>(), args.dims = tables->lane_dims.as<uint32_t>();
        args.runs = tables->lane_groups.as<fp_run_t>(), args.runs_count = (uint32_t)engine.lane_groups.size();
    }
    else {
        args.hashers = tables->wave_hashers.as<fp_hasher_t>(), args.dims = tables->wave_dims.as<uint32_t>();
        args.runs = tables->wave_runs.as<fp_run_t>(), args.runs_count = (uint32_t)engine.wave_runs.size();
    }
    args.dimensions = (uint32_t)dims;
    args.units = identity ? nullptr : units_buffer.as<fp_unit_t>();
    args.units_count = identity ? count : units.size();
    args.segment = segment;
    args.hashes = hashes, args.counts = counts, args.hashes_stride = hashes_stride, args.counts_stride = counts_stride;
    args.partial_minimums = partial_minimums.as<double>(), args.partial_counts = partial_counts.as<uint32_t>();
    args.positions = (uint32_t)(lanes ? engine.lane_dims.size() : engine.wave_dims.size());
    if (error == hipSuccess) error = launch_kernel(engine, lanes, args, stream);
    if (error == hipSuccess && partials) {
        fp_merge_t merge {long_buffer.as<fp_long_text_t>(), long_texts.size(), args.positions, args.dims,
                          partial_minimums.as<double>(), partial_counts.as<uint32_t>(), hashes, counts,
                          hashes_stride, counts_stride};
        uint64_t const threads = long_texts.size() * args.positions;
        hipLaunchKernelGGL(fingerprint_merge_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, stream,
                           merge);
        error = hipGetLastError();
    }
    if (error == hipSuccess) error = hipEventRecord(finish, stream);
    if (error == hipSuccess && !device_outputs) {
        error = hipMemcpy2DAsync(min_hashes, min_hashes_stride, hashes, dims * 4, dims * 4, count,
                                 hipMemcpyDeviceToHost, stream);
        if (error == hipSuccess)
            error = hipMemcpy2DAsync(min_counts, min_counts_stride, counts, dims * 4, dims * 4, count,
                                     hipMemcpyDeviceToHost, stream);
    }
    if (error == hipSuccess) error = hipStreamSynchronize(stream);
    float elapsed = 0;
    if (error == hipSuccess && hipEventElapsedTime(&elapsed, begin, finish) == hipSuccess)
        engine.last_kernel_ms = elapsed;
    if (begin) (void)hipEventDestroy(begin);
    if (finish) (void)hipEventDestroy(finish);
#undef SZS_FP_TRY
    if (error != hipSuccess) return from_hip(error, error_message, "Fingerprinting failed");
    return ok(error_message);
}

/** The C ABI boundary: no C++ exception (a failed host allocation, for one) crosses it. */
template <typename tape_type_>
sz_status_t fingerprint_call(void *engine_handle, szs_device_scope_t device, tape_type_ const *texts,
                             sz_u32_t *min_hashes, sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                             sz_size_t min_counts_stride, char const **error_message) {
    return guarded(error_message, [&]() {
        return fingerprint_call_unguarded(engine_handle, device, texts, min_hashes, min_hashes_stride, min_counts,
                                          min_counts_stride, error_message);
    });
}

#pragma endregion

} // namespace

extern "C" {

sz_status_t szs_fingerprints_init(sz_size_t dimensions, sz_size_t alphabet_size, sz_size_t const *window_widths,
                                  sz_size_t window_widths_count, sz_u64_t seed, sz_memory_allocator_t const *alloc,
                                  sz_capability_t capabilities, szs_fingerprints_t *engine,
                                  char const **error_message) {
    (void)alphabet_size, (void)alloc, (void)capabilities; // the f64 hasher ignores the alphabet (serial.hpp:426-432)
    if (!engine) return fail(sz_unexpected_dimensions_k, error_message, "Engine pointer cannot be null");
    fp_engine_t *created = new (std::nothrow) fp_engine_t();
    if (!created) return fail(sz_bad_alloc_k, error_message, "Failed to allocate Fingerprints engine");
    static_assert(sizeof(sz_size_t) == sizeof(uint64_t), "64-bit sizes expected");
    sz_status_t const status = guarded(error_message, [&]() {
        return build_engine(*created, dimensions, reinterpret_cast<uint64_t const *>(window_widths),
                            window_widths_count, seed, error_message);
    });
    if (status != sz_success_k) {
        delete created;
        return status;
    }
    *engine = created;
    return ok(error_message);
}

sz_status_t szs_fingerprints_sequence(szs_fingerprints_t engine, szs_device_scope_t device, sz_sequence_t const *texts,
                                      sz_u32_t *min_hashes, sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                      sz_size_t min_counts_stride, char const **error_message) {
    if (!texts) return fail(sz_unexpected_dimensions_k, error_message, "Input texts cannot be null");
    return guarded(error_message, [&]() {
        gathered_t gathered;
        gathered.gather(texts);
        return fingerprint_call(engine, device, &gathered.view, min_hashes, min_hashes_stride, min_counts,
                                min_counts_stride, error_message);
    });
}

sz_status_t szs_fingerprints_u64tape(szs_fingerprints_t engine, szs_device_scope_t device,
                                     sz_sequence_u64tape_t const *texts, sz_u32_t *min_hashes,
                                     sz_size_t min_hashes_stride, sz_u32_t *min_counts, sz_size_t min_counts_stride,
                                     char const **error_message) {
    return fingerprint_call(engine, device, texts, min_hashes, min_hashes_stride, min_counts, min_counts_stride,
                            error_message);
}

sz_status_t szs_fingerprints_u32tape(szs_fingerprints_t engine, szs_device_scope_t device,
                                     sz_sequence_u32tape_t const *texts, sz_u32_t *min_hashes,
                                     sz_size_t min_hashes_stride, sz_u32_t *min_counts, sz_size_t min_counts_stride,
                                     char const **error_message) {
    return fingerprint_call(engine, device, texts, min_hashes, min_hashes_stride, min_counts, min_counts_stride,
                            error_message);
}

void szs_fingerprints_free(szs_fingerprints_t engine) { delete reinterpret_cast<fp_engine_t *>(engine); }

sz_status_t szs_fingerprints_utf8_init(sz_size_t dimensions, sz_size_t alphabet_size, sz_size_t const *window_widths,
                                       sz_size_t window_widths_count, sz_u64_t seed,
                                       sz_memory_allocator_t const *alloc, sz_capability_t capabilities,
                                       szs_fingerprints_utf8_t *engine, char const **error_message) {
    return szs_fingerprints_init(dimensions, alphabet_size, window_widths, window_widths_count, seed, alloc,
                                 capabilities, engine, error_message);
}

sz_status_t szs_fingerprints_utf8_sequence(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                           sz_sequence_t const *texts, sz_u32_t *min_hashes,
                                           sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                           sz_size_t min_counts_stride, char const **error_message) {
    return szs_fingerprints_sequence(engine, device, texts, min_hashes, min_hashes_stride, min_counts,
                                     min_counts_stride, error_message);
}

sz_status_t szs_fingerprints_utf8_u32tape(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                          sz_sequence_u32tape_t const *texts, sz_u32_t *min_hashes,
                                          sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                          sz_size_t min_counts_stride, char const **error_message) {
    return szs_fingerprints_u32tape(engine, device, texts, min_hashes, min_hashes_stride, min_counts,
                                    min_counts_stride, error_message);
}

sz_status_t szs_fingerprints_utf8_u64tape(szs_fingerprints_utf8_t engine, szs_device_scope_t device,
                                          sz_sequence_u64tape_t const *texts, sz_u32_t *min_hashes,
                                          sz_size_t min_hashes_stride, sz_u32_t *min_counts,
                                          sz_size_t min_counts_stride, char const **error_message) {
    return szs_fingerprints_u64tape(engine, device, texts, min_hashes, min_hashes_stride, min_counts,
                                    min_counts_stride, error_message);
}

void szs_fingerprints_utf8_free(szs_fingerprints_utf8_t engine) { szs_fingerprints_free(engine); }

sz_size_t szs_fingerprints_dimensions(szs_fingerprints_t engine) {
    return engine ? (sz_size_t) reinterpret_cast<fp_engine_t *>(engine)->dimensions : 0;
}

double szs_fingerprints_last_kernel_ms(szs_fingerprints_t engine) {
    return engine ? reinterpret_cast<fp_engine_t *>(engine)->last_kernel_ms : 0.0;
}

sz_status_t szs_fingerprints_tune(szs_fingerprints_t engine, int kernel, sz_size_t segment_bytes,
                                  char const **error_message) {
    if (!engine) return fail(sz_unexpected_dimensions_k, error_message, "Engine must be initialized");
    if (kernel < fp_kernel_auto_k || kernel > fp_kernel_lane_k)
        return fail(sz_unexpected_dimensions_k, error_message, "Kernel must be 0 (automatic), 1 (wave) or 2 (lane)");
    if (segment_bytes && (segment_bytes < 64 || segment_bytes > (1ull << 30)))
        return fail(sz_unexpected_dimensions_k, error_message, "Segment length must be 0 or within [64, 2^30]");
    fp_engine_t &state = *reinterpret_cast<fp_engine_t *>(engine);
    std::lock_guard<std::mutex> lock(state.mutex);
    state.kernel = kernel;
    state.segment_bytes = segment_bytes;
    return ok(error_message);
}

} // extern "C"
